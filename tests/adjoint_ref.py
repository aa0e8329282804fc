"""FP64 CPU reference of the topology-fixed boolean adjoint (test
infrastructure; the product runs it on the GPU through sbg_adjoint).

A derived vertex is X = x + t (y - x), t = ((p0 - x).N) / ((y - x).N),
N = (p1 - p0) x (p2 - p0): the line through input vertices x, y meets the plane
through p0, p1, p2.  The volume of a face set is sum v0.(v1 x v2) / 6.  The
vector-Jacobian products below are checked against central differences in
test_adjoint_ref.py, which is what pins them."""
import numpy as np


def derived(xyz, src):
    x, y, p0, p1, p2 = (xyz[src[:, k]] for k in range(5))
    e, w = y - x, p0 - x
    N = np.cross(p1 - p0, p2 - p0)
    t = np.einsum("ij,ij->i", w, N) / np.einsum("ij,ij->i", e, N)
    return x + t[:, None] * e


def all_vertices(xyz, src):
    return np.concatenate([xyz, derived(xyz, src)]) if len(src) else xyz


def face_terms(allv, faces):
    a, b, c = allv[faces[:, 0]], allv[faces[:, 1]], allv[faces[:, 2]]
    return np.einsum("ij,ij->i", a, np.cross(b, c)) / 6.0


def volume(xyz, src, faces):
    return float(np.sum(face_terms(all_vertices(xyz, src), faces)))


def volume_difference(xyz_plus, xyz_minus, src, faces):
    """V(plus) - V(minus) summed face by face, so the large total cancels
    first; evaluated in the inputs' precision (np.longdouble for checks)."""
    return np.sum(face_terms(all_vertices(xyz_plus, src), faces) - face_terms(all_vertices(xyz_minus, src), faces))


def vjp(xyz, src, faces, vol_weight=1.0, grad_derived=None):
    """vol_weight * dV/dxyz + J^T grad_derived, J = d(derived)/d(xyz)."""
    nin = len(xyz)
    allv = all_vertices(xyz, src)
    g = np.zeros_like(allv)
    if vol_weight:
        a, b, c = allv[faces[:, 0]], allv[faces[:, 1]], allv[faces[:, 2]]
        w6 = vol_weight / 6.0
        np.add.at(g, faces[:, 0], np.cross(b, c) * w6)
        np.add.at(g, faces[:, 1], np.cross(c, a) * w6)
        np.add.at(g, faces[:, 2], np.cross(a, b) * w6)
    if len(src):
        gd = g[nin:].copy()
        if grad_derived is not None:
            gd += grad_derived
        x, y, p0, p1, p2 = (xyz[src[:, k]] for k in range(5))
        e, w = y - x, p0 - x
        c, d = p2 - p0, p1 - p0
        N = np.cross(d, c)
        bb = np.einsum("ij,ij->i", e, N)
        t = np.einsum("ij,ij->i", w, N) / bb
        s = np.einsum("ij,ij->i", gd, e) / bb
        m = w - t[:, None] * e
        gsN = gd - s[:, None] * N
        gp1 = s[:, None] * np.cross(c, m)
        gp2 = s[:, None] * np.cross(m, d)
        gp0 = s[:, None] * N - gp1 - gp2
        for k, v in enumerate(((1.0 - t)[:, None] * gsN, t[:, None] * gsN, gp0, gp1, gp2)):
            np.add.at(g, src[:, k], v)
    return g[:nin]


def directions(n, count, seed):
    """`count` seeded unit directions in R^(n x 3)."""
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((count, n, 3))
    return u / np.linalg.norm(u.reshape(count, -1), axis=1)[:, None, None]
