"""Differentiable boolean: gradients of the union volume and of the new
(intersection) vertex positions with respect to every input vertex
coordinate, the combinatorial topology held fixed.

The topology comes from one drop-in combine (SolidBoolean over the GPU
path): the kept faces and, for every new vertex, the two input vertices of
the edge and the three of the plane whose intersection made it.  Values and
vector-Jacobian products then come from the hand-written adjoint kernel
behind the C-ABI (sbg_adjoint), wrapped as a torch.autograd.Function.
Inputs are the first mesh's vertices followed by the second's, (nA + nB) x 3
FP64.
"""
from __future__ import annotations

import numpy as np
import torch

from . import dropin

NONE = 0xFFFFFFFF
OPS = ("union", "diff", "intersect")


class Topology:
    """Faces of one boolean result over [inputs | derived] and the edge/plane
    input indices (derived x 5) of every derived vertex."""

    def __init__(self, va, ta, vb, tb, op: str = "union"):
        if op not in OPS:
            raise ValueError(f"op must be one of {OPS}")
        va = np.ascontiguousarray(va, np.float64).reshape(-1, 3)
        vb = np.ascontiguousarray(vb, np.float64).reshape(-1, 3)
        A, B = dropin.SolidMesh(va, ta), dropin.SolidMesh(vb, tb)
        A.prepare()
        B.prepare()
        sb = dropin.SolidBoolean(A, B)
        if not sb.combine():
            raise RuntimeError("combine() failed on these meshes; no topology to differentiate")
        src = sb.vertex_sources().astype(np.int64)
        faces = {"union": sb.fetch_union, "diff": sb.fetch_diff, "intersect": sb.fetch_intersect}[op]()
        self.nin = len(va) + len(vb)
        derived = src[:, 1] != NONE
        remap = np.empty(len(src), np.int64)
        remap[~derived] = src[~derived, 0]
        remap[derived] = self.nin + np.arange(int(derived.sum()))
        self.faces = np.ascontiguousarray(remap[faces], np.uint32)
        self.src = np.ascontiguousarray(src[derived], np.uint32)
        self.result_vertices = sb.result_vertices()
        self.derived_rows = np.nonzero(derived)[0]  # result-vertex index of each derived vertex
        self.input_xyz = np.concatenate([va, vb])
        self.op = op
        self._keep = (A, B, sb)

    @property
    def nderived(self) -> int:
        return len(self.src)


class _BooleanFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, topo, gpu):
        x = xyz.detach().cpu().numpy().astype(np.float64)
        derived, vol, _ = gpu.adjoint(x, topo.src, topo.faces, want_grad=False)
        ctx.topo, ctx.gpu = topo, gpu
        ctx.save_for_backward(xyz)
        return (torch.tensor(vol, dtype=torch.float64, device=xyz.device),
                torch.from_numpy(derived).to(xyz.device))

    @staticmethod
    def backward(ctx, g_vol, g_derived):
        (xyz,) = ctx.saved_tensors
        w = 0.0 if g_vol is None else float(g_vol)
        gd = None if g_derived is None else g_derived.detach().cpu().numpy().astype(np.float64)
        _, _, grad = ctx.gpu.adjoint(xyz.detach().cpu().numpy().astype(np.float64), ctx.topo.src, ctx.topo.faces,
                                     vol_weight=w, grad_derived=gd)
        return torch.from_numpy(grad).to(xyz.device), None, None


def volume_and_vertices(xyz: torch.Tensor, topo: Topology, gpu):
    """(volume of the result faces, derived vertex positions) as differentiable
    functions of the input vertices xyz ((nA + nB) x 3, float64)."""
    return _BooleanFn.apply(xyz, topo, gpu)
