// Topology-fixed adjoint of a boolean result (no reference counterpart: the
// reference computes no derivatives).
//
// Every result vertex is either an input vertex or a derived vertex: the point
// where the line through two input vertices x, y (an edge of one triangle)
// meets the plane through three others p0, p1, p2 (the other triangle), which
// is how every Guigue-Devillers segment endpoint is made
// (thirdparty/GuigueDevillers03/tri_tri_intersect.c:297-356).  With
//   e = y - x, w = p0 - x, N = (p1 - p0) x (p2 - p0), t = (w.N) / (e.N),
//   X = x + t e,
// the vector-Jacobian product of an upstream gradient g is, with
// s = (g.e) / (e.N), m = w - t e, c = p2 - p0, d = p1 - p0:
//   g_x = (1 - t)(g - s N),  g_y = t (g - s N),
//   g_p1 = s (c x m),  g_p2 = s (m x d),  g_p0 = s N - g_p1 - g_p2.
// The volume of a closed face set is sum over faces of v0.(v1 x v2) / 6, with
// d/dv0 = (v1 x v2) / 6 and cyclic.  Sums over faces and derived vertices use
// FP64 atomics, so the last bits of a gradient may differ between runs.
#include "sbg_device.h"
#include "sbg_internal.h"

#pragma clang fp contract(off)

namespace sbg {

__device__ __forceinline__ V3 load3(const double *__restrict__ p, uint32_t i)
{
    return {p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2]};
}

__device__ __forceinline__ V3 vadd(const V3 &a, const V3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 vscale(const V3 &a, double s) { return {a.x * s, a.y * s, a.z * s}; }

__device__ __forceinline__ void atomic_add3(double *__restrict__ g, uint32_t i, const V3 &v)
{
    atomicAdd(g + 3 * (size_t)i, v.x);
    atomicAdd(g + 3 * (size_t)i + 1, v.y);
    atomicAdd(g + 3 * (size_t)i + 2, v.z);
}

// positions of all vertices [inputs | derived] into all (nin + nd) x 3
__global__ __launch_bounds__(256) void k_derived(const double *__restrict__ xyz, uint32_t nin,
                                                 const uint32_t *__restrict__ src, uint32_t nd,
                                                 double *__restrict__ all)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < nin) {
        for (int k = 0; k < 3; ++k)
            all[3 * (size_t)i + k] = xyz[3 * (size_t)i + k];
        return;
    }
    const uint32_t d = i - nin;
    if (d >= nd)
        return;
    const uint32_t *s = src + 5 * (size_t)d;
    const V3 x = load3(xyz, s[0]), y = load3(xyz, s[1]);
    const V3 p0 = load3(xyz, s[2]), p1 = load3(xyz, s[3]), p2 = load3(xyz, s[4]);
    const V3 e = vsub(y, x), w = vsub(p0, x);
    const V3 N = vcross(vsub(p1, p0), vsub(p2, p0));
    const double t = vdot(w, N) / vdot(e, N);
    const V3 X = vadd(x, vscale(e, t));
    all[3 * (size_t)i] = X.x;
    all[3 * (size_t)i + 1] = X.y;
    all[3 * (size_t)i + 2] = X.z;
}

// volume (block partial sums into *vol) and face gradients into gall
__global__ __launch_bounds__(256) void k_volume(const double *__restrict__ all, const uint32_t *__restrict__ faces,
                                                uint32_t nf, double weight, double *vol, double *__restrict__ gall)
{
    __shared__ double s_sum[256];
    const uint32_t f = blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (f < nf) {
        const uint32_t a = faces[3 * (size_t)f], b = faces[3 * (size_t)f + 1], c = faces[3 * (size_t)f + 2];
        const V3 va = load3(all, a), vb = load3(all, b), vc = load3(all, c);
        v = vdot(va, vcross(vb, vc)) / 6.0;
        if (gall && weight != 0.0) {
            const double w6 = weight / 6.0;
            atomic_add3(gall, a, vscale(vcross(vb, vc), w6));
            atomic_add3(gall, b, vscale(vcross(vc, va), w6));
            atomic_add3(gall, c, vscale(vcross(va, vb), w6));
        }
    }
    s_sum[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            s_sum[threadIdx.x] += s_sum[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(vol, s_sum[0]);
}

// derived vertices pass their gradient (from the faces plus grad_derived) to
// their five inputs
__global__ __launch_bounds__(256) void k_derived_vjp(const double *__restrict__ xyz, uint32_t nin,
                                                     const uint32_t *__restrict__ src, uint32_t nd,
                                                     const double *__restrict__ gradDerived, double *__restrict__ gall)
{
    const uint32_t d = blockIdx.x * 256 + threadIdx.x;
    if (d >= nd)
        return;
    V3 g = load3(gall, nin + d);
    if (gradDerived)
        g = vadd(g, load3(gradDerived, d));
    const uint32_t *sv = src + 5 * (size_t)d;
    const V3 x = load3(xyz, sv[0]), y = load3(xyz, sv[1]);
    const V3 p0 = load3(xyz, sv[2]), p1 = load3(xyz, sv[3]), p2 = load3(xyz, sv[4]);
    const V3 e = vsub(y, x), w = vsub(p0, x);
    const V3 c = vsub(p2, p0), dd = vsub(p1, p0);
    const V3 N = vcross(dd, c);
    const double b = vdot(e, N);
    const double t = vdot(w, N) / b;
    const double s = vdot(g, e) / b;
    const V3 m = vsub(w, vscale(e, t));
    const V3 gsN = vsub(g, vscale(N, s));
    const V3 gp1 = vscale(vcross(c, m), s);
    const V3 gp2 = vscale(vcross(m, dd), s);
    const V3 gp0 = vsub(vsub(vscale(N, s), gp1), gp2);
    atomic_add3(gall, sv[0], vscale(gsN, 1.0 - t));
    atomic_add3(gall, sv[1], vscale(gsN, t));
    atomic_add3(gall, sv[2], gp0);
    atomic_add3(gall, sv[3], gp1);
    atomic_add3(gall, sv[4], gp2);
}

// all: (nin + nd) x 3 positions out; *vol and gall ((nin + nd) x 3, may be
// null for no gradient) zero on entry
void launch_adjoint(const double *xyz, uint32_t nin, const uint32_t *src, uint32_t nd, const uint32_t *faces,
                    uint32_t nf, double weight, const double *gradDerived, double *all, double *vol, double *gall,
                    hipStream_t s)
{
    const uint64_t n = (uint64_t)nin + nd;
    if (n)
        hipLaunchKernelGGL(k_derived, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, xyz, nin, src, nd, all);
    if (nf)
        hipLaunchKernelGGL(k_volume, dim3((nf + 255) / 256), dim3(256), 0, s, all, faces, nf, weight, vol, gall);
    if (gall && nd)
        hipLaunchKernelGGL(k_derived_vjp, dim3((nd + 255) / 256), dim3(256), 0, s, xyz, nin, src, nd, gradDerived,
                           gall);
}

}  // namespace sbg
