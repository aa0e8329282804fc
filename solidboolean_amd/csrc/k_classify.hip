// Inside/outside classification by ray parity (reference: isPointInMesh,
// src/solidboolean.cpp:48-92, its three ray directions :31-35, and the vote of
// decideGroupSide :482-510).
//
// One lane per (query point, axis), axis-major, and for triangle queries in
// the querying mesh's Morton order, so neighbouring lanes trace parallel rays
// from neighbouring points.  The ray is the segment p -> p + axis with
// axis = (DBL_MAX, eps, eps) and permutations; its FP64 box is rounded outward
// to FP32 to prune the target's LBVH; every leaf whose exact FP64 box
// closed-overlaps the segment's box (the reference's candidate set, :55-63)
// gets the FP64 plane test (src/vector3.h:264-280) and the three-normal inside
// test (:78-86) in the reference's operation order.  Hits are deduplicated by
// PositionKey (src/positionkey.cpp:25-37); the lane's answer is the parity of
// the distinct keys (:89).
//
// k_rays_fast keeps its traversal stack in LDS and up to 4 distinct keys in
// registers; a lane that needs more appends its (point, axis) to an overflow
// list, which k_rays_list finishes with a 64-deep private stack and 32 keys.
// A lane past those limits raises `overflow` and the call fails: nothing is
// guessed.
#include "sbg_device.h"
#include "sbg_internal.h"

#pragma clang fp contract(off)

namespace sbg {

constexpr int kFastDepth = 40;
constexpr int kFastKeys = 4;
constexpr int kListStack = 64;
constexpr int kListKeys = 32;

struct RBox {
    float lx, ly, lz, hx, hy, hz;
};

__device__ __forceinline__ bool rbox_overlap(float alx, float aly, float alz, float ahx, float ahy, float ahz,
                                             const RBox &b)
{
    return alx <= b.hx && ahx >= b.lx && aly <= b.hy && ahy >= b.ly && alz <= b.hz && ahz >= b.lz;
}

__global__ __launch_bounds__(256) void k_mark_intersected(const uint64_t *__restrict__ pairs,
                                                          const uint8_t *__restrict__ status, uint32_t n, int side,
                                                          uint8_t *__restrict__ flags)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || status[i] != 1)
        return;
    const uint64_t pr = pairs[i];
    flags[side == 0 ? (uint32_t)(pr >> 32) : (uint32_t)pr] = 1;
}

// (v0 + v1 + v2) / 3.0 (src/solidboolean.cpp:497-499)
__global__ __launch_bounds__(256) void k_centroids(const double *__restrict__ tri64, const uint8_t *__restrict__ flags,
                                                   uint32_t nt, double *__restrict__ pts)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nt || (flags && flags[t]))
        return;
    const double *r = tri64 + (size_t)kTri64Stride * t;
    pts[3 * (size_t)t] = (r[0] + r[4] + r[8]) / 3.0;
    pts[3 * (size_t)t + 1] = (r[1] + r[5] + r[9]) / 3.0;
    pts[3 * (size_t)t + 2] = (r[2] + r[6] + r[10]) / 3.0;
}

struct Ray {
    V3 p, u;
    Box64 rb;
    RBox rf;
};

__device__ __forceinline__ Ray ray_setup(const double *__restrict__ pts, uint32_t i, uint32_t axis)
{
    const double eps = 2.220446049250313080847e-16, big = 1.7976931348623157e308;
    Ray r;
    r.p = {pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]};
    const V3 dir = {axis == 0 ? big : eps, axis == 1 ? big : eps, axis == 2 ? big : eps};
    const V3 e = {r.p.x + dir.x, r.p.y + dir.y, r.p.z + dir.z};
    r.rb = {{big, big, big}, {-big, -big, -big}};
    box_update(r.rb, r.p);
    box_update(r.rb, e);
    r.rf = {f_down(r.rb.mn.x), f_down(r.rb.mn.y), f_down(r.rb.mn.z),
            f_up(r.rb.mx.x),   f_up(r.rb.mx.y),   f_up(r.rb.mx.z)};
    r.u = vsub(e, r.p);
    return r;
}

// distinct PositionKeys of one ray; CAP is small enough to stay in registers
// when every index below is a compile-time constant
template <int CAP>
struct KeySet {
    Key3 k[CAP];
    int n = 0;
    bool full = false;
    __device__ __forceinline__ void add(const Key3 &key)
    {
#pragma unroll
        for (int j = 0; j < CAP; ++j)
            if (j < n && k[j].x == key.x && k[j].y == key.y && k[j].z == key.z)
                return;
        if (n == CAP) {
            full = true;
            return;
        }
#pragma unroll
        for (int j = 0; j < CAP; ++j)
            if (j == n)
                k[j] = key;
        ++n;
    }
};

// candidate triangle t: FP64 box test, plane hit, three-normal inside test
template <class Keys>
__device__ __forceinline__ void ray_leaf(uint32_t t, const double *__restrict__ tri64,
                                         const double *__restrict__ normals, const Ray &r, Keys &keys)
{
    const double2 *x = reinterpret_cast<const double2 *>(tri64 + (size_t)kTri64Stride * t);
    const double2 r0 = x[0], r1 = x[1], r2 = x[2], r3 = x[3], r4 = x[4], r5 = x[5];
    const V3 v0 = {r0.x, r0.y, r1.x}, v1 = {r2.x, r2.y, r3.x}, v2 = {r4.x, r4.y, r5.x};
    if (!box64_overlap(tri_box64(v0, v1, v2), r.rb))
        return;
    const V3 n = {normals[3 * (size_t)t], normals[3 * (size_t)t + 1], normals[3 * (size_t)t + 2]};
    const V3 w = vsub(r.p, v0);
    const double d = n.x * r.u.x + n.y * r.u.y + n.z * r.u.z;
    const double num = (-n.x) * w.x + (-n.y) * w.y + (-n.z) * w.z;
    if (__builtin_fabs(d) <= 2.220446049250313080847e-16)
        return;
    const double s = num / d;
    if (s < 0 || s > 1 || __builtin_isnan(s) || __builtin_isinf(s))
        return;
    const V3 hit = {r.p.x + s * r.u.x, r.p.y + s * r.u.y, r.p.z + s * r.u.z};
    const V3 n0 = tri_normal(hit, v0, v1);
    const V3 n1 = tri_normal(hit, v1, v2);
    const V3 n2 = tri_normal(hit, v2, v0);
    if (!(vdot(n0, n1) > 0 && vdot(n0, n2) > 0))
        return;
    keys.add(position_key(hit));
}

__global__ __launch_bounds__(256) void k_rays_fast(const BNode *__restrict__ nodes, uint32_t root,
                                                   const float *__restrict__ rootBox, const double *__restrict__ tri64,
                                                   const double *__restrict__ normals, const double *__restrict__ pts,
                                                   const uint32_t *__restrict__ order, const uint8_t *__restrict__ skip,
                                                   uint32_t n, uint8_t *__restrict__ bits, uint32_t *__restrict__ ovfList,
                                                   uint32_t *ovfCount, uint32_t maxVisits)
{
    __shared__ uint32_t s_stack[kFastDepth][256];
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= 3u * n)
        return;
    const uint32_t axis = g / n;
    const uint32_t q = g - axis * n;
    const uint32_t i = order ? order[q] : q;
    if (skip && skip[i])
        return;
    const Ray r = ray_setup(pts, i, axis);
    KeySet<kFastKeys> keys;
    bool ovf = false;
    if (rbox_overlap(rootBox[0], rootBox[1], rootBox[2], rootBox[3], rootBox[4], rootBox[5], r.rf)) {
        if (root & kLeafBit) {
            ray_leaf(root & ~kLeafBit, tri64, normals, r, keys);
        } else {
            uint32_t *stack = &s_stack[0][threadIdx.x];  // entry k at stack[k * 256]
            int sp = 0;
            uint32_t cur = root;
            uint32_t visits = 0;
            // visit cur; descend into an overlapping child, push the other
            for (;;) {
                if (++visits > maxVisits || keys.full) {
                    ovf = true;
                    break;
                }
                const BNode *nd = nodes + cur;
                const float4 a = nd->a, b = nd->b, c = nd->c;
                const int4 d = nd->d;
                const bool o0 = rbox_overlap(a.x, a.y, a.z, a.w, b.x, b.y, r.rf);
                const bool o1 = rbox_overlap(b.z, b.w, c.x, c.y, c.z, c.w, r.rf);
                const uint32_t c0 = (uint32_t)d.x, c1 = (uint32_t)d.y;
                bool go0 = false, go1 = false;
                if (o0) {
                    if (c0 & kLeafBit)
                        ray_leaf(c0 & ~kLeafBit, tri64, normals, r, keys);
                    else
                        go0 = true;
                }
                if (o1) {
                    if (c1 & kLeafBit)
                        ray_leaf(c1 & ~kLeafBit, tri64, normals, r, keys);
                    else
                        go1 = true;
                }
                if (go0 && go1) {
                    if (sp == kFastDepth) {
                        ovf = true;
                        break;
                    }
                    stack[sp * 256] = c1;
                    ++sp;
                    cur = c0;
                } else if (go0) {
                    cur = c0;
                } else if (go1) {
                    cur = c1;
                } else if (sp > 0) {
                    --sp;
                    cur = stack[sp * 256];
                } else {
                    break;
                }
            }
        }
    }
    if (ovf || keys.full) {
        const uint32_t slot = atomicAdd(ovfCount, 1u);
        ovfList[slot] = g;  // slot < 3n: every lane appends at most once
        return;
    }
    bits[(size_t)axis * n + i] = (uint8_t)(keys.n & 1);
}

// lanes the fast kernel handed over: (axis, query) encoded as g = axis * n + q
__global__ __launch_bounds__(256) void k_rays_list(const BNode *__restrict__ nodes, uint32_t root,
                                                   const double *__restrict__ tri64,
                                                   const double *__restrict__ normals, const double *__restrict__ pts,
                                                   const uint32_t *__restrict__ order, uint32_t n,
                                                   uint8_t *__restrict__ bits, const uint32_t *__restrict__ list,
                                                   const uint32_t *__restrict__ count, uint32_t *overflow,
                                                   uint32_t maxVisits)
{
    const uint32_t total = *count;
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < total; k += gridDim.x * 256) {
        const uint32_t g = list[k];
        const uint32_t axis = g / n;
        const uint32_t q = g - axis * n;
        const uint32_t i = order ? order[q] : q;
        const Ray r = ray_setup(pts, i, axis);
        KeySet<kListKeys> keys;
        bool ovf = false;
        uint32_t stack[kListStack];
        int sp = 0;
        if (root & kLeafBit)
            ray_leaf(root & ~kLeafBit, tri64, normals, r, keys);
        else
            stack[sp++] = root;
        uint32_t visits = 0;
        while (sp > 0) {
            if (++visits > maxVisits) {
                ovf = true;
                break;
            }
            const BNode *nd = nodes + stack[--sp];
            const float4 a = nd->a, b = nd->b, c = nd->c;
            const int4 d = nd->d;
            const uint32_t c0 = (uint32_t)d.x, c1 = (uint32_t)d.y;
            if (rbox_overlap(b.z, b.w, c.x, c.y, c.z, c.w, r.rf)) {
                if (c1 & kLeafBit)
                    ray_leaf(c1 & ~kLeafBit, tri64, normals, r, keys);
                else if (sp < kListStack)
                    stack[sp++] = c1;
                else
                    ovf = true;
            }
            if (rbox_overlap(a.x, a.y, a.z, a.w, b.x, b.y, r.rf)) {
                if (c0 & kLeafBit)
                    ray_leaf(c0 & ~kLeafBit, tri64, normals, r, keys);
                else if (sp < kListStack)
                    stack[sp++] = c0;
                else
                    ovf = true;
            }
        }
        if (ovf || keys.full)
            atomicOr(overflow, 1u);
        bits[(size_t)axis * n + i] = (uint8_t)(keys.n & 1);
    }
}

// decideGroupSide's vote (src/solidboolean.cpp:492-508): inside iff
// (float)inside / 3 > 0.5, i.e. at least 2 of the 3 rays.
__global__ __launch_bounds__(256) void k_vote(const uint8_t *__restrict__ bits, const uint8_t *__restrict__ skip,
                                              uint32_t n, uint8_t *__restrict__ out)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    if (skip && skip[i]) {
        out[i] = 2;
        return;
    }
    const int inside = bits[i] + bits[n + i] + bits[2 * (size_t)n + i];
    out[i] = ((float)inside / 3.0f > 0.5f) ? 1 : 0;
}

void launch_mark_intersected(const uint64_t *pairs, const uint8_t *status, uint32_t n, int side,
                             uint8_t *flags, hipStream_t s)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_mark_intersected, dim3((n + 255) / 256), dim3(256), 0, s, pairs, status, n, side, flags);
}

void launch_centroids(const double *tri64, const uint8_t *flags, uint32_t nt, double *pts, hipStream_t s)
{
    if (nt == 0)
        return;
    hipLaunchKernelGGL(k_centroids, dim3((nt + 255) / 256), dim3(256), 0, s, tri64, flags, nt, pts);
}

void launch_rays(const BNode *nodes, uint32_t root, const float *rootBox, const double *tri64,
                 const double *normals, uint32_t ntTarget, const double *pts, const uint32_t *order,
                 const uint8_t *skip, uint32_t n, uint8_t *axisBits, uint32_t *ovfList, uint32_t *ovfCount,
                 uint32_t *overflow, hipStream_t s)
{
    if (n == 0)
        return;
    const uint64_t lanes = 3ull * n;
    hipLaunchKernelGGL(k_rays_fast, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, s, nodes, root, rootBox,
                       tri64, normals, pts, order, skip, n, axisBits, ovfList, ovfCount, ntTarget + 1u);
    hipLaunchKernelGGL(k_rays_list, dim3(256), dim3(256), 0, s, nodes, root, tri64, normals, pts, order, n, axisBits,
                       ovfList, ovfCount, overflow, ntTarget + 1u);
}

void launch_vote(const uint8_t *axisBits, const uint8_t *skip, uint32_t n, uint8_t *out, hipStream_t s)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_vote, dim3((n + 255) / 256), dim3(256), 0, s, axisBits, skip, n, out);
}

}  // namespace sbg
