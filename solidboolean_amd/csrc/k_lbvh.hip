// Linear BVH over Morton-sorted triangles (Karras, "Maximizing Parallelism in
// the Construction of BVHs, Octrees, and k-d Trees", HPG 2012), replacing the
// reference's recursive top-down tree (src/axisalignedboundingboxtree.cpp:27-141).
// The pair set found against it is the same: every node box contains its
// members' boxes in both trees, and leaf pairs are rechecked with the exact FP64
// closed test (src/axisalignedboundingboxtree.h:62-67).
//
// Node i (0 <= i < n-1) is internal; node 0 is the root.  Leaves are referenced
// as kLeafBit | triangle, so traversal never needs a leaf -> triangle lookup.
// The refit stores each child's FP32 box inside its parent, so one 64-B node
// load yields both child boxes.
#include "sbg_internal.h"

namespace sbg {

// common-prefix length of sorted keys i and j, ties broken by index
__device__ __forceinline__ int delta(const uint32_t *__restrict__ codes, int n, int i, int j)
{
    if (j < 0 || j >= n)
        return -1;
    const uint32_t a = codes[i], b = codes[j];
    if (a == b)
        return 32 + __clz((uint32_t)(i ^ j));
    return __clz(a ^ b);
}

__global__ __launch_bounds__(256) void k_hierarchy(const uint32_t *__restrict__ codes,
                                                   const uint32_t *__restrict__ ids, int n,
                                                   BNode *__restrict__ nodes, uint32_t *__restrict__ leafParent)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n - 1)
        return;
    const int d = (delta(codes, n, i, i + 1) - delta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(codes, n, i, i - d);
    int lmax = 2;
    while (delta(codes, n, i, i + lmax * d) > dmin)
        lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(codes, n, i, i + (l + t) * d) > dmin)
            l += t;
    const int j = i + l * d;
    const int dnode = delta(codes, n, i, j);
    int s = 0;
    int t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(codes, n, i, i + (s + t) * d) > dnode)
            s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j;
    const int hi = i < j ? j : i;
    int c0, c1;
    if (lo == gamma) {
        c0 = (int)(kLeafBit | ids[gamma]);
        leafParent[gamma] = (uint32_t)i;
    } else {
        c0 = gamma;
        nodes[gamma].d.z = i;
    }
    if (hi == gamma + 1) {
        c1 = (int)(kLeafBit | ids[gamma + 1]);
        leafParent[gamma + 1] = (uint32_t)i;
    } else {
        c1 = gamma + 1;
        nodes[gamma + 1].d.z = i;
    }
    nodes[i].d.x = c0;
    nodes[i].d.y = c1;
    if (i == 0)
        nodes[0].d.z = -1;
}

__device__ __forceinline__ void store_child_box(BNode *nd, int slot, float4 lo, float4 hi)
{
    if (slot == 0) {
        nd->a = make_float4(lo.x, lo.y, lo.z, hi.x);
        nd->b.x = hi.y;
        nd->b.y = hi.z;
    } else {
        nd->b.z = lo.x;
        nd->b.w = lo.y;
        nd->c = make_float4(lo.z, hi.x, hi.y, hi.z);
    }
}

// one thread per leaf position; the second thread to reach a node merges it
__global__ __launch_bounds__(256) void k_refit(const uint32_t *__restrict__ ids, const float *__restrict__ box32,
                                               int n, BNode *nodes, const uint32_t *__restrict__ leafParent,
                                               uint32_t *flags, float *rootBox)
{
    const int pos = blockIdx.x * 256 + threadIdx.x;
    if (pos >= n)
        return;
    const uint32_t tri = ids[pos];
    float4 lo = reinterpret_cast<const float4 *>(box32 + (size_t)kBox32Stride * tri)[0];
    float4 hi = reinterpret_cast<const float4 *>(box32 + (size_t)kBox32Stride * tri)[1];
    int node = (int)leafParent[pos];
    int child = (int)(kLeafBit | tri);
    // a Karras tree over n leaves is at most n - 1 deep; the bound keeps every
    // lane's walk finite whatever the input
    for (int step = 0; node >= 0 && step < n; ++step) {
        BNode *nd = nodes + node;
        const int slot = (nd->d.x == child) ? 0 : 1;
        store_child_box(nd, slot, lo, hi);
        __threadfence();
        const uint32_t prev = __hip_atomic_fetch_add(flags + node, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == 0)
            return;  // the sibling's thread finishes this node
        __threadfence();
        // the sibling's box was written by another lane, possibly on another
        // XCD: read it through volatile loads after the acquire fence
        const volatile float *f = reinterpret_cast<const volatile float *>(nd);
        lo = make_float4(fminf(f[0], f[6]), fminf(f[1], f[7]), fminf(f[2], f[8]), 0.0f);
        hi = make_float4(fmaxf(f[3], f[9]), fmaxf(f[4], f[10]), fmaxf(f[5], f[11]), 0.0f);
        child = node;
        node = nd->d.z;
    }
    if (node >= 0)
        return;
    rootBox[0] = lo.x;
    rootBox[1] = lo.y;
    rootBox[2] = lo.z;
    rootBox[3] = hi.x;
    rootBox[4] = hi.y;
    rootBox[5] = hi.z;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// plain streaming copy, 16 B per lane per iteration (counter calibration)
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

void launch_copy16(const void *src, void *dst, size_t n16, hipStream_t s)
{
    if (n16 == 0)
        return;
    hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst, n16);
}

void launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, value, n);
}

void launch_hierarchy(const uint32_t *codes, const uint32_t *ids, uint32_t n, BNode *nodes,
                      uint32_t *leafParent, hipStream_t s)
{
    if (n < 2)
        return;
    hipLaunchKernelGGL(k_hierarchy, dim3((n - 1 + 255) / 256), dim3(256), 0, s, codes, ids, (int)n, nodes,
                       leafParent);
}

void launch_refit(const uint32_t *ids, const float *box32, uint32_t n, BNode *nodes,
                  const uint32_t *leafParent, uint32_t *flags, float *rootBox, hipStream_t s)
{
    if (n < 2)
        return;
    launch_fill_u32(flags, 0u, n - 1, s);
    hipLaunchKernelGGL(k_refit, dim3((n + 255) / 256), dim3(256), 0, s, ids, box32, (int)n, nodes, leafParent,
                       flags, rootBox);
}

}  // namespace sbg
