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

// 64-bit keys (mesh id above the Morton code, for the combined tree of a batch)
__device__ __forceinline__ int delta(const uint64_t *__restrict__ codes, int n, int i, int j)
{
    if (j < 0 || j >= n)
        return -1;
    const uint64_t a = codes[i], b = codes[j];
    if (a == b)
        return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

// Leaves are refit in tiles of kRefitTile sorted positions; a node whose leaf
// range crosses a tile boundary is a "top" node, finished by k_refit_top.
constexpr int kRefitTile = 1024;
constexpr int kRefitBlock = 256;

template <typename K>
__global__ __launch_bounds__(256) void k_hierarchy(const K *__restrict__ codes,
                                                   const uint32_t *__restrict__ ids, int n,
                                                   BNode *__restrict__ nodes, uint32_t *__restrict__ leafParent,
                                                   uint32_t *__restrict__ topList, uint32_t *topCount)
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
    nodes[i].d.w = j;  // the node covers sorted leaves [min(i, j), max(i, j)]
    if (i == 0)
        nodes[0].d.z = -1;
    if (lo / kRefitTile != hi / kRefitTile)
        topList[atomicAdd(topCount, 1u)] = (uint32_t)i;
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

__device__ __forceinline__ void write_root(float *rootBox, float4 lo, float4 hi)
{
    rootBox[0] = lo.x;
    rootBox[1] = lo.y;
    rootBox[2] = lo.z;
    rootBox[3] = hi.x;
    rootBox[4] = hi.y;
    rootBox[5] = hi.z;
}

// Refit, stage 1: one workgroup per tile of kRefitTile sorted leaves.  Every
// leaf walks up; at a node whose leaf range lies inside the tile the first of
// its two children to arrive leaves its box in LDS and stops, the second
// merges and continues.  All hand-offs stay inside the workgroup (LDS +
// workgroup-scope fences), so no lane ever waits and no device-wide fence is
// needed.  At a node whose range leaves the tile the lane stores its box in
// the node, counts the arrival in topFlags and stops.
__global__ __launch_bounds__(kRefitBlock) void k_refit_tiles(const uint32_t *__restrict__ ids,
                                                             const float *__restrict__ box32, int n,
                                                             BNode *__restrict__ nodes,
                                                             const uint32_t *__restrict__ leafParent,
                                                             uint32_t *__restrict__ topFlags, float *rootBox)
{
    __shared__ float s_box[kRefitTile][2][6];
    __shared__ uint32_t s_flag[kRefitTile];
    const int tile0 = blockIdx.x * kRefitTile;
    const int tile1 = min(tile0 + kRefitTile, n);
    for (int k = threadIdx.x; k < kRefitTile; k += kRefitBlock)
        s_flag[k] = 0;
    __syncthreads();
    for (int pos = tile0 + threadIdx.x; pos < tile1; pos += kRefitBlock) {
        const uint32_t tri = ids[pos];
        float4 lo = reinterpret_cast<const float4 *>(box32 + (size_t)kBox32Stride * tri)[0];
        float4 hi = reinterpret_cast<const float4 *>(box32 + (size_t)kBox32Stride * tri)[1];
        int node = (int)leafParent[pos];
        int child = (int)(kLeafBit | tri);
        // each step moves to a strictly larger leaf range, so a walk ends
        // within the tile's depth; the bound only guards malformed input
        for (int step = 0; step < kRefitTile; ++step) {
            BNode *nd = nodes + node;
            const int4 d = nd->d;
            const int slot = (d.x == child) ? 0 : 1;
            store_child_box(nd, slot, lo, hi);
            const int lo_leaf = min(node, d.w), hi_leaf = max(node, d.w);
            if (lo_leaf < tile0 || hi_leaf >= tile1 || (lo_leaf / kRefitTile) != (hi_leaf / kRefitTile)) {
                atomicAdd(topFlags + node, 1u);
                break;
            }
            const int local = node - tile0;
            float *mine = s_box[local][slot];
            mine[0] = lo.x;
            mine[1] = lo.y;
            mine[2] = lo.z;
            mine[3] = hi.x;
            mine[4] = hi.y;
            mine[5] = hi.z;
            __threadfence_block();
            if (atomicAdd(&s_flag[local], 1u) == 0)
                break;
            __threadfence_block();
            const volatile float *o = s_box[local][1 - slot];
            lo = make_float4(fminf(lo.x, o[0]), fminf(lo.y, o[1]), fminf(lo.z, o[2]), 0.0f);
            hi = make_float4(fmaxf(hi.x, o[3]), fmaxf(hi.y, o[4]), fmaxf(hi.z, o[5]), 0.0f);
            child = node;
            node = d.z;
            if (node < 0) {  // the whole tree fits in this tile
                write_root(rootBox, lo, hi);
                break;
            }
        }
    }
}

// Refit, stage 2: one workgroup finishes the top nodes (those spanning tiles)
// from a ready list.  A top node is ready once both children are complete
// (topFlags == 2).  Each round merges every ready node's two child boxes into
// its parent's slot and counts the arrival in the parent's flag; the arrival
// that brings a parent to 2 appends it to the next round's list.  Every node
// is merged once, so the work is O(ntop) and the number of rounds is the
// height of the top tree.  All hand-offs stay inside the workgroup (barriers
// and workgroup-scope fences); the loop ends when a round's list is empty,
// which the root's merge guarantees within ntop + 1 rounds.
__global__ __launch_bounds__(1024) void k_refit_top(BNode *__restrict__ nodes, const uint32_t *__restrict__ topList,
                                                    const uint32_t *__restrict__ topCount, uint32_t *topFlags,
                                                    uint32_t *__restrict__ readyA, uint32_t *__restrict__ readyB,
                                                    float *rootBox)
{
    __shared__ uint32_t s_n[2];
    const uint32_t ntop = *topCount;
    if (threadIdx.x == 0) {
        s_n[0] = 0;
        s_n[1] = 0;
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < ntop; k += 1024) {
        const uint32_t p = topList[k];
        if (topFlags[p] == 2)
            readyA[atomicAdd(&s_n[0], 1u)] = p;
    }
    int cur = 0;
    for (uint32_t round = 0; round <= ntop; ++round) {
        __syncthreads();
        const uint32_t m = s_n[cur];
        if (m == 0)
            break;
        uint32_t *in = cur ? readyB : readyA;
        uint32_t *out = cur ? readyA : readyB;
        if (threadIdx.x == 0)
            s_n[cur ^ 1] = 0;
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < m; k += 1024) {
            const uint32_t p = in[k];
            const volatile float *b = reinterpret_cast<const volatile float *>(nodes + p);
            const float4 lo = make_float4(fminf(b[0], b[6]), fminf(b[1], b[7]), fminf(b[2], b[8]), 0.0f);
            const float4 hi = make_float4(fmaxf(b[3], b[9]), fmaxf(b[4], b[10]), fmaxf(b[5], b[11]), 0.0f);
            const int parent = nodes[p].d.z;
            if (parent < 0) {
                write_root(rootBox, lo, hi);
                continue;
            }
            BNode *pn = nodes + parent;
            store_child_box(pn, pn->d.x == (int)p ? 0 : 1, lo, hi);
            __threadfence_block();
            if (atomicAdd(topFlags + parent, 1u) == 1u)
                out[atomicAdd(&s_n[cur ^ 1], 1u)] = (uint32_t)parent;
        }
        cur ^= 1;
    }
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

__global__ void k_copy_u32(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        dst[i] = src[i];
}

void launch_copy_u32(const uint32_t *src, uint32_t *dst, size_t n, hipStream_t s)
{
    if (n)
        hipLaunchKernelGGL(k_copy_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
}

void launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, value, n);
}

// work: 4n + 1 u32 = top flags [0, n-1), top list [n, 2n-1), top count [2n],
// two ready lists of the top refit [2n + 1, 3n + 1) and [3n + 1, 4n + 1)
void launch_hierarchy(const uint32_t *codes, const uint32_t *ids, uint32_t n, BNode *nodes,
                      uint32_t *leafParent, uint32_t *work, hipStream_t s)
{
    if (n < 2)
        return;
    launch_fill_u32(work, 0u, (size_t)2 * n + 1, s);
    hipLaunchKernelGGL(k_hierarchy<uint32_t>, dim3((n - 1 + 255) / 256), dim3(256), 0, s, codes, ids, (int)n, nodes,
                       leafParent, work + n, work + 2 * (size_t)n);
}

void launch_hierarchy64(const uint64_t *codes, const uint32_t *ids, uint32_t n, BNode *nodes,
                        uint32_t *leafParent, uint32_t *work, hipStream_t s)
{
    if (n < 2)
        return;
    launch_fill_u32(work, 0u, (size_t)2 * n + 1, s);
    hipLaunchKernelGGL(k_hierarchy<uint64_t>, dim3((n - 1 + 255) / 256), dim3(256), 0, s, codes, ids, (int)n, nodes,
                       leafParent, work + n, work + 2 * (size_t)n);
}

void launch_refit(const uint32_t *ids, const float *box32, uint32_t n, BNode *nodes,
                  const uint32_t *leafParent, uint32_t *work, float *rootBox, hipStream_t s)
{
    if (n < 2)
        return;
    hipLaunchKernelGGL(k_refit_tiles, dim3((n + kRefitTile - 1) / kRefitTile), dim3(kRefitBlock), 0, s, ids, box32,
                       (int)n, nodes, leafParent, work, rootBox);
    hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(1024), 0, s, nodes, work + n, work + 2 * (size_t)n, work,
                       work + 2 * (size_t)n + 1, work + 3 * (size_t)n + 1, rootBox);
}

}  // namespace sbg
