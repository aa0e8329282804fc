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
constexpr int kRefitTile = 1024;  // leaves per workgroup, one per lane

// Scratch shared by the hierarchy and the refit (2n + 1 u32): the top-node
// count, then per sorted leaf position two bytes saying whether a top node's
// range starts there (its suffix union is needed) or ends there (its prefix
// union is needed), then the top-node list.  Only the first two parts start
// at zero.
struct TopWork {
    uint32_t *count;
    uint8_t *need;  // [2 pos]: a range starts at pos, [2 pos + 1]: one ends at pos
    uint32_t *list;
};
__host__ __device__ inline TopWork top_work(uint32_t *work, uint32_t n)
{
    return {work, reinterpret_cast<uint8_t *>(work + 1), work + n};
}
__host__ __device__ inline size_t top_work_zeroed(uint32_t n) { return 1 + ((size_t)n + 1) / 2; }

template <typename K>
__global__ __launch_bounds__(256) void k_hierarchy(const K *__restrict__ codes,
                                                   const uint32_t *__restrict__ ids, int n,
                                                   BNode *__restrict__ nodes, uint32_t *__restrict__ leafParent,
                                                   TopWork top)
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
    if (lo / kRefitTile != hi / kRefitTile) {
        top.list[atomicAdd(top.count, 1u)] = (uint32_t)i;
        top.need[2 * (size_t)lo] = 1;
        top.need[2 * (size_t)hi + 1] = 1;
    }
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

// Refit, stage 1: one workgroup per tile of kRefitTile sorted leaves, one
// leaf per lane.  Every leaf walks up; at a node whose leaf range lies inside
// the tile the first of its two children to arrive leaves its box in LDS and
// stops, the second merges and continues.  All hand-offs stay inside the
// workgroup (LDS + workgroup-scope fences), so no lane ever waits and no
// device-wide fence is needed.  At a node whose range leaves the tile (a
// "top" node) the lane stores its box in the node and stops.
//
// The workgroup also writes the union of the tile's leaf boxes up to a leaf
// position (prefix) where a top node's range ends and from it (suffix) where
// one starts, and the tile's box.  Box unions are exact (min / max), so stage
// 2 can assemble any top node's box from these pieces in any order.
__device__ __forceinline__ void put6(float *dst, const float v[6])
{
#pragma unroll
    for (int k = 0; k < 6; ++k)
        dst[k] = v[k];
}

__global__ __launch_bounds__(kRefitTile) void k_refit_tiles(const uint32_t *__restrict__ ids,
                                                            const float *__restrict__ box32, int n,
                                                            BNode *__restrict__ nodes,
                                                            const uint32_t *__restrict__ leafParent,
                                                            const uint8_t *__restrict__ need,
                                                            float *__restrict__ prefix, float *__restrict__ suffix,
                                                            float *__restrict__ tileBox, float *rootBox)
{
    __shared__ float s_box[kRefitTile][2][6];
    __shared__ uint32_t s_flag[kRefitTile];
    __shared__ float s_wave[kRefitTile / 64][6];
    __shared__ int4 s_d[kRefitTile];  // links of internal nodes [tile0, tile0 + kRefitTile)
    const int tile0 = blockIdx.x * kRefitTile;
    const int tile1 = min(tile0 + kRefitTile, n);
    const int pos = tile0 + threadIdx.x;
    const bool valid = pos < tile1;
    s_flag[threadIdx.x] = 0;
    // A node whose leaf range lies inside the tile has its index inside the
    // tile too (internal node i covers leaves [min(i, j), max(i, j)]), so the
    // walks below read the links of those nodes from one coalesced load here
    // instead of one dependent global load per level.
    if (pos < n - 1)
        s_d[threadIdx.x] = nodes[pos].d;
    __syncthreads();
    float4 lo = make_float4(__builtin_inff(), __builtin_inff(), __builtin_inff(), 0.0f);
    float4 hi = make_float4(-__builtin_inff(), -__builtin_inff(), -__builtin_inff(), 0.0f);
    if (valid) {
        const uint32_t tri = ids[pos];
        lo = reinterpret_cast<const float4 *>(box32 + (size_t)kBox32Stride * tri)[0];
        hi = reinterpret_cast<const float4 *>(box32 + (size_t)kBox32Stride * tri)[1];
        float4 wl = lo, wh = hi;
        int node = (int)leafParent[pos];
        int child = (int)(kLeafBit | tri);
        // Each step moves to a strictly larger leaf range, so a walk ends
        // within the tile's depth; the bound only guards malformed input.
        // The walk reads and writes LDS only: a node inside the tile gets its
        // child boxes from LDS after the walk, and the one global store is the
        // box the walk carries into its top node.  (A global store inside
        // the loop would hold every step's fence until the store lands.)
        bool outside = false;  // the walk stopped at a top node with no links in LDS
        for (int step = 0; step < kRefitTile; ++step) {
            if (node < tile0 || node >= tile0 + kRefitTile || node >= n - 1) {
                outside = true;
                break;
            }
            const int local = node - tile0;
            const int4 d = s_d[local];
            const int slot = (d.x == child) ? 0 : 1;
            const int lo_leaf = min(node, d.w), hi_leaf = max(node, d.w);
            if (lo_leaf < tile0 || hi_leaf >= tile1 || (lo_leaf / kRefitTile) != (hi_leaf / kRefitTile)) {
                store_child_box(nodes + node, slot, wl, wh);
                break;
            }
            float *mine = s_box[local][slot];
            mine[0] = wl.x;
            mine[1] = wl.y;
            mine[2] = wl.z;
            mine[3] = wh.x;
            mine[4] = wh.y;
            mine[5] = wh.z;
            __threadfence_block();
            if (atomicAdd(&s_flag[local], 1u) == 0)
                break;
            __threadfence_block();
            const float *o = s_box[local][1 - slot];
            wl = make_float4(fminf(wl.x, o[0]), fminf(wl.y, o[1]), fminf(wl.z, o[2]), 0.0f);
            wh = make_float4(fmaxf(wh.x, o[3]), fmaxf(wh.y, o[4]), fmaxf(wh.z, o[5]), 0.0f);
            child = node;
            node = d.z;
            if (node < 0) {  // the whole tree fits in this tile
                write_root(rootBox, wl, wh);
                break;
            }
        }
        if (outside) {
            BNode *nd = nodes + node;
            store_child_box(nd, nd->d.x == child ? 0 : 1, wl, wh);
        }
    }
    __syncthreads();  // every walk's LDS boxes
    // nodes inside the tile: both child boxes, from LDS
    if (pos < n - 1) {
        const int4 d = s_d[threadIdx.x];
        const int lo_leaf = min(pos, d.w), hi_leaf = max(pos, d.w);
        if (lo_leaf >= tile0 && hi_leaf < tile1 && (lo_leaf / kRefitTile) == (hi_leaf / kRefitTile)) {
            const float *b0 = s_box[threadIdx.x][0], *b1 = s_box[threadIdx.x][1];
            BNode *nd = nodes + pos;
            nd->a = make_float4(b0[0], b0[1], b0[2], b0[3]);
            nd->b = make_float4(b0[4], b0[5], b1[0], b1[1]);
            nd->c = make_float4(b1[2], b1[3], b1[4], b1[5]);
        }
    }
    // prefix / suffix unions over the tile's positions: wave scans, then the
    // waves' totals
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float pv[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z}, sv[6] = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const float up = __shfl_up(pv[k], off, 64), dn = __shfl_down(sv[k], off, 64);
            if (lane >= off)
                pv[k] = k < 3 ? fminf(pv[k], up) : fmaxf(pv[k], up);
            if (lane + off < 64)
                sv[k] = k < 3 ? fminf(sv[k], dn) : fmaxf(sv[k], dn);
        }
    }
    if (lane == 63)
        put6(s_wave[wave], pv);
    __syncthreads();
    for (int w = 0; w < kRefitTile / 64; ++w) {
        if (w < wave)
#pragma unroll
            for (int k = 0; k < 6; ++k)
                pv[k] = k < 3 ? fminf(pv[k], s_wave[w][k]) : fmaxf(pv[k], s_wave[w][k]);
        if (w > wave)
#pragma unroll
            for (int k = 0; k < 6; ++k)
                sv[k] = k < 3 ? fminf(sv[k], s_wave[w][k]) : fmaxf(sv[k], s_wave[w][k]);
    }
    if (valid) {
        const uchar2 nd2 = reinterpret_cast<const uchar2 *>(need)[pos];
        if (nd2.y)
            put6(prefix + 6 * (size_t)pos, pv);
        if (nd2.x)
            put6(suffix + 6 * (size_t)pos, sv);
    }
    if (threadIdx.x == kRefitTile - 1)
        put6(tileBox + 6 * (size_t)blockIdx.x, pv);
}

// Sparse table over tile boxes: level j holds the union of tiles [t, t + 2^j).
// One workgroup.  Up to kLdsTiles tiles the levels are built in place in LDS
// (read, barrier, write, barrier) and copied out as they complete; beyond that
// each level is a pass over global memory.  The LDS is sized to the tiles at
// launch (24 B per tile): a workgroup asking for the whole 144 KB waits for a
// CU with nothing else resident, which the ray-grid build beside it delays.
constexpr uint32_t kLdsTiles = 6144;  // 144 KB of LDS (gfx950 gives a workgroup 160 KB)
constexpr int kLdsTileRounds = kLdsTiles / 1024;

__global__ __launch_bounds__(1024) void k_tile_table(float *__restrict__ table, uint32_t ntiles, int levels)
{
    extern __shared__ float s_tiles[];  // 6 x ntiles when ntiles <= kLdsTiles
    auto s_t = [&](int k, uint32_t t) -> float & { return s_tiles[(size_t)k * ntiles + t]; };
    if (ntiles <= kLdsTiles) {
        for (uint32_t t = threadIdx.x; t < ntiles; t += 1024)
            for (int k = 0; k < 6; ++k)
                s_t(k, t) = table[6 * (size_t)t + k];
        __syncthreads();
        for (int j = 1; j < levels; ++j) {
            const uint32_t half = 1u << (j - 1);
            float v[kLdsTileRounds][6];
#pragma unroll
            for (int r = 0; r < kLdsTileRounds; ++r) {
                const uint32_t t = threadIdx.x + 1024u * r;
                if (t + 2 * half <= ntiles)
#pragma unroll
                    for (int k = 0; k < 6; ++k)
                        v[r][k] = k < 3 ? fminf(s_t(k, t), s_t(k, t + half)) : fmaxf(s_t(k, t), s_t(k, t + half));
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < kLdsTileRounds; ++r) {
                const uint32_t t = threadIdx.x + 1024u * r;
                if (t + 2 * half <= ntiles)
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        s_t(k, t) = v[r][k];
                        table[6 * ((size_t)j * ntiles + t) + k] = v[r][k];
                    }
            }
            __syncthreads();
        }
        return;
    }
    for (int j = 1; j < levels; ++j) {
        const uint32_t half = 1u << (j - 1);
        const float *src = table + 6 * (size_t)(j - 1) * ntiles;
        float *dst = table + 6 * (size_t)j * ntiles;
        for (uint32_t t = threadIdx.x; t + 2 * half <= ntiles; t += 1024)
            for (int k = 0; k < 6; ++k) {
                const float a = src[6 * (size_t)t + k], b = src[6 * (size_t)(t + half) + k];
                dst[6 * (size_t)t + k] = k < 3 ? fminf(a, b) : fmaxf(a, b);
            }
        __syncthreads();
    }
}

// Refit, stage 2: every top node at once.  A top node's leaf range [lo, hi]
// runs from inside tile tl to inside tile th > tl; its box is the suffix of tl
// at lo, the prefix of th at hi, and (when th > tl + 1) two sparse-table
// entries covering the whole tiles between.  It goes into the parent's child
// slot (or the root box), so nothing waits on anything.
__global__ __launch_bounds__(256) void k_refit_top(BNode *__restrict__ nodes, const uint32_t *__restrict__ topList,
                                                   const uint32_t *__restrict__ topCount,
                                                   const float *__restrict__ prefix, const float *__restrict__ suffix,
                                                   const float *__restrict__ table, uint32_t ntiles, float *rootBox)
{
    const uint32_t ntop = *topCount;
    for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < ntop; k += gridDim.x * 256) {
        const uint32_t p = topList[k];
        const int4 d = nodes[p].d;
        const uint32_t lo = p < (uint32_t)d.w ? p : (uint32_t)d.w;
        const uint32_t hi = p < (uint32_t)d.w ? (uint32_t)d.w : p;
        const uint32_t tl = lo / kRefitTile, th = hi / kRefitTile;
        float v[6];
        const float *a = suffix + 6 * (size_t)lo, *b = prefix + 6 * (size_t)hi;
#pragma unroll
        for (int c = 0; c < 6; ++c)
            v[c] = c < 3 ? fminf(a[c], b[c]) : fmaxf(a[c], b[c]);
        if (th > tl + 1) {
            const uint32_t first = tl + 1, len = th - 1 - first + 1;
            const int j = 31 - __clz(len);
            const float *x = table + 6 * ((size_t)j * ntiles + first);
            const float *y = table + 6 * ((size_t)j * ntiles + (th - 1) - (1u << j) + 1);
#pragma unroll
            for (int c = 0; c < 6; ++c)
                v[c] = c < 3 ? fminf(v[c], fminf(x[c], y[c])) : fmaxf(v[c], fmaxf(x[c], y[c]));
        }
        const float4 blo = make_float4(v[0], v[1], v[2], 0.0f), bhi = make_float4(v[3], v[4], v[5], 0.0f);
        if (d.z < 0) {
            write_root(rootBox, blo, bhi);
        } else {
            BNode *pn = nodes + d.z;
            store_child_box(pn, pn->d.x == (int)p ? 0 : 1, blo, bhi);
        }
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

uint32_t hierarchy_zeroed_u32(uint32_t n) { return (uint32_t)top_work_zeroed(n); }

// work: top_work(work, n), its first top_work_zeroed(n) u32 already zero
void launch_hierarchy(const uint32_t *codes, const uint32_t *ids, uint32_t n, BNode *nodes,
                      uint32_t *leafParent, uint32_t *work, hipStream_t s)
{
    if (n < 2)
        return;
    hipLaunchKernelGGL(k_hierarchy<uint32_t>, dim3((n - 1 + 255) / 256), dim3(256), 0, s, codes, ids, (int)n, nodes,
                       leafParent, top_work(work, n));
}

void launch_hierarchy64(const uint64_t *codes, const uint32_t *ids, uint32_t n, BNode *nodes,
                        uint32_t *leafParent, uint32_t *work, hipStream_t s)
{
    if (n < 2)
        return;
    launch_fill_u32(work, 0u, top_work_zeroed(n), s);
    hipLaunchKernelGGL(k_hierarchy<uint64_t>, dim3((n - 1 + 255) / 256), dim3(256), 0, s, codes, ids, (int)n, nodes,
                       leafParent, top_work(work, n));
}

static int table_levels(uint32_t ntiles)
{
    int L = 1;
    while ((1u << L) <= ntiles)
        ++L;
    return L;
}

size_t refit_scratch_floats(uint32_t n)
{
    const uint32_t ntiles = (n + kRefitTile - 1) / kRefitTile;
    return 12 * (size_t)n + 6 * (size_t)ntiles * table_levels(ntiles) + 6;
}

// scratch: refit_scratch_floats(n) floats (prefix, suffix, tile table)
void launch_refit(const uint32_t *ids, const float *box32, uint32_t n, BNode *nodes,
                  const uint32_t *leafParent, uint32_t *work, float *scratch, float *rootBox, hipStream_t s)
{
    if (n < 2)
        return;
    const uint32_t ntiles = (n + kRefitTile - 1) / kRefitTile;
    float *prefix = scratch, *suffix = scratch + 6 * (size_t)n, *table = scratch + 12 * (size_t)n;
    const TopWork top = top_work(work, n);
    hipLaunchKernelGGL(k_refit_tiles, dim3(ntiles), dim3(kRefitTile), 0, s, ids, box32, (int)n, nodes, leafParent,
                       (const uint8_t *)top.need, prefix, suffix, table, rootBox);
    if (ntiles < 2)
        return;
    const size_t lds = ntiles <= kLdsTiles ? sizeof(float) * 6 * ntiles : 0;
    if (lds > 65536)  // past the default cap on dynamic LDS (the error, if any, surfaces at the launch check)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_tile_table),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_tile_table, dim3(1), dim3(1024), lds, s, table, ntiles, table_levels(ntiles));
    hipLaunchKernelGGL(k_refit_top, dim3(256), dim3(256), 0, s, nodes, (const uint32_t *)top.list,
                       (const uint32_t *)top.count, prefix, suffix, table, ntiles, rootBox);
}

}  // namespace sbg
