// Tree-vs-tree candidate search: the set the reference's simultaneous DFS emits
// (src/axisalignedboundingboxtree.h:54-95 via src/solidboolean.cpp:94-101),
// i.e. every (a, b) whose FP64 triangle boxes closed-overlap.
//
// Breadth-first over node pairs: each level's kernel reads the frontier written
// by the previous one.  A thread loads the two nodes of its pair (64 B each,
// both child boxes inside), tests the up-to-4 child combinations with the FP32
// outward-rounded boxes (a superset test), and appends surviving internal pairs
// to the next frontier.  Leaf-leaf survivors are rechecked with the exact FP64
// closed box test on the 96-B FP64 records before they are appended to the
// candidate queue.  Appends are wave-aggregated: one atomic per wave per queue,
// offsets from a 64-lane prefix sum.  Queues past capacity raise `overflow`;
// the host grows them and repeats the search, so nothing is ever dropped.
#include "sbg_device.h"
#include "sbg_internal.h"

#pragma clang fp contract(off)

namespace sbg {

struct FBox {
    float lx, ly, lz, hx, hy, hz;
};

__device__ __forceinline__ bool fbox_overlap(const FBox &a, const FBox &b)
{
    return a.lx <= b.hx && a.hx >= b.lx && a.ly <= b.hy && a.hy >= b.ly && a.lz <= b.hz && a.hz >= b.lz;
}

__device__ __forceinline__ FBox leaf_box(const float *__restrict__ box32, uint32_t tri)
{
    const float4 *p = reinterpret_cast<const float4 *>(box32 + (size_t)kBox32Stride * tri);
    const float4 lo = p[0], hi = p[1];
    return {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
}

__device__ __forceinline__ Box64 record_box(const double *__restrict__ tri64, uint32_t t)
{
    const double2 *p = reinterpret_cast<const double2 *>(tri64 + (size_t)kTri64Stride * t);
    const double2 r0 = p[0], r1 = p[1], r2 = p[2], r3 = p[3], r4 = p[4], r5 = p[5];
    return tri_box64({r0.x, r0.y, r1.x}, {r2.x, r2.y, r3.x}, {r4.x, r4.y, r5.x});
}

// wave-wide exclusive prefix of `cnt`; returns this lane's offset in the queue
// (one atomic per wave) or ~0u when the queue would overflow.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t cnt, uint32_t *counter, uint32_t cap, uint32_t *overflow,
                                                 uint32_t &waveBaseOut)
{
    const int lane = threadIdx.x & 63;
    uint32_t incl = cnt;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= off)
            incl += v;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    uint32_t base = 0;
    if (lane == 63 && total)
        base = atomicAdd(counter, total);
    base = __shfl(base, 63, 64);
    waveBaseOut = base;
    if (total && base + total > cap && lane == 0)
        atomicOr(overflow, 1u);
    return base + incl - cnt;
}

// Children of a node reference: one (a leaf) or two, with their FP32 boxes.
struct Kids {
    uint32_t c0, c1;
    FBox b0, b1;
    bool two;
};

__device__ __forceinline__ Kids kids_of(const BNode *__restrict__ nodes, const float *__restrict__ box32, uint32_t ref)
{
    Kids k;
    if (ref & kLeafBit) {
        k.c0 = k.c1 = ref;
        k.b0 = k.b1 = leaf_box(box32, ref & ~kLeafBit);
        k.two = false;
        return k;
    }
    const BNode *nd = nodes + ref;
    const float4 a = nd->a, b = nd->b, cc = nd->c;
    const int4 d = nd->d;
    k.c0 = (uint32_t)d.x;
    k.c1 = (uint32_t)d.y;
    k.b0 = {a.x, a.y, a.z, a.w, b.x, b.y};
    k.b1 = {b.z, b.w, cc.x, cc.y, cc.z, cc.w};
    k.two = true;
    return k;
}

// Expansion of one node pair: child pair s = 2u + v (u over A's children, v
// over B's) goes to the frontier (bit s of next) or, when both are leaves
// whose exact FP64 boxes overlap, to the candidates (bit s of cand).  Every
// field is a named register, so nothing lands in scratch memory.
struct Expansion {
    uint32_t a0, a1, b0, b1;
    uint32_t next, cand;
};

// 0: pruned, 1: to the frontier, 2: candidate
__device__ __forceinline__ uint32_t expand_one(const TraverseArgs &t, uint32_t ca, uint32_t cb, const FBox &ba,
                                               const FBox &bb)
{
    if (!fbox_overlap(ba, bb))
        return 0u;
    if ((ca & kLeafBit) && (cb & kLeafBit))
        return box64_overlap(record_box(t.tri64A, ca & ~kLeafBit), record_box(t.tri64B, cb & ~kLeafBit)) ? 2u : 0u;
    return 1u;
}

__device__ __forceinline__ Expansion expand_pair(const TraverseArgs &t, uint2 pr, bool valid)
{
    Expansion x = {0, 0, 0, 0, 0, 0};
    if (!valid)
        return x;
    const Kids ka = kids_of(t.nodesA, t.boxA, pr.x);
    const Kids kb = kids_of(t.nodesB, t.boxB, pr.y);
    x.a0 = ka.c0;
    x.a1 = ka.c1;
    x.b0 = kb.c0;
    x.b1 = kb.c1;
    const uint32_t r0 = expand_one(t, ka.c0, kb.c0, ka.b0, kb.b0);
    const uint32_t r1 = kb.two ? expand_one(t, ka.c0, kb.c1, ka.b0, kb.b1) : 0u;
    const uint32_t r2 = ka.two ? expand_one(t, ka.c1, kb.c0, ka.b1, kb.b0) : 0u;
    const uint32_t r3 = ka.two && kb.two ? expand_one(t, ka.c1, kb.c1, ka.b1, kb.b1) : 0u;
    x.next = (r0 == 1u ? 1u : 0u) | (r1 == 1u ? 2u : 0u) | (r2 == 1u ? 4u : 0u) | (r3 == 1u ? 8u : 0u);
    x.cand = (r0 == 2u ? 1u : 0u) | (r1 == 2u ? 2u : 0u) | (r2 == 2u ? 4u : 0u) | (r3 == 2u ? 8u : 0u);
    return x;
}

// child pair of slot s (a compile-time constant at every use)
__device__ __forceinline__ uint2 slot_pair(const Expansion &x, int s)
{
    return make_uint2(s < 2 ? x.a0 : x.a1, (s & 1) ? x.b1 : x.b0);
}

__device__ __forceinline__ uint64_t cand_key(const Expansion &x, int s)
{
    const uint2 p = slot_pair(x, s);
    return ((uint64_t)(p.x & ~kLeafBit) << 32) | (p.y & ~kLeafBit);
}

// Workgroup size of the grid-wide breadth-first levels (the launches' grids
// are given in units of 256 lanes).
constexpr int kExpandThreads = 256;
constexpr int kExpandWaves = kExpandThreads / 64;

// Offsets of this lane's na next-frontier pairs and nb candidates in their
// queues, for a workgroup of kExpandThreads lanes: the two counts travel
// packed in one 32-bit prefix sum (at most 4 each per lane, so a workgroup's
// totals stay below 2^16), and the workgroup takes one atomic per queue.
// Every lane of the workgroup calls it.
__device__ __forceinline__ uint2 group_reserve2(uint32_t na, uint32_t nb, uint32_t *countA, uint32_t capA,
                                                uint32_t *countB, uint32_t capB, uint32_t *overflow)
{
    __shared__ uint32_t s_wave[kExpandWaves], s_pre[kExpandWaves], s_base[2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t mine = na | nb << 16;
    uint32_t incl = mine;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= off)
            incl += v;
    }
    if (lane == 63)
        s_wave[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < kExpandWaves; ++w) {
            s_pre[w] = run;
            run += s_wave[w];
        }
        const uint32_t ta = run & 0xffffu, tb = run >> 16;
        const uint32_t ba = ta ? atomicAdd(countA, ta) : 0u, bb = tb ? atomicAdd(countB, tb) : 0u;
        if ((ta && ba + ta > capA) || (tb && bb + tb > capB))
            atomicOr(overflow, 1u);
        s_base[0] = ba;
        s_base[1] = bb;
    }
    __syncthreads();
    const uint32_t off = s_pre[wave] + incl - mine;
    return make_uint2(s_base[0] + (off & 0xffffu), s_base[1] + (off >> 16));
}

// One breadth-first level over the whole grid: frontier `in` (inCount pairs)
// to `out`, candidates appended.  Workgroups of kExpandThreads lanes.
__device__ __forceinline__ void expand_level(const TraverseArgs &t, const uint2 *__restrict__ in,
                                             const uint32_t *__restrict__ inCount, uint2 *__restrict__ out,
                                             uint32_t *outCount, uint32_t outCap, uint64_t *__restrict__ cand,
                                             uint32_t *candCount, uint32_t candCap, uint32_t *overflow)
{
    // both frontier buffers hold outCap pairs; a count past it has already
    // raised `overflow` and the whole search is repeated with larger queues
    const uint32_t n = *inCount < outCap ? *inCount : outCap;
    const uint32_t stride = gridDim.x * kExpandThreads;
    for (uint32_t base = blockIdx.x * kExpandThreads; base < n; base += stride) {  // uniform over the workgroup
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < n;
        const Expansion x = expand_pair(t, valid ? in[i] : make_uint2(0, 0), valid);
        const uint2 o = group_reserve2(__popc(x.next), __popc(x.cand), outCount, outCap, candCount, candCap,
                                       overflow);
        const uint32_t on = o.x, oc = o.y;
        uint32_t kn = 0, kc = 0;
#pragma unroll
        for (int sl = 0; sl < 4; ++sl) {
            if ((x.next >> sl) & 1u) {
                if (on + kn < outCap)
                    out[on + kn] = slot_pair(x, sl);
                ++kn;
            }
            if ((x.cand >> sl) & 1u) {
                if (oc + kc < candCap)
                    cand[oc + kc] = cand_key(x, sl);
                ++kc;
            }
        }
    }
}

__global__ __launch_bounds__(kExpandThreads) void k_expand(TraverseArgs t, const uint2 *__restrict__ in,
                                                const uint32_t *__restrict__ inCount, uint2 *__restrict__ out,
                                                uint32_t *outCount, uint32_t outCap, uint64_t *__restrict__ cand,
                                                uint32_t *candCount, uint32_t candCap, uint32_t *overflow)
{
    expand_level(t, in, inCount, out, outCount, outCap, cand, candCount, candCap, overflow);
}

// ---- top of the search in one workgroup -------------------------------------
// The first levels of a tree-vs-tree search hold a handful of node pairs, and a
// grid-wide launch per level costs a launch's latency for a few lanes of work
// (and waits for free slots behind any kernel that fills the chip).  One
// workgroup of 1024 lanes expands level after level instead, frontier in the
// two global frontier buffers, next-frontier slots handed out from LDS, until
// a level holds more than topMax pairs, the level cap is reached or the
// frontier empties.  cnt[l] gets every level's size as the grid-wide kernel
// would leave it; *levelOut gets the first level not expanded here.
__global__ __launch_bounds__(1024) void k_search_top(TraverseArgs t, uint2 *__restrict__ front0,
                                                     uint2 *__restrict__ front1, uint32_t *__restrict__ cnt,
                                                     uint32_t fcap, uint64_t *__restrict__ cand, uint32_t *candCount,
                                                     uint32_t candCap, uint32_t *overflow, int maxLevels,
                                                     uint32_t topMax, uint32_t *levelOut)
{
    __shared__ uint32_t s_next;
    const int lane = threadIdx.x & 63;
    uint32_t nIn = cnt[0];
    int level = 0;
    while (level < maxLevels && nIn > 0 && nIn <= topMax) {
        const uint2 *in = (level & 1) ? front1 : front0;
        uint2 *out = (level & 1) ? front0 : front1;
        if (threadIdx.x == 0)
            s_next = 0;
        __syncthreads();
        for (uint32_t b = 0; b < nIn; b += 1024) {
            const uint32_t i = b + threadIdx.x;
            const bool valid = i < nIn;
            const Expansion x = expand_pair(t, valid ? in[i] : make_uint2(0, 0), valid);
            // next-frontier slots: a wave prefix, one LDS atomic per wave
            const uint32_t nn = __popc(x.next);
            uint32_t incl = nn;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t v = __shfl_up(incl, off, 64);
                if (lane >= off)
                    incl += v;
            }
            const uint32_t total = __shfl(incl, 63, 64);
            uint32_t wbase = 0;
            if (lane == 63 && total)
                wbase = atomicAdd(&s_next, total);
            wbase = __shfl(wbase, 63, 64);
            const uint32_t on = wbase + incl - nn;
            if (total && wbase + total > fcap && lane == 0)
                atomicOr(overflow, 1u);
            uint32_t wb;
            const uint32_t oc = wave_reserve(__popc(x.cand), candCount, candCap, overflow, wb);
            uint32_t kn = 0, kc = 0;
#pragma unroll
            for (int sl = 0; sl < 4; ++sl) {
                if ((x.next >> sl) & 1u) {
                    if (on + kn < fcap)
                        out[on + kn] = slot_pair(x, sl);
                    ++kn;
                }
                if ((x.cand >> sl) & 1u) {
                    if (oc + kc < candCap)
                        cand[oc + kc] = cand_key(x, sl);
                    ++kc;
                }
            }
        }
        // the next level reads this one's output: workgroup-scope ordering of
        // the global frontier writes, and the count from LDS
        __syncthreads();
        nIn = s_next;
        if (threadIdx.x == 0)
            cnt[level + 1] = nIn;
        ++level;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *levelOut = (uint32_t)level;
}

// Breadth-first level *topLevel + j over the whole grid, or nothing when that
// is at or past lastLevel (the launches after the top kernel are queued before
// it is known where it stopped).
__global__ __launch_bounds__(kExpandThreads) void k_expand_at(TraverseArgs t, uint2 *__restrict__ front0,
                                                   uint2 *__restrict__ front1, uint32_t *__restrict__ cnt,
                                                   const uint32_t *__restrict__ topLevel, uint32_t j,
                                                   uint32_t lastLevel, uint32_t fcap, uint64_t *__restrict__ cand,
                                                   uint32_t *candCount, uint32_t candCap, uint32_t *overflow)
{
    const uint32_t level = *topLevel + j;
    if (level >= lastLevel)
        return;
    const uint2 *in = (level & 1) ? front1 : front0;
    uint2 *out = (level & 1) ? front0 : front1;
    expand_level(t, in, cnt + level, out, cnt + level + 1, fcap, cand, candCount, candCap, overflow);
}

// Persistent depth-first stage.  One wave per workgroup keeps a stack of node
// pairs in LDS (spilling to a per-wave region of global memory when full).
// Each step pops up to 64 pairs (one per lane), expands them as k_expand
// does, appends candidates with one atomic per wave, and pushes the surviving
// child pairs back with a 64-lane prefix sum.  An empty wave claims the next
// few seed pairs (the breadth-first frontier) from a shared counter.  Waves never wait for
// each other: a wave leaves when its own stacks are empty and no seed is left.
// Spill overflow or the iteration cap raise `overflow` (the host then repeats
// the search breadth-first), so nothing is dropped silently.
constexpr int kDfsLds = 1024;
constexpr uint32_t kDfsMaxIter = 1u << 24;

// With levelBase, the seeds are the frontier of level min(*levelBase +
// levelAdd, lastLevel): `seeds` / `seedsAlt` are the even / odd levels'
// frontier buffers and seedCount the per-level counts.
__global__ __launch_bounds__(64) void k_traverse_dfs(TraverseArgs t, const uint2 *__restrict__ seeds,
                                                     const uint2 *__restrict__ seedsAlt,
                                                     const uint32_t *__restrict__ seedCount, uint32_t seedCap,
                                                     const uint32_t *__restrict__ levelBase, uint32_t levelAdd,
                                                     uint32_t lastLevel, uint32_t *head, uint2 *__restrict__ spill,
                                                     uint32_t spillPerWave, uint64_t *__restrict__ cand,
                                                     uint32_t *candCount, uint32_t candCap, uint32_t *overflow)
{
    __shared__ uint2 s_stack[kDfsLds];
    const int lane = threadIdx.x;
    if (levelBase) {
        uint32_t level = *levelBase + levelAdd;
        level = level < lastLevel ? level : lastLevel;
        seeds = (level & 1) ? seedsAlt : seeds;
        seedCount += level;
    }
    const uint32_t nseed = *seedCount < seedCap ? *seedCount : seedCap;
    // seeds per claim: enough for every wave to start busy, at most one per lane
    uint32_t grab = nseed / gridDim.x;
    grab = grab < 1u ? 1u : (grab > 64u ? 64u : grab);
    uint2 *gstack = spill + (size_t)blockIdx.x * spillPerWave;
    int sp = 0;          // LDS entries (same value in every lane)
    uint32_t gsp = 0;    // spilled entries (same value in every lane)
    uint32_t iter = 0;
    for (; iter < kDfsMaxIter; ++iter) {
        if (sp == 0 && gsp > 0) {
            const uint32_t m = gsp < (uint32_t)(kDfsLds / 2) ? gsp : (uint32_t)(kDfsLds / 2);
            for (uint32_t k = lane; k < m; k += 64)
                s_stack[k] = gstack[gsp - m + k];
            gsp -= m;
            sp = (int)m;
            __syncthreads();
        }
        if (sp == 0) {
            uint32_t start = 0;
            if (lane == 0)
                start = atomicAdd(head, grab);
            start = __shfl(start, 0, 64);
            if (start >= nseed)
                break;
            const uint32_t m = nseed - start < grab ? nseed - start : grab;
            if ((uint32_t)lane < m)
                s_stack[lane] = seeds[start + lane];
            sp = (int)m;
            __syncthreads();
        }
        const int take = sp < 64 ? sp : 64;
        const bool valid = lane < take;
        uint2 pr = make_uint2(0, 0);
        if (valid)
            pr = s_stack[sp - take + lane];
        sp -= take;
        __syncthreads();
        const Expansion x = expand_pair(t, pr, valid);
        const uint32_t nn = __popc(x.next);
        uint32_t wb;
        const uint32_t oc = wave_reserve(__popc(x.cand), candCount, candCap, overflow, wb);
        uint32_t kc = 0;
#pragma unroll
        for (int sl = 0; sl < 4; ++sl)
            if ((x.cand >> sl) & 1u) {
                if (oc + kc < candCap)
                    cand[oc + kc] = cand_key(x, sl);
                ++kc;
            }
        // push children: LDS when they fit, else the wave's spill region
        uint32_t incl = nn;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t v = __shfl_up(incl, off, 64);
            if (lane >= off)
                incl += v;
        }
        const uint32_t total = __shfl(incl, 63, 64);
        const uint32_t mine = incl - nn;
        if (sp + (int)total <= kDfsLds) {
            uint32_t kn = 0;
#pragma unroll
            for (int sl = 0; sl < 4; ++sl)
                if ((x.next >> sl) & 1u)
                    s_stack[sp + mine + kn++] = slot_pair(x, sl);
            sp += (int)total;
        } else if (gsp + total <= spillPerWave) {
            uint32_t kn = 0;
#pragma unroll
            for (int sl = 0; sl < 4; ++sl)
                if ((x.next >> sl) & 1u)
                    gstack[gsp + mine + kn++] = slot_pair(x, sl);
            gsp += total;
        } else {
            if (lane == 0)
                atomicOr(overflow, 4u);
            break;
        }
        __syncthreads();
    }
    if (iter == kDfsMaxIter && lane == 0)
        atomicOr(overflow, 8u);
}

void launch_traverse_dfs(const TraverseArgs &t, const uint2 *seeds, const uint32_t *seedCount, uint32_t seedCap,
                         uint32_t *head, uint2 *spill, uint32_t spillPerWave, int waves, uint64_t *cand,
                         uint32_t *candCount, uint32_t candCap, uint32_t *overflow, hipStream_t s)
{
    hipLaunchKernelGGL(k_traverse_dfs, dim3(waves), dim3(64), 0, s, t, seeds, seeds, seedCount, seedCap,
                       (const uint32_t *)nullptr, 0u, 0u, head, spill, spillPerWave, cand, candCount, candCap,
                       overflow);
}

// The single-pair search after k_seed: the top kernel, `extra` grid-wide
// levels after it (each a no-op past lastLevel), then the depth-first stage
// from the level reached.  levelWord: one u32 of scratch.
void launch_search_levels(const TraverseArgs &t, uint2 *front0, uint2 *front1, uint32_t *cnt, uint32_t fcap,
                          uint32_t topMax, int lastLevel, int extra, uint32_t *levelWord, uint32_t *head,
                          uint2 *spill, uint32_t spillPerWave, int waves, uint64_t *cand, uint32_t *candCount,
                          uint32_t candCap, uint32_t *overflow, int expandGrid, hipStream_t s)
{
    hipLaunchKernelGGL(k_search_top, dim3(1), dim3(1024), 0, s, t, front0, front1, cnt, fcap, cand, candCount,
                       candCap, overflow, lastLevel, topMax, levelWord);
    for (int j = 0; j < extra; ++j)
        hipLaunchKernelGGL(k_expand_at, dim3(expandGrid * 256 / kExpandThreads), dim3(kExpandThreads), 0, s, t,
                           front0, front1, cnt,
                           (const uint32_t *)levelWord, (uint32_t)j, (uint32_t)lastLevel, fcap, cand, candCount,
                           candCap, overflow);
    hipLaunchKernelGGL(k_traverse_dfs, dim3(waves), dim3(64), 0, s, t, (const uint2 *)front0, (const uint2 *)front1,
                       (const uint32_t *)cnt, fcap, (const uint32_t *)levelWord, (uint32_t)extra,
                       (uint32_t)lastLevel, head, spill, spillPerWave, cand, candCount, candCap, overflow);
}

// root pair: both meshes non-empty; a root may itself be a leaf (1 triangle)
__global__ void k_seed(TraverseArgs t, uint32_t rootA, uint32_t rootB, const float *__restrict__ rootBoxA,
                       const float *__restrict__ rootBoxB, uint2 *frontier, uint32_t *count, uint64_t *cand,
                       uint32_t *candCount)
{
    if (threadIdx.x != 0)
        return;
    FBox a = {rootBoxA[0], rootBoxA[1], rootBoxA[2], rootBoxA[3], rootBoxA[4], rootBoxA[5]};
    FBox b = {rootBoxB[0], rootBoxB[1], rootBoxB[2], rootBoxB[3], rootBoxB[4], rootBoxB[5]};
    if (!fbox_overlap(a, b))
        return;
    if ((rootA & kLeafBit) && (rootB & kLeafBit)) {
        const uint32_t ta = rootA & ~kLeafBit, tb = rootB & ~kLeafBit;
        if (box64_overlap(record_box(t.tri64A, ta), record_box(t.tri64B, tb))) {
            cand[0] = ((uint64_t)ta << 32) | tb;
            *candCount = 1;
        }
        return;
    }
    frontier[0] = make_uint2(rootA, rootB);
    *count = 1;
}

void launch_expand(const TraverseArgs &t, const uint2 *in, const uint32_t *inCount, uint2 *out,
                   uint32_t *outCount, uint32_t outCap, uint64_t *cand, uint32_t *candCount,
                   uint32_t candCap, uint32_t *overflow, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_expand, dim3(grid * 256 / kExpandThreads), dim3(kExpandThreads), 0, s, t, in, inCount,
                       out, outCount, outCap, cand, candCount, candCap, overflow);
}

void launch_seed(const TraverseArgs &t, uint32_t rootA, uint32_t rootB, const float *rootBoxA,
                 const float *rootBoxB, uint2 *frontier, uint32_t *count, uint64_t *cand,
                 uint32_t *candCount, hipStream_t s)
{
    hipLaunchKernelGGL(k_seed, dim3(1), dim3(64), 0, s, t, rootA, rootB, rootBoxA, rootBoxB, frontier, count,
                       cand, candCount);
}

}  // namespace sbg
