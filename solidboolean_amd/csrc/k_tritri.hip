// Batched Guigue-Devillers triangle-triangle test with segment construction
// (reference: thirdparty/GuigueDevillers03/tri_tri_intersect.c:285-472, called
// per candidate pair by src/solidboolean.cpp:103-122 and :315-341).
//
// One lane per candidate pair.  Two precisions, identical results:
//  * FP64: the reference's exact operation sequence (sbg_device.h gd_fp64)
//    on the pair's two 96-B FP64 records.
//  * FP32-filtered, two kernels: k_gd_filter evaluates every orientation
//    predicate the reference would evaluate from the 48-B FP32 records and
//    certifies its sign against a per-pair forward error bound.  A pair whose
//    certified path ends in a rejection is decided there; every other pair (a
//    certified intersection, or any uncertain sign) is written to its wave
//    chunk's slice of a list that k_gd_list re-runs with the full FP64
//    sequence, so every status and every segment equals the FP64 reference's.
//    Splitting the two keeps waves from running both paths for the few lanes
//    that hit; per-chunk slices keep the list free of shared atomics.
// Per pair and launch: 8 B pair + 1 B status + 2 x 96 B (FP64) or 2 x 48 B
// (filtered, + 2 x 96 B + 4 B list entry for hits / fallbacks) + 48 B segment
// per hit.
#include "sbg_device.h"
#include "sbg_internal.h"

#pragma clang fp contract(off)

namespace sbg {

__device__ __forceinline__ void load_tri64(const double *__restrict__ rec, uint32_t t, V3 &p, V3 &q, V3 &r)
{
    const double2 *x = reinterpret_cast<const double2 *>(rec + (size_t)kTri64Stride * t);
    const double2 a = x[0], b = x[1], c = x[2], d = x[3], e = x[4], f = x[5];
    p = {a.x, a.y, b.x};
    q = {c.x, c.y, d.x};
    r = {e.x, e.y, f.x};
}

// FP32 record: vertices + (largest box extent, largest |coordinate|) upper bounds
__device__ __forceinline__ void load_tri32(const float *__restrict__ rec, uint32_t t, F3 &p, F3 &q, F3 &r,
                                           float &ext, float &mag)
{
    const float4 *x = reinterpret_cast<const float4 *>(rec + (size_t)kTri32Stride * t);
    const float4 a = x[0], b = x[1], c = x[2];
    p = {a.x, a.y, a.z};
    q = {b.x, b.y, b.z};
    r = {c.x, c.y, c.z};
    ext = a.w;
    mag = b.w;
}

// Per-pair bound on |d32 - d64| for every predicate d = DOT(a, CROSS(b, c)) of
// the pair.  Every vertex of A lies within extA of p1 and every vertex of B
// within extB of p2 on each axis (extents of the FP64 boxes, rounded up), and
// FP32 rounding moves each coordinate by at most uM, so every coordinate
// difference among the six vertices, as computed in FP32, is at most
//   D = (max_axis |p1 - p2| + extA + extB + 2 eta)(1 + 4u),
// with eta = 2uM(1 + 2u) >= the input-rounding error of a difference and M
// the larger |coordinate|.  This holds for any pair, overlapping or not.  With
// P <= 6D^3, Q <= 18D^2 and S1 <= 9D in the per-predicate bound derived in
// sbg_device.h,
//   |d32 - d64| <= 27 eta D^2 + 45 eta^2 D + 50 eta^3 + 72 u D^3,
// taken here with a further 1e-4 relative margin for the rounding of this
// very expression and 1e-36 absolute for FP32 gradual underflow.
__device__ __forceinline__ float pair_bound(const F3 &p1, const F3 &p2, float extA, float extB, float magA,
                                            float magB)
{
    const float u = 5.9604644775390625e-8f;  // 2^-24
    const float eta = 2.0f * u * fmaxf(magA, magB) * (1.0f + 2.0f * u);
    const float gap = fmaxf(fmaxf(__builtin_fabsf(p1.x - p2.x), __builtin_fabsf(p1.y - p2.y)),
                            __builtin_fabsf(p1.z - p2.z));
    const float D = (gap + extA + extB + 2.0f * eta) * (1.0f + 4.0f * u);
    const float E = 27.0f * eta * D * D + 45.0f * eta * eta * D + 50.0f * eta * eta * eta + 72.0f * u * D * D * D;
    return E * 1.0001f + 1e-36f;
}

// certified sign of DOT(a, n) with n = CROSS(b, c) precomputed, or 0.  The
// 2^-99 floor keeps |d64| >= 2^-100, so the reference's FP64 products
// dp1 * dq1 (tri_tri_intersect.c:421,438) cannot underflow to zero.
__device__ __forceinline__ int fsign_n(const F3 &a, const F3 &n, float E)
{
    const float d = a.x * n.x + a.y * n.y + a.z * n.z;
    if (__builtin_fabsf(d) > E + 1.6e-30f)
        return d > 0.0f ? 1 : -1;
    return 0;
}

__device__ __forceinline__ F3 fcross(const F3 &b, const F3 &c)
{
    return {b.y * c.z - b.z * c.y, b.z * c.x - b.x * c.z, b.x * c.y - b.y * c.x};
}

__device__ __forceinline__ int fsign(const F3 &a, const F3 &b, const F3 &c, float E)
{
    return fsign_n(a, fcross(b, c), E);
}

// CONSTRUCT_INTERSECTION's predicates (tri_tri_intersect.c:286-296, 324-332):
// 1 = intersection (segment computed later in FP64), 0 = rejection, -2 = uncertain
__device__ __forceinline__ int f_construct(const F3 &p1, const F3 &q1, const F3 &r1, const F3 &p2, const F3 &q2,
                                           const F3 &r2, float E)
{
    const F3 v = fsub(p2, p1);
    F3 v1 = fsub(q1, p1);
    F3 v2 = fsub(r2, p1);
    int s = fsign(v, v1, v2, E);
    if (!s)
        return -2;
    if (s > 0) {
        v1 = fsub(r1, p1);
        s = fsign(v, v1, v2, E);
        if (!s)
            return -2;
        return s < 0 ? 1 : 0;
    }
    v2 = fsub(q2, p1);
    s = fsign(v, v1, v2, E);
    if (!s)
        return -2;
    return s < 0 ? 0 : 1;
}

// TRI_TRI_INTER_3D (tri_tri_intersect.c:360-385) for non-zero signs
__device__ __forceinline__ int f_inter3d(const F3 &p1, const F3 &q1, const F3 &r1, const F3 &p2, const F3 &q2,
                                         const F3 &r2, int sp2, int sq2, int sr2, float E)
{
    if (sp2 > 0) {
        if (sq2 > 0)
            return f_construct(p1, r1, q1, r2, p2, q2, E);
        if (sr2 > 0)
            return f_construct(p1, r1, q1, q2, r2, p2, E);
        return f_construct(p1, q1, r1, p2, q2, r2, E);
    }
    if (sq2 < 0)
        return f_construct(p1, q1, r1, r2, p2, q2, E);
    if (sr2 < 0)
        return f_construct(p1, q1, r1, q2, r2, p2, E);
    return f_construct(p1, r1, q1, p2, q2, r2, E);
}

// Returns a certified GdExit rejection (0, 1, 2), GD_HIT when the certified path
// ends in an intersection, or -1 when some sign is uncertain.
__device__ __forceinline__ int gd_filter(const F3 &p1, const F3 &q1, const F3 &r1, const F3 &p2, const F3 &q2,
                                         const F3 &r2, float E)
{
    const F3 N2 = fcross(fsub(p2, r2), fsub(q2, r2));
    const int sp1 = fsign_n(fsub(p1, r2), N2, E);
    const int sq1 = fsign_n(fsub(q1, r2), N2, E);
    const int sr1 = fsign_n(fsub(r1, r2), N2, E);
    if (!sp1 || !sq1 || !sr1)
        return -1;
    if (sp1 == sq1 && sp1 == sr1)
        return GD_REJECT_PLANE1;
    const F3 N1 = fcross(fsub(q1, p1), fsub(r1, p1));
    const int sp2 = fsign_n(fsub(p2, r1), N1, E);
    const int sq2 = fsign_n(fsub(q2, r1), N1, E);
    const int sr2 = fsign_n(fsub(r2, r1), N1, E);
    if (!sp2 || !sq2 || !sr2)
        return -1;
    if (sp2 == sq2 && sp2 == sr2)
        return GD_REJECT_PLANE2;
    int r;
    if (sp1 > 0) {
        if (sq1 > 0)
            r = f_inter3d(r1, p1, q1, p2, r2, q2, sp2, sr2, sq2, E);
        else if (sr1 > 0)
            r = f_inter3d(q1, r1, p1, p2, r2, q2, sp2, sr2, sq2, E);
        else
            r = f_inter3d(p1, q1, r1, p2, q2, r2, sp2, sq2, sr2, E);
    } else {
        if (sq1 < 0)
            r = f_inter3d(r1, p1, q1, p2, q2, r2, sp2, sq2, sr2, E);
        else if (sr1 < 0)
            r = f_inter3d(q1, r1, p1, p2, q2, r2, sp2, sq2, sr2, E);
        else
            r = f_inter3d(p1, q1, r1, p2, r2, q2, sp2, sr2, sq2, E);
    }
    if (r == -2)
        return -1;
    return r ? GD_HIT : GD_REJECT_INTERVAL;
}

// ---- wave chunks ------------------------------------------------------------
// Pairs are processed in wave chunks of 64 * items consecutive pairs, one wave
// per chunk.  The filter writes the chunk's undecided pairs into the chunk's
// own slice of the list (list[base, base + count)), so no queue counter is
// shared between waves and nothing needs an atomic; k_gd_list then gives each
// chunk's slice to one wave.  Statistics are kept as wave-uniform counts and
// written once per chunk (8 u32: five exits, fallback, filtered, spare) for a
// one-block reduction.
constexpr int kStatSlots = 8;

__device__ __forceinline__ uint32_t wave_popc(bool pred) { return (uint32_t)__popcll(__ballot(pred)); }

// ex[k] for k < 5, else 0, without indexing a register array by a lane value
__device__ __forceinline__ uint32_t pick5(const uint32_t ex[5], int k)
{
    return k == 0 ? ex[0] : k == 1 ? ex[1] : k == 2 ? ex[2] : k == 3 ? ex[3] : k == 4 ? ex[4] : 0u;
}

__device__ __forceinline__ void gd_pair_fp64(const double *__restrict__ tri64A, const double *__restrict__ tri64B,
                                             uint64_t pr, uint32_t i, uint8_t *__restrict__ status,
                                             double *__restrict__ seg, int &exitKind)
{
    const uint32_t a = (uint32_t)(pr >> 32), b = (uint32_t)pr;
    V3 p1, q1, r1, p2, q2, r2, s, t;
    load_tri64(tri64A, a, p1, q1, r1);
    load_tri64(tri64B, b, p2, q2, r2);
    exitKind = gd_fp64(p1, q1, r1, p2, q2, r2, s, t);
    if (exitKind == GD_HIT) {
        double2 *o = reinterpret_cast<double2 *>(seg + 6 * (size_t)i);
        o[0] = make_double2(s.x, s.y);
        o[1] = make_double2(s.z, t.x);
        o[2] = make_double2(t.y, t.z);
    }
    status[i] = exitKind == GD_HIT ? 1 : (exitKind == GD_COPLANAR ? 2 : 0);
}

// plain FP64 over every pair
__global__ __launch_bounds__(256) void k_gd_fp64(const double *__restrict__ tri64A, const double *__restrict__ tri64B,
                                                 const uint64_t *__restrict__ pairs, uint32_t cap,
                                                 const uint32_t *__restrict__ count, uint32_t items,
                                                 uint8_t *__restrict__ status, double *__restrict__ seg,
                                                 uint32_t *__restrict__ partial)
{
    const int lane = threadIdx.x & 63;
    const uint32_t w = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    const uint64_t base = (uint64_t)w * 64u * items;
    if (base >= cap)
        return;
    const uint32_t n = count && *count < cap ? *count : cap;
    uint32_t ex[5] = {0, 0, 0, 0, 0};
    for (uint32_t it = 0; it < items && base + 64u * it < n; ++it) {
        const uint64_t i = base + 64u * it + lane;
        const bool valid = i < n;
        int exitKind = -1;
        if (valid)
            gd_pair_fp64(tri64A, tri64B, pairs[i], (uint32_t)i, status, seg, exitKind);
        if (partial) {
#pragma unroll
            for (int k = 0; k < 5; ++k)
                ex[k] += wave_popc(valid && exitKind == k);
        }
    }
    if (partial && lane < kStatSlots)
        partial[(size_t)w * kStatSlots + lane] = pick5(ex, lane);
}

// FP32 certificates; each chunk's undecided pairs go to its slice of `list`
__global__ __launch_bounds__(256) void k_gd_filter(const float *__restrict__ tri32A, const float *__restrict__ tri32B,
                                                   const uint64_t *__restrict__ pairs, uint32_t cap,
                                                   const uint32_t *__restrict__ count, uint32_t items,
                                                   uint8_t *__restrict__ status, uint32_t *__restrict__ list,
                                                   uint32_t *__restrict__ chunkCount, uint32_t *__restrict__ partial)
{
    const int lane = threadIdx.x & 63;
    const uint32_t w = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    const uint64_t base = (uint64_t)w * 64u * items;
    if (base >= cap)
        return;
    // with a device-side count, chunks past it still clear their slice count
    const uint32_t n = count && *count < cap ? *count : cap;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t cnt = 0, e0 = 0, e1 = 0, e2 = 0, fb = 0, ft = 0;
    for (uint32_t it = 0; it < items && base + 64u * it < n; ++it) {
        const uint64_t i = base + 64u * it + lane;
        const bool valid = i < n;
        int f = -3;
        if (valid) {
            const uint64_t pr = pairs[i];
            F3 p1, q1, r1, p2, q2, r2;
            float extA, magA, extB, magB;
            load_tri32(tri32A, (uint32_t)(pr >> 32), p1, q1, r1, extA, magA);
            load_tri32(tri32B, (uint32_t)pr, p2, q2, r2, extB, magB);
            f = gd_filter(p1, q1, r1, p2, q2, r2, pair_bound(p1, p2, extA, extB, magA, magB));
            if (f >= 0 && f <= GD_REJECT_INTERVAL)
                status[i] = 0;
        }
        const bool pass = valid && (f < 0 || f == GD_HIT);
        const unsigned long long m = __ballot(pass);
        if (pass)
            list[base + cnt + (uint32_t)__popcll(m & below)] = (uint32_t)i;
        cnt += (uint32_t)__popcll(m);
        if (partial) {
            e0 += wave_popc(valid && f == 0);
            e1 += wave_popc(valid && f == 1);
            e2 += wave_popc(valid && f == 2);
            fb += wave_popc(valid && f < 0);
            ft += wave_popc(valid && f >= 0 && f <= GD_REJECT_INTERVAL);
        }
    }
    if (lane == 0)
        chunkCount[w] = cnt;
    if (partial && lane < kStatSlots) {
        const uint32_t v = lane == 0 ? e0 : lane == 1 ? e1 : lane == 2 ? e2 : lane == 5 ? fb : lane == 6 ? ft : 0u;
        partial[(size_t)w * kStatSlots + lane] = v;
    }
}

// Small launches (one wave chunk of 64 pairs per wave): the FP32 certificates
// and, in the same lane, the FP64 sequence for a pair they leave undecided or
// certify as a hit.  At a few tens of thousands of pairs the two-kernel split
// costs a second launch's latency for nothing; results are the same.
__global__ __launch_bounds__(256) void k_gd_fused(const float *__restrict__ tri32A, const float *__restrict__ tri32B,
                                                  const double *__restrict__ tri64A,
                                                  const double *__restrict__ tri64B,
                                                  const uint64_t *__restrict__ pairs, uint32_t n,
                                                  uint8_t *__restrict__ status, double *__restrict__ seg,
                                                  uint32_t *__restrict__ partial)
{
    const int lane = threadIdx.x & 63;
    const uint32_t w = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint64_t i = (uint64_t)w * 64u + lane;
    if ((uint64_t)w * 64u >= n)
        return;
    const bool valid = i < n;
    int f = -3, exitKind = -1;
    if (valid) {
        const uint64_t pr = pairs[i];
        F3 p1, q1, r1, p2, q2, r2;
        float extA, magA, extB, magB;
        load_tri32(tri32A, (uint32_t)(pr >> 32), p1, q1, r1, extA, magA);
        load_tri32(tri32B, (uint32_t)pr, p2, q2, r2, extB, magB);
        f = gd_filter(p1, q1, r1, p2, q2, r2, pair_bound(p1, p2, extA, extB, magA, magB));
        if (f >= 0 && f <= GD_REJECT_INTERVAL) {
            status[i] = 0;
            exitKind = f;
        } else {
            gd_pair_fp64(tri64A, tri64B, pr, (uint32_t)i, status, seg, exitKind);
        }
    }
    if (partial) {
        uint32_t ex[5];
#pragma unroll
        for (int k = 0; k < 5; ++k)
            ex[k] = wave_popc(valid && exitKind == k);
        const uint32_t fb = wave_popc(valid && f < 0);
        const uint32_t ft = wave_popc(valid && f >= 0 && f <= GD_REJECT_INTERVAL);
        if (lane < kStatSlots)
            partial[(size_t)w * kStatSlots + lane] = lane < 5 ? pick5(ex, lane) : lane == 5 ? fb : lane == 6 ? ft : 0u;
    }
}

// FP64 on each chunk's undecided pairs; adds its exits to the chunk's counts
__global__ __launch_bounds__(256) void k_gd_list(const double *__restrict__ tri64A, const double *__restrict__ tri64B,
                                                 const uint64_t *__restrict__ pairs, uint32_t nChunks, uint32_t items,
                                                 const uint32_t *__restrict__ list,
                                                 const uint32_t *__restrict__ chunkCount,
                                                 uint8_t *__restrict__ status, double *__restrict__ seg,
                                                 uint32_t *__restrict__ partial)
{
    const int lane = threadIdx.x & 63;
    const uint32_t w = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (w >= nChunks)
        return;
    const uint64_t base = (uint64_t)w * 64u * items;
    const uint32_t m = chunkCount[w];
    uint32_t ex[5] = {0, 0, 0, 0, 0};
    for (uint32_t k0 = 0; k0 < m; k0 += 64) {
        const uint32_t k = k0 + lane;
        const bool valid = k < m;
        int exitKind = -1;
        if (valid) {
            const uint32_t i = list[base + k];
            gd_pair_fp64(tri64A, tri64B, pairs[i], i, status, seg, exitKind);
        }
        if (partial) {
#pragma unroll
            for (int e = 0; e < 5; ++e)
                ex[e] += wave_popc(valid && exitKind == e);
        }
    }
    if (partial && lane < 5)
        partial[(size_t)w * kStatSlots + lane] += pick5(ex, lane);
}

// one block: sums the per-chunk counts into the launch's statistics
__global__ __launch_bounds__(256) void k_gd_stats(const uint32_t *__restrict__ partial, uint32_t nChunks,
                                                  GdStats *__restrict__ stats)
{
    __shared__ unsigned long long s_sum[256];
    unsigned long long acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t w = threadIdx.x; w < nChunks; w += 256) {
        const uint32_t *p = partial + (size_t)w * kStatSlots;
#pragma unroll
        for (int k = 0; k < 7; ++k)
            acc[k] += p[k];
    }
    for (int k = 0; k < 7; ++k) {
        s_sum[threadIdx.x] = acc[k];
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h)
                s_sum[threadIdx.x] += s_sum[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (k < 5)
                stats->exits[k] += s_sum[0];
            else if (k == 5)
                stats->fallback += s_sum[0];
            else
                stats->filtered += s_sum[0];
        }
        __syncthreads();
    }
}

// host triangle arrays: a9/b9 hold 9 doubles per pair (p, q, r)
__global__ __launch_bounds__(256) void k_tritri_raw(const double *__restrict__ a9, const double *__restrict__ b9,
                                                    uint32_t n, uint8_t *__restrict__ status,
                                                    double *__restrict__ seg)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const double *a = a9 + 9 * (size_t)i;
    const double *b = b9 + 9 * (size_t)i;
    V3 p1 = {a[0], a[1], a[2]}, q1 = {a[3], a[4], a[5]}, r1 = {a[6], a[7], a[8]};
    V3 p2 = {b[0], b[1], b[2]}, q2 = {b[3], b[4], b[5]}, r2 = {b[6], b[7], b[8]};
    V3 s = {0, 0, 0}, t = {0, 0, 0};
    const int e = gd_fp64(p1, q1, r1, p2, q2, r2, s, t);
    status[i] = e == GD_HIT ? 1 : (e == GD_COPLANAR ? 2 : 0);
    double *o = seg + 6 * (size_t)i;
    o[0] = s.x;
    o[1] = s.y;
    o[2] = s.z;
    o[3] = t.x;
    o[4] = t.y;
    o[5] = t.z;
}

// ---- hit compaction (pair order preserved) --------------------------------
constexpr int kCompactBlock = 256;

__global__ __launch_bounds__(kCompactBlock) void k_hit_count(const uint8_t *__restrict__ status, uint32_t n,
                                                             uint32_t *__restrict__ blockCounts)
{
    const uint32_t i = blockIdx.x * kCompactBlock + threadIdx.x;
    const bool hit = i < n && status[i] == 1;
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0)
        s_cnt = 0;
    __syncthreads();
    const unsigned long long m = __ballot(hit);
    if ((threadIdx.x & 63) == 0)
        atomicAdd(&s_cnt, (uint32_t)__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0)
        blockCounts[blockIdx.x] = s_cnt;
}

__global__ __launch_bounds__(kCompactBlock) void k_hit_write(const uint64_t *__restrict__ pairs,
                                                             const uint8_t *__restrict__ status,
                                                             const double *__restrict__ seg, uint32_t n,
                                                             const uint32_t *__restrict__ blockOffsets,
                                                             uint64_t *__restrict__ hitPairs,
                                                             double *__restrict__ hitSeg, uint32_t *nHits)
{
    const uint32_t i = blockIdx.x * kCompactBlock + threadIdx.x;
    const bool hit = i < n && status[i] == 1;
    __shared__ uint32_t s_wave[kCompactBlock / 64];
    const unsigned long long m = __ballot(hit);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        s_wave[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = blockOffsets[blockIdx.x];
    for (int w = 0; w < wave; ++w)
        off += s_wave[w];
    off += (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (hit) {
        hitPairs[off] = pairs[i];
        for (int k = 0; k < 6; ++k)
            hitSeg[6 * (size_t)off + k] = seg[6 * (size_t)i + k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        uint32_t tot = blockOffsets[blockIdx.x];
        for (int w = 0; w < kCompactBlock / 64; ++w)
            tot += s_wave[w];
        *nHits = tot;
    }
}

int compact_blocks(uint32_t n) { return (int)((n + kCompactBlock - 1) / kCompactBlock); }

void launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s);

// up to this many pairs the filtered precision runs as one fused launch
constexpr uint32_t kFusedMaxPairs = 1u << 18;

// pairs per wave chunk / 64: one chunk per wave until about 8192 waves are
// available, then up to 16 pairs per lane
static uint32_t gd_items(uint32_t n)
{
    const uint32_t per = (uint32_t)((n + 64ull * 8192u - 1) / (64ull * 8192u));
    return per < 1u ? 1u : (per > 16u ? 16u : per);
}

static uint32_t gd_chunks(uint32_t n, uint32_t items) { return (uint32_t)((n + 64ull * items - 1) / (64ull * items)); }

size_t tritri_work_u32(uint32_t n)
{
    const uint32_t chunks = gd_chunks(n, gd_items(n));
    return (size_t)n + (size_t)chunks * (1 + kStatSlots) + 4;
}

// work: tritri_work_u32(n) u32 (list slices, chunk counts, per-chunk statistics);
// count (may be null): device-side number of pairs, at most n, read at run time
void launch_tritri(const double *tri64A, const double *tri64B, const float *tri32A, const float *tri32B,
                   const uint64_t *pairs, uint32_t n, const uint32_t *count, int precision, uint8_t *status,
                   double *seg, uint32_t *work, GdStats *stats, hipStream_t s)
{
    if (n == 0)
        return;
    const uint32_t items = gd_items(n);
    const uint32_t chunks = gd_chunks(n, items);
    const dim3 grid(xcd_grid((chunks + 3) / 4));
    uint32_t *list = work;
    uint32_t *chunkCount = work + n;
    uint32_t *partial = stats ? chunkCount + chunks : nullptr;
    if (precision == 1 && !count && items == 1 && n <= kFusedMaxPairs) {
        hipLaunchKernelGGL(k_gd_fused, grid, dim3(256), 0, s, tri32A, tri32B, tri64A, tri64B, pairs, n, status, seg,
                           partial);
    } else if (precision == 1) {
        hipLaunchKernelGGL(k_gd_filter, grid, dim3(256), 0, s, tri32A, tri32B, pairs, n, count, items, status, list,
                           chunkCount, partial);
        hipLaunchKernelGGL(k_gd_list, grid, dim3(256), 0, s, tri64A, tri64B, pairs, chunks, items, list, chunkCount,
                           status, seg, partial);
    } else {
        hipLaunchKernelGGL(k_gd_fp64, grid, dim3(256), 0, s, tri64A, tri64B, pairs, n, count, items, status, seg,
                           partial);
    }
    if (stats)
        hipLaunchKernelGGL(k_gd_stats, dim3(1), dim3(256), 0, s, partial, chunks, stats);
}

void launch_tritri_raw(const double *a9, const double *b9, uint32_t n, uint8_t *status, double *seg,
                       hipStream_t s)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_tritri_raw, dim3((n + 255) / 256), dim3(256), 0, s, a9, b9, n, status, seg);
}

void launch_exclusive_scan(uint32_t *data, size_t m, hipStream_t s);

void launch_compact_hits(const uint64_t *pairs, const uint8_t *status, const double *seg, uint32_t n,
                         uint32_t *blockCounts, uint64_t *hitPairs, double *hitSeg, uint32_t *nHits,
                         hipStream_t s)
{
    if (n == 0)
        return;
    const int nb = compact_blocks(n);
    hipLaunchKernelGGL(k_hit_count, dim3(nb), dim3(kCompactBlock), 0, s, status, n, blockCounts);
    launch_exclusive_scan(blockCounts, (size_t)nb, s);
    hipLaunchKernelGGL(k_hit_write, dim3(nb), dim3(kCompactBlock), 0, s, pairs, status, seg, n, blockCounts,
                       hitPairs, hitSeg, nHits);
}

}  // namespace sbg
