// Stable LSD radix sort, 8-bit digits, for u32 keys (+ u32 values) and u64 keys.
//
// Per pass: (1) every block counts the digits of its tile (256 x ITEMS keys)
// into LDS and writes a 256 x nblocks histogram column; (2) one block per digit
// turns the histogram (digit-major) into exclusive offsets; (3) every block
// loads its whole tile into registers, then ranks it in order, 256 keys per
// round, each key among equal digits with a 64-lane ballot match (stable within
// the wave), adds the counts of earlier waves and earlier rounds kept in LDS,
// and scatters.  ITEMS shrinks with n so that short sorts still spread over
// many workgroups.  Per key and pass: read 4 (8) B key (+4 B value) twice,
// write once.
#include "sbg_internal.h"

namespace sbg {

constexpr int kSortBlock = 256;
constexpr int kWaves = kSortBlock / 64;

// keys per thread for a sort of n keys
static int sort_items(size_t n) { return n >= ((size_t)1 << 20) ? 16 : (n >= ((size_t)1 << 17) ? 8 : 4); }

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, int shift, uint32_t mask)
{
    return (uint32_t)(key >> shift) & mask;
}

// count (may be null): the keys are the first min(n, *count) of the n the
// launch covers (a list whose length is known on the device only); the rest
// are neither read nor moved
__device__ __forceinline__ size_t sort_len(size_t n, const uint32_t *count)
{
    return count && *count < n ? *count : n;
}

template <typename K, int ITEMS>
__global__ __launch_bounds__(kSortBlock) void k_hist(const K *__restrict__ keys, size_t n,
                                                     const uint32_t *__restrict__ count, int shift, uint32_t mask,
                                                     uint32_t *__restrict__ hist, int nblocks)
{
    n = sort_len(n, count);
    __shared__ uint32_t s_cnt[256];
    s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * (kSortBlock * ITEMS);
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        size_t idx = base + (size_t)r * kSortBlock + threadIdx.x;
        if (idx < n)
            atomicAdd(&s_cnt[digit_of(keys[idx], shift, mask)], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = s_cnt[threadIdx.x];
}

// exclusive scan of m entries in place, one block of 1024 threads
__global__ __launch_bounds__(1024) void k_scan(uint32_t *__restrict__ data, size_t m)
{
    __shared__ uint32_t s_part[1024];
    const size_t per = (m + 1023) / 1024;
    const size_t lo = (size_t)threadIdx.x * per;
    const size_t hi = lo + per < m ? lo + per : m;
    uint32_t sum = 0;
    for (size_t i = lo; i < hi; ++i)
        sum += data[i];
    s_part[threadIdx.x] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan of the 1024 partials
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t v = threadIdx.x >= (unsigned)off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? s_part[threadIdx.x - 1] : 0;
    for (size_t i = lo; i < hi; ++i) {
        uint32_t v = data[i];
        data[i] = run;
        run += v;
    }
}

// ---- exclusive scan of a long u32 array, three launches --------------------
// Tiles of 4096 entries (16 per thread): tile sums, a one-block scan of the
// sums, then each tile rescans its entries from its base.
constexpr int kScanTile = 4096;

__global__ __launch_bounds__(256) void k_scan_tile_sum(const uint32_t *__restrict__ data, size_t m,
                                                       uint32_t *__restrict__ tileSum)
{
    __shared__ uint32_t s[256];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t sum = 0;
    for (int k = 0; k < kScanTile / 256; ++k) {
        const size_t i = base + (size_t)k * 256 + threadIdx.x;
        if (i < m)
            sum += data[i];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        tileSum[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_scan_tile_apply(uint32_t *__restrict__ data, size_t m,
                                                         const uint32_t *__restrict__ tileBase)
{
    __shared__ uint32_t s[256];
    constexpr int per = kScanTile / 256;
    const size_t lo = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * per;
    uint32_t v[per];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        v[k] = lo + k < m ? data[lo + k] : 0u;
        sum += v[k];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = tileBase[blockIdx.x] + (threadIdx.x ? s[threadIdx.x - 1] : 0u);
#pragma unroll
    for (int k = 0; k < per; ++k) {
        if (lo + k < m)
            data[lo + k] = run;
        run += v[k];
    }
}

size_t scan_large_temp(size_t m) { return (m + kScanTile - 1) / kScanTile + 1; }

void launch_scan_large(uint32_t *data, size_t m, uint32_t *temp, hipStream_t s)
{
    if (m == 0)
        return;
    const size_t tiles = (m + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tile_sum, dim3((unsigned)tiles), dim3(256), 0, s, data, m, temp);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, temp, tiles);
    hipLaunchKernelGGL(k_scan_tile_apply, dim3((unsigned)tiles), dim3(256), 0, s, data, m, temp);
}

// exclusive scan of each digit's row of the histogram (one workgroup per
// digit); the row's total goes to rowTotal[digit]
__global__ __launch_bounds__(256) void k_scan_rows(uint32_t *__restrict__ hist, int nblocks,
                                                   uint32_t *__restrict__ rowTotal)
{
    __shared__ uint32_t s[256];
    uint32_t *row = hist + (size_t)blockIdx.x * nblocks;
    const int per = (nblocks + 255) / 256;
    const int lo = threadIdx.x * per;
    const int hi = lo + per < nblocks ? lo + per : nblocks;
    uint32_t sum = 0;
    for (int i = lo; i < hi; ++i)
        sum += row[i];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t v = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? s[threadIdx.x - 1] : 0;
    for (int i = lo; i < hi; ++i) {
        const uint32_t v = row[i];
        row[i] = run;
        run += v;
    }
    if (threadIdx.x == 255)
        rowTotal[blockIdx.x] = s[255];
}

template <typename K, bool HAS_VAL, int ITEMS>
__global__ __launch_bounds__(kSortBlock) void k_scatter(const K *__restrict__ keys, const uint32_t *__restrict__ vals,
                                                        K *__restrict__ keysOut, uint32_t *__restrict__ valsOut,
                                                        size_t n, const uint32_t *__restrict__ count, int shift,
                                                        uint32_t mask, const uint32_t *__restrict__ offsets,
                                                        int nblocks, const uint32_t *__restrict__ rowTotal)
{
    n = sort_len(n, count);
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_wcnt[kWaves][256];
    // the tile first, so that its loads overlap the offset set-up
    const size_t base = (size_t)blockIdx.x * (kSortBlock * ITEMS);
    K kreg[ITEMS];
    uint32_t vreg[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const size_t idx = base + (size_t)r * kSortBlock + threadIdx.x;
        kreg[r] = idx < n ? keys[idx] : (K)0;
        vreg[r] = (HAS_VAL && idx < n) ? vals[idx] : 0u;
    }
    // digit bases: exclusive scan of the 256 row totals
    const uint32_t total = rowTotal[threadIdx.x];
    s_base[threadIdx.x] = total;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t v = threadIdx.x >= (unsigned)off ? s_base[threadIdx.x - off] : 0;
        __syncthreads();
        s_base[threadIdx.x] += v;
        __syncthreads();
    }
    const uint32_t digitBase = s_base[threadIdx.x] - total;
    __syncthreads();
    s_base[threadIdx.x] = digitBase + offsets[(size_t)threadIdx.x * nblocks + blockIdx.x];
    for (int w = 0; w < kWaves; ++w)
        s_wcnt[w][threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const unsigned long long ltmask = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const size_t idx = base + (size_t)r * kSortBlock + threadIdx.x;
        const bool valid = idx < n;
        const K key = kreg[r];
        const uint32_t val = vreg[r];
        const uint32_t d = digit_of(key, shift, mask);
        // lanes of this wave holding the same digit
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t rank = __popcll(peers & ltmask);
        if (valid && rank == 0)
            s_wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t pos = 0;
        if (valid) {
            pos = s_base[d] + rank;
            for (int w = 0; w < wave; ++w)
                pos += s_wcnt[w][d];
        }
        __syncthreads();
        {
            uint32_t tot = 0;
            for (int w = 0; w < kWaves; ++w) {
                tot += s_wcnt[w][threadIdx.x];
                s_wcnt[w][threadIdx.x] = 0;
            }
            s_base[threadIdx.x] += tot;
        }
        __syncthreads();
        if (valid) {
            keysOut[pos] = key;
            if (HAS_VAL)
                valsOut[pos] = val;
        }
    }
}

// A list of (first << 32 | second) keys already sorted by their high word:
// each run of equal high words is sorted by the low word in place, one lane
// per run (the lane at the run's first entry), by insertion.  A run longer
// than maxRun is left as it is and raises *flag (the caller then sorts the
// whole list by both words).  Runs of ~0 keys (padding) are skipped.  Lanes
// only ever move entries within their own run, so the high word another lane
// reads at a run's border never changes under it.
__global__ __launch_bounds__(256) void k_sort_runs(uint64_t *keys, uint32_t n, const uint32_t *__restrict__ count,
                                                   uint32_t maxRun, uint32_t *flag)
{
    n = (uint32_t)sort_len(n, count);
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const uint64_t k = keys[i];
    const uint32_t hi = (uint32_t)(k >> 32);
    if (k == ~0ull || (i > 0 && (uint32_t)(keys[i - 1] >> 32) == hi))
        return;
    uint32_t j = i + 1;
    while (j < n && (uint32_t)(keys[j] >> 32) == hi) {
        if (++j - i > maxRun) {
            atomicOr(flag, 1u);
            return;
        }
    }
    for (uint32_t x = i + 1; x < j; ++x) {
        const uint64_t v = keys[x];
        uint32_t y = x;
        for (; y > i && keys[y - 1] > v; --y)
            keys[y] = keys[y - 1];
        keys[y] = v;
    }
}

void launch_sort_runs(uint64_t *keys, size_t n, const uint32_t *count, uint32_t maxRun, uint32_t *flag,
                      hipStream_t s)
{
    if (n)
        hipLaunchKernelGGL(k_sort_runs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, keys, (uint32_t)n, count,
                           maxRun, flag);
}

void launch_exclusive_scan(uint32_t *data, size_t m, hipStream_t s)
{
    if (m == 0)
        return;
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, data, m);
}

int sort_blocks(size_t n)
{
    const size_t tile = (size_t)kSortBlock * sort_items(n);
    return (int)((n + tile - 1) / tile);
}
// 256 x nblocks histogram + 256 row totals
size_t sort_hist_entries(size_t n) { return (size_t)256 * (size_t)(sort_blocks(n) > 0 ? sort_blocks(n) : 1) + 256; }

template <typename K, bool HAS_VAL>
static void one_pass(const K *kin, const uint32_t *vin, K *kout, uint32_t *vout, size_t n, const uint32_t *count,
                     int shift, int bits, uint32_t *hist, hipStream_t s)
{
    const int nb = sort_blocks(n);
    const uint32_t mask = (1u << bits) - 1u;
    uint32_t *rowTotal = hist + (size_t)256 * nb;
    const int items = sort_items(n);
    if (items == 16)
        hipLaunchKernelGGL((k_hist<K, 16>), dim3(nb), dim3(kSortBlock), 0, s, kin, n, count, shift, mask, hist, nb);
    else if (items == 8)
        hipLaunchKernelGGL((k_hist<K, 8>), dim3(nb), dim3(kSortBlock), 0, s, kin, n, count, shift, mask, hist, nb);
    else
        hipLaunchKernelGGL((k_hist<K, 4>), dim3(nb), dim3(kSortBlock), 0, s, kin, n, count, shift, mask, hist, nb);
    hipLaunchKernelGGL(k_scan_rows, dim3(256), dim3(256), 0, s, hist, nb, rowTotal);
    if (items == 16)
        hipLaunchKernelGGL((k_scatter<K, HAS_VAL, 16>), dim3(nb), dim3(kSortBlock), 0, s, kin, vin, kout, vout, n,
                           count, shift, mask, hist, nb, rowTotal);
    else if (items == 8)
        hipLaunchKernelGGL((k_scatter<K, HAS_VAL, 8>), dim3(nb), dim3(kSortBlock), 0, s, kin, vin, kout, vout, n,
                           count, shift, mask, hist, nb, rowTotal);
    else
        hipLaunchKernelGGL((k_scatter<K, HAS_VAL, 4>), dim3(nb), dim3(kSortBlock), 0, s, kin, vin, kout, vout, n,
                           count, shift, mask, hist, nb, rowTotal);
}

int radix_sort_u32(uint32_t *keys, uint32_t *vals, uint32_t *keysAlt, uint32_t *valsAlt, size_t n,
                   int beginBit, int endBit, uint32_t *hist, hipStream_t s)
{
    if (n <= 1)
        return 0;
    int flip = 0;
    for (int shift = beginBit; shift < endBit; shift += 8) {
        const int bits = endBit - shift < 8 ? endBit - shift : 8;
        uint32_t *ki = flip ? keysAlt : keys, *ko = flip ? keys : keysAlt;
        uint32_t *vi = flip ? valsAlt : vals, *vo = flip ? vals : valsAlt;
        if (vals)
            one_pass<uint32_t, true>(ki, vi, ko, vo, n, nullptr, shift, bits, hist, s);
        else
            one_pass<uint32_t, false>(ki, nullptr, ko, nullptr, n, nullptr, shift, bits, hist, s);
        flip ^= 1;
    }
    return flip;
}

int radix_sort_u64(uint64_t *keys, uint64_t *keysAlt, size_t n, const int *bitRanges, int nRanges,
                   uint32_t *hist, hipStream_t s, const uint32_t *count)
{
    if (n <= 1)
        return 0;
    int flip = 0;
    for (int r = 0; r < nRanges; ++r) {
        const int beginBit = bitRanges[2 * r], endBit = bitRanges[2 * r + 1];
        for (int shift = beginBit; shift < endBit; shift += 8) {
            const int bits = endBit - shift < 8 ? endBit - shift : 8;
            uint64_t *ki = flip ? keysAlt : keys, *ko = flip ? keys : keysAlt;
            one_pass<uint64_t, false>(ki, nullptr, ko, nullptr, n, count, shift, bits, hist, s);
            flip ^= 1;
        }
    }
    return flip;
}

}  // namespace sbg
