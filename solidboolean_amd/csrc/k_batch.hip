// Batched booleans: many independent mesh pairs (A_k, B_k) held as one
// combined mesh (mesh 2k = A_k, mesh 2k + 1 = B_k).  Each of them runs the
// reference's per-pair hot path (src/solidboolean.cpp:94-122 and :482-510);
// here they share every launch:
//  * one LBVH over all triangles whose keys put the mesh id above a Morton code
//    taken within the mesh's own centroid bounds, so every mesh is a subtree;
//  * one candidate search seeded with every pair's two subtree roots;
//  * one Guigue-Devillers launch over all candidates;
//  * one ray-grid build and one ray launch, each mesh with its own grid.
// The kernels below are the batch-specific steps; the rest are the single-pair
// kernels run over the combined arrays.
#include "sbg_device.h"
#include "sbg_internal.h"

namespace sbg {

// order-preserving int image of a float, for atomic min / max
__device__ __forceinline__ int ord_of(float f)
{
    const int b = __float_as_int(f);
    return b >= 0 ? b : b ^ 0x7fffffff;
}
__device__ __forceinline__ float float_of(int o) { return __int_as_float(o >= 0 ? o : o ^ 0x7fffffff); }

// Per-mesh centroid bounds, as ord_of images (min xyz, max xyz); mb holds
// (INT_MAX x 3, INT_MIN x 3) per mesh on entry.  Pass 1: a wave whose
// triangles all belong to one mesh (all but the waves that straddle a mesh
// boundary) stores its reduced bounds in its own slot of `waves`; a straddling
// wave reduces each run of one mesh's lanes (a segmented scan: a mesh's
// triangles are contiguous) and the run's last lane applies 6 atomics.  Pass
// 2: one workgroup per mesh reduces the slots of the waves inside the mesh
// and applies 6 atomics.  No address takes more than a handful of atomics,
// whatever the mesh sizes.
__global__ __launch_bounds__(256) void k_batch_bounds(const float *__restrict__ box32,
                                                      const uint32_t *__restrict__ meshOf, uint32_t nt,
                                                      int *__restrict__ mb, int *__restrict__ waves)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    const bool valid = t < nt;
    if (!__any(valid))
        return;  // whole wave past the end
    const uint32_t m = valid ? meshOf[t] : ~0u;
    int v[6] = {0x7fffffff, 0x7fffffff, 0x7fffffff, (int)0x80000000, (int)0x80000000, (int)0x80000000};
    if (valid) {
        const float3 c = box_centre(box32, t);
        v[0] = v[3] = ord_of(c.x);
        v[1] = v[4] = ord_of(c.y);
        v[2] = v[5] = ord_of(c.z);
    }
    const uint32_t m0 = __shfl(m, 0, 64);  // lane 0 is valid: valid lanes are a prefix of the wave
    if (__all(!valid || m == m0)) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                v[k] = min(v[k], __shfl_xor(v[k], off, 64));
                v[3 + k] = max(v[3 + k], __shfl_xor(v[3 + k], off, 64));
            }
        }
        const int lane = threadIdx.x & 63;
        if (lane < 6)
            waves[6 * (size_t)(t / 64) + lane] = lane == 0 ? v[0] : lane == 1 ? v[1] : lane == 2 ? v[2]
                                                 : lane == 3 ? v[3] : lane == 4 ? v[4] : v[5];
    } else {
        const int lane = threadIdx.x & 63;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t mu = __shfl_up(m, off, 64);
            int u[6];
#pragma unroll
            for (int k = 0; k < 6; ++k)
                u[k] = __shfl_up(v[k], off, 64);
            if (lane >= off && mu == m) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    v[k] = min(v[k], u[k]);
                    v[3 + k] = max(v[3 + k], u[3 + k]);
                }
            }
        }
        const uint32_t mn = __shfl_down(m, 1, 64);
        if (valid && (lane == 63 || mn != m)) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                atomicMin(mb + 6 * (size_t)m + k, v[k]);
                atomicMax(mb + 6 * (size_t)m + 3 + k, v[3 + k]);
            }
        }
    }
}

// one workgroup per mesh over the waves lying wholly inside it
__global__ __launch_bounds__(256) void k_batch_bounds_merge(const int *__restrict__ waves,
                                                            const uint32_t *__restrict__ triBegin, uint32_t nt,
                                                            int *__restrict__ mb)
{
    __shared__ int s[6][256];
    const uint32_t m = blockIdx.x;
    const uint32_t t0 = triBegin[m], t1 = triBegin[m + 1];
    const uint32_t w0 = (t0 + 63) / 64;
    const uint32_t w1 = t1 == nt ? (t1 + 63) / 64 : t1 / 64;  // a last, partial wave is whole
    int v[6] = {0x7fffffff, 0x7fffffff, 0x7fffffff, (int)0x80000000, (int)0x80000000, (int)0x80000000};
    for (uint32_t w = w0 + threadIdx.x; w < w1; w += 256)
        for (int k = 0; k < 3; ++k) {
            v[k] = min(v[k], waves[6 * (size_t)w + k]);
            v[3 + k] = max(v[3 + k], waves[6 * (size_t)w + 3 + k]);
        }
    for (int k = 0; k < 6; ++k)
        s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int k = 0; k < 6; ++k)
                s[k][threadIdx.x] = k < 3 ? min(s[k][threadIdx.x], s[k][threadIdx.x + h])
                                          : max(s[k][threadIdx.x], s[k][threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x < 3 && w0 < w1) {
        atomicMin(mb + 6 * (size_t)m + threadIdx.x, s[threadIdx.x][0]);
        atomicMax(mb + 6 * (size_t)m + 3 + threadIdx.x, s[3 + threadIdx.x][0]);
    }
}

// key = mesh id | Morton code within the mesh's bounds (top mortonBits of 30) |
// index within the mesh (localBits)
__global__ __launch_bounds__(256) void k_batch_keys(const float *__restrict__ box32,
                                                    const uint32_t *__restrict__ meshOf,
                                                    const uint32_t *__restrict__ triBegin, const int *__restrict__ mb,
                                                    uint32_t nt, int localBits, int mortonBits,
                                                    uint64_t *__restrict__ keys)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nt)
        return;
    const uint32_t m = meshOf[t];
    const int *b = mb + 6 * (size_t)m;
    const float bounds[6] = {float_of(b[0]), float_of(b[1]), float_of(b[2]),
                             float_of(b[3]), float_of(b[4]), float_of(b[5])};
    const uint32_t code = morton30(box_centre(box32, t), bounds) >> (30 - mortonBits);
    keys[t] = ((uint64_t)m << (mortonBits + localBits)) | ((uint64_t)code << localBits) | (uint64_t)(t - triBegin[m]);
}

// sorted keys -> hierarchy codes (key without the local index) and triangle ids
__global__ __launch_bounds__(256) void k_batch_codes(const uint64_t *__restrict__ keys, uint32_t n, int localBits,
                                                     int mortonBits, const uint32_t *__restrict__ triBegin,
                                                     uint64_t *__restrict__ codes, uint32_t *__restrict__ ids)
{
    const uint32_t pos = blockIdx.x * 256 + threadIdx.x;
    if (pos >= n)
        return;
    const uint64_t key = keys[pos];
    const uint64_t code = key >> localBits;
    const uint32_t m = (uint32_t)(code >> mortonBits);
    codes[pos] = code;
    ids[pos] = triBegin[m] + (uint32_t)(key & ((1ull << localBits) - 1ull));
}

__device__ __forceinline__ void put_box(float *dst, float4 lo, float4 hi)
{
    dst[0] = lo.x;
    dst[1] = lo.y;
    dst[2] = lo.z;
    dst[3] = hi.x;
    dst[4] = hi.y;
    dst[5] = hi.z;
}

// Subtree root of every mesh and its FP32 box, one lane per mesh.  A mesh's
// sorted positions form one run, [triBegin[m], triBegin[m + 1]) (the mesh id
// leads the key); a one-triangle run is its own root (a leaf), otherwise the
// root is the internal node whose leaf range is exactly the run (the keys of
// the run share the mesh-id prefix, so the radix tree has such a node), and
// an internal node's index is one end of its range, so it is node `first`
// or node `last`.  meshRoot holds ~0u on entry, which stays for empty meshes.
__global__ __launch_bounds__(256) void k_mesh_roots(const BNode *__restrict__ nodes,
                                                    const uint32_t *__restrict__ ids,
                                                    const uint32_t *__restrict__ triBegin, uint32_t nmeshes,
                                                    uint32_t n, const float *__restrict__ box32,
                                                    const float *__restrict__ rootBox, uint32_t *__restrict__ meshRoot,
                                                    float *__restrict__ meshBox)
{
    const uint32_t m = blockIdx.x * 256 + threadIdx.x;
    if (m >= nmeshes)
        return;
    const uint32_t first = triBegin[m], end = triBegin[m + 1];
    if (first == end)
        return;
    const uint32_t last = end - 1;
    float *dst = meshBox + 6 * (size_t)m;
    if (first == last) {
        const uint32_t t = ids[first];
        meshRoot[m] = kLeafBit | t;
        const float4 *bx = reinterpret_cast<const float4 *>(box32 + (size_t)kBox32Stride * t);
        put_box(dst, bx[0], bx[1]);
        return;
    }
    // node `last` exists only below n - 1; a run ending there is node `first`'s
    const uint32_t i = (last >= n - 1 || nodes[first].d.w == (int)last) ? first : last;
    meshRoot[m] = i;
    if (i == 0) {
        for (int k = 0; k < 6; ++k)
            dst[k] = rootBox[k];
        return;
    }
    const BNode *pn = nodes + nodes[i].d.z;
    const float4 a = pn->a, b = pn->b, c = pn->c;
    if (pn->d.x == (int)i)
        put_box(dst, make_float4(a.x, a.y, a.z, 0.0f), make_float4(a.w, b.x, b.y, 0.0f));
    else
        put_box(dst, make_float4(b.z, b.w, c.x, 0.0f), make_float4(c.y, c.z, c.w, 0.0f));
}

__device__ __forceinline__ bool box6_overlap(const float *a, const float *b)
{
    return a[0] <= b[3] && a[3] >= b[0] && a[1] <= b[4] && a[4] >= b[1] && a[2] <= b[5] && a[5] >= b[2];
}

// pair k's root pair: to the frontier, or straight to the candidates when both
// meshes are single triangles whose FP64 boxes overlap
__global__ __launch_bounds__(256) void k_batch_seed(TraverseArgs t, const uint32_t *__restrict__ meshRoot,
                                                    const float *__restrict__ meshBox, uint32_t npairs,
                                                    uint2 *__restrict__ frontier, uint32_t *frontCount,
                                                    uint32_t frontCap, uint64_t *__restrict__ cand,
                                                    uint32_t *candCount, uint32_t candCap, uint32_t *overflow)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= npairs)
        return;
    const uint32_t ra = meshRoot[2 * k], rb = meshRoot[2 * k + 1];
    if (ra == ~0u || rb == ~0u || !box6_overlap(meshBox + 12 * (size_t)k, meshBox + 12 * (size_t)k + 6))
        return;
    if ((ra & kLeafBit) && (rb & kLeafBit)) {
        const uint32_t ta = ra & ~kLeafBit, tb = rb & ~kLeafBit;
        const double *pa = t.tri64A + (size_t)kTri64Stride * ta, *pb = t.tri64B + (size_t)kTri64Stride * tb;
        const Box64 ba = tri_box64({pa[0], pa[1], pa[2]}, {pa[4], pa[5], pa[6]}, {pa[8], pa[9], pa[10]});
        const Box64 bb = tri_box64({pb[0], pb[1], pb[2]}, {pb[4], pb[5], pb[6]}, {pb[8], pb[9], pb[10]});
        if (!box64_overlap(ba, bb))
            return;
        const uint32_t slot = atomicAdd(candCount, 1u);
        if (slot < candCap)
            cand[slot] = ((uint64_t)ta << 32) | tb;
        else
            atomicOr(overflow, 1u);
        return;
    }
    const uint32_t slot = atomicAdd(frontCount, 1u);
    if (slot < frontCap)
        frontier[slot] = make_uint2(ra, rb);
    else
        atomicOr(overflow, 1u);
}

// both triangles of every hit are intersected
__global__ __launch_bounds__(256) void k_mark_both(const uint64_t *__restrict__ pairs,
                                                   const uint8_t *__restrict__ status,
                                                   const uint32_t *__restrict__ count, uint32_t cap,
                                                   uint8_t *__restrict__ flags)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const uint32_t n = *count < cap ? *count : cap;
    if (i >= n || status[i] != 1)
        return;
    const uint64_t pr = pairs[i];
    flags[(uint32_t)(pr >> 32)] = 1;
    flags[(uint32_t)pr] = 1;
}

__global__ __launch_bounds__(256) void k_batch_bounds_init(int *__restrict__ mb, uint32_t nmeshes)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < 6 * nmeshes)
        mb[i] = (i % 6) < 3 ? 0x7fffffff : (int)0x80000000;
}

// ---- launchers ---------------------------------------------------------------
static unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

void launch_batch_bounds_init(int *mb, uint32_t nmeshes, hipStream_t s)
{
    if (nmeshes)
        hipLaunchKernelGGL(k_batch_bounds_init, dim3(blocks_for(6ull * nmeshes)), dim3(256), 0, s, mb, nmeshes);
}

// waves: 6 * ceil(nt / 64) ints of scratch
void launch_batch_bounds(const float *box32, const uint32_t *meshOf, const uint32_t *triBegin, uint32_t nmeshes,
                         uint32_t nt, int *mb, int *waves, hipStream_t s)
{
    if (!nt)
        return;
    hipLaunchKernelGGL(k_batch_bounds, dim3(blocks_for(nt)), dim3(256), 0, s, box32, meshOf, nt, mb, waves);
    hipLaunchKernelGGL(k_batch_bounds_merge, dim3(nmeshes), dim3(256), 0, s, waves, triBegin, nt, mb);
}

void launch_batch_keys(const float *box32, const uint32_t *meshOf, const uint32_t *triBegin, const int *mb,
                       uint32_t nt, int localBits, int mortonBits, uint64_t *keys, hipStream_t s)
{
    if (nt)
        hipLaunchKernelGGL(k_batch_keys, dim3(blocks_for(nt)), dim3(256), 0, s, box32, meshOf, triBegin, mb, nt,
                           localBits, mortonBits, keys);
}

void launch_batch_codes(const uint64_t *keys, uint32_t n, int localBits, int mortonBits, const uint32_t *triBegin,
                        uint64_t *codes, uint32_t *ids, hipStream_t s)
{
    if (n)
        hipLaunchKernelGGL(k_batch_codes, dim3(blocks_for(n)), dim3(256), 0, s, keys, n, localBits, mortonBits,
                           triBegin, codes, ids);
}

void launch_mesh_roots(const BNode *nodes, const uint32_t *ids, const uint32_t *triBegin, uint32_t nmeshes,
                       uint32_t n, const float *box32, const float *rootBox, uint32_t *meshRoot, float *meshBox,
                       hipStream_t s)
{
    if (n && nmeshes)
        hipLaunchKernelGGL(k_mesh_roots, dim3(blocks_for(nmeshes)), dim3(256), 0, s, nodes, ids, triBegin, nmeshes, n,
                           box32, rootBox, meshRoot, meshBox);
}

void launch_batch_seed(const TraverseArgs &t, const uint32_t *meshRoot, const float *meshBox, uint32_t npairs,
                       uint2 *frontier, uint32_t *frontCount, uint32_t frontCap, uint64_t *cand, uint32_t *candCount,
                       uint32_t candCap, uint32_t *overflow, hipStream_t s)
{
    if (npairs)
        hipLaunchKernelGGL(k_batch_seed, dim3(blocks_for(npairs)), dim3(256), 0, s, t, meshRoot, meshBox, npairs,
                           frontier, frontCount, frontCap, cand, candCount, candCap, overflow);
}

void launch_mark_both(const uint64_t *pairs, const uint8_t *status, const uint32_t *count, uint32_t cap,
                      uint8_t *flags, hipStream_t s)
{
    if (cap)
        hipLaunchKernelGGL(k_mark_both, dim3(blocks_for(cap)), dim3(256), 0, s, pairs, status, count, cap, flags);
}

}  // namespace sbg
