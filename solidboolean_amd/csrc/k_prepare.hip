// Triangle records, normals, leaf boxes and Morton keys: the device side of
// SolidMesh::prepare (reference: src/solidmesh.cpp:42-76).
//
// One fused pass per mesh reads each triangle's 3 indices and 9 FP64
// coordinates once and writes every per-triangle record the later stages
// stream: the 96-B FP64 record, the 48-B FP32 record, the FP64 normal
// (src/vector3.h:155-176) and the outward-rounded FP32 box.  Algorithmic bytes
// per triangle: 12 (indices) + 72 (coordinate gathers) + 96 + 48 + 24 + 32.
#include "sbg_device.h"
#include "sbg_internal.h"

#pragma clang fp contract(off)

namespace sbg {

constexpr int kRecBlock = 256;

__global__ __launch_bounds__(kRecBlock) void k_records(const double *__restrict__ V,
                                                       const uint32_t *__restrict__ T, uint32_t nt,
                                                       double *__restrict__ tri64, float *__restrict__ tri32,
                                                       float *__restrict__ box32, float *__restrict__ blockBounds)
{
    const uint32_t i = blockIdx.x * kRecBlock + threadIdx.x;
    float cmin[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
    float cmax[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    if (i < nt) {
        const uint32_t ia = T[3 * i], ib = T[3 * i + 1], ic = T[3 * i + 2];
        V3 a = {V[3 * ia], V[3 * ia + 1], V[3 * ia + 2]};
        V3 b = {V[3 * ib], V[3 * ib + 1], V[3 * ib + 2]};
        V3 c = {V[3 * ic], V[3 * ic + 1], V[3 * ic + 2]};
        // FP64 record: each vertex's pad slot carries one normal component, so
        // a ray's triangle test reads one 96-B record and nothing else
        const V3 n = tri_normal(a, b, c);
        double2 *r64 = reinterpret_cast<double2 *>(tri64 + (size_t)kTri64Stride * i);
        r64[0] = make_double2(a.x, a.y);
        r64[1] = make_double2(a.z, n.x);
        r64[2] = make_double2(b.x, b.y);
        r64[3] = make_double2(b.z, n.y);
        r64[4] = make_double2(c.x, c.y);
        r64[5] = make_double2(c.z, n.z);
        // exact FP64 min/max (AxisAlignedBoudingBox::update), then rounded outward
        Box64 bb = tri_box64(a, b, c);
        // FP32 record; its pad slots carry upper bounds of the triangle's
        // largest box extent and of its largest |coordinate| (the FP32
        // filter's error bound is built from them)
        const double ext = fmax(fmax(bb.mx.x - bb.mn.x, bb.mx.y - bb.mn.y), bb.mx.z - bb.mn.z);
        const double mag = fmax(fmax(fmax(fabs(bb.mn.x), fabs(bb.mx.x)), fmax(fabs(bb.mn.y), fabs(bb.mx.y))),
                                fmax(fabs(bb.mn.z), fabs(bb.mx.z)));
        float4 *r32 = reinterpret_cast<float4 *>(tri32 + (size_t)kTri32Stride * i);
        r32[0] = make_float4((float)a.x, (float)a.y, (float)a.z, f_up(ext * (1.0 + 0x1p-40)));
        r32[1] = make_float4((float)b.x, (float)b.y, (float)b.z, f_up(mag));
        r32[2] = make_float4((float)c.x, (float)c.y, (float)c.z, 0.0f);
        float4 lo = make_float4(f_down(bb.mn.x), f_down(bb.mn.y), f_down(bb.mn.z), 0.0f);
        float4 hi = make_float4(f_up(bb.mx.x), f_up(bb.mx.y), f_up(bb.mx.z), 0.0f);
        float4 *bx = reinterpret_cast<float4 *>(box32 + (size_t)kBox32Stride * i);
        bx[0] = lo;
        bx[1] = hi;
        // centroid of the FP32 box, for the Morton frame
        float cx = 0.5f * lo.x + 0.5f * hi.x, cy = 0.5f * lo.y + 0.5f * hi.y, cz = 0.5f * lo.z + 0.5f * hi.z;
        cmin[0] = cmax[0] = cx;
        cmin[1] = cmax[1] = cy;
        cmin[2] = cmax[2] = cz;
    }
    // block reduction of the centroid bounds
    __shared__ float s[6][kRecBlock / 64];
    for (int k = 0; k < 3; ++k) {
        float lo = cmin[k], hi = cmax[k];
        for (int off = 32; off > 0; off >>= 1) {
            lo = fminf(lo, __shfl_xor(lo, off));
            hi = fmaxf(hi, __shfl_xor(hi, off));
        }
        if ((threadIdx.x & 63) == 0) {
            s[k][threadIdx.x >> 6] = lo;
            s[3 + k][threadIdx.x >> 6] = hi;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        float v = s[threadIdx.x][0];
        for (int w = 1; w < kRecBlock / 64; ++w)
            v = threadIdx.x < 3 ? fminf(v, s[threadIdx.x][w]) : fmaxf(v, s[threadIdx.x][w]);
        blockBounds[6 * blockIdx.x + threadIdx.x] = v;
    }
}

__global__ __launch_bounds__(256) void k_bounds_reduce(const float *__restrict__ blockBounds, int nblocks,
                                                       float *__restrict__ bounds)
{
    __shared__ float s[6][256];
    float v[6] = {__builtin_inff(), __builtin_inff(), __builtin_inff(),
                  -__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
    for (int b = threadIdx.x; b < nblocks; b += 256)
        for (int k = 0; k < 6; ++k)
            v[k] = k < 3 ? fminf(v[k], blockBounds[6 * b + k]) : fmaxf(v[k], blockBounds[6 * b + k]);
    for (int k = 0; k < 6; ++k)
        s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride)
            for (int k = 0; k < 6; ++k)
                s[k][threadIdx.x] = k < 3 ? fminf(s[k][threadIdx.x], s[k][threadIdx.x + stride])
                                          : fmaxf(s[k][threadIdx.x], s[k][threadIdx.x + stride]);
        __syncthreads();
    }
    if (threadIdx.x < 6)
        bounds[threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_morton(const float *__restrict__ box32, const float *__restrict__ bounds,
                                                uint32_t nt, uint32_t *__restrict__ keys,
                                                uint32_t *__restrict__ vals, uint32_t *__restrict__ zero,
                                                uint32_t nzero)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < nzero)
        zero[i] = 0u;
    if (i >= nt)
        return;
    // 8 bits per axis: a 3-pass sort, and the tree's order is as good at
    // these sizes (equal codes keep triangle order, which is local too)
    keys[i] = morton30(box_centre(box32, i), bounds) >> (30 - kMortonBits);
    vals[i] = i;
}

// exact FP64 boxes (min xyz, max xyz) from the FP64 records
__global__ __launch_bounds__(256) void k_boxes64(const double *__restrict__ tri64, uint32_t nt,
                                                 double *__restrict__ out)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nt)
        return;
    const double *r = tri64 + (size_t)kTri64Stride * i;
    const Box64 b = tri_box64({r[0], r[1], r[2]}, {r[4], r[5], r[6]}, {r[8], r[9], r[10]});
    double *o = out + 6 * (size_t)i;
    o[0] = b.mn.x;
    o[1] = b.mn.y;
    o[2] = b.mn.z;
    o[3] = b.mx.x;
    o[4] = b.mx.y;
    o[5] = b.mx.z;
}

// FP64 normals (the records' pad slots) as xyz triples
__global__ __launch_bounds__(256) void k_normals64(const double *__restrict__ tri64, uint32_t nt,
                                                   double *__restrict__ out)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nt)
        return;
    const double *r = tri64 + (size_t)kTri64Stride * i;
    out[3 * (size_t)i] = r[3];
    out[3 * (size_t)i + 1] = r[7];
    out[3 * (size_t)i + 2] = r[11];
}

void launch_normals64(const double *tri64, uint32_t nt, double *out, hipStream_t s)
{
    if (nt == 0)
        return;
    hipLaunchKernelGGL(k_normals64, dim3((nt + 255) / 256), dim3(256), 0, s, tri64, nt, out);
}

void launch_boxes64(const double *tri64, uint32_t nt, double *out, hipStream_t s)
{
    if (nt == 0)
        return;
    hipLaunchKernelGGL(k_boxes64, dim3((nt + 255) / 256), dim3(256), 0, s, tri64, nt, out);
}

int records_blocks(uint32_t nt) { return (int)((nt + kRecBlock - 1) / kRecBlock); }

void launch_records(const double *V, const uint32_t *T, uint32_t nt, double *tri64, float *tri32, float *box32,
                    float *blockBounds, int nblocks, hipStream_t s)
{
    if (nt == 0)
        return;
    hipLaunchKernelGGL(k_records, dim3(nblocks), dim3(kRecBlock), 0, s, V, T, nt, tri64, tri32, box32,
                       blockBounds);
}

void launch_bounds_reduce(const float *blockBounds, int nblocks, float *bounds, hipStream_t s)
{
    hipLaunchKernelGGL(k_bounds_reduce, dim3(1), dim3(256), 0, s, blockBounds, nblocks, bounds);
}

void launch_morton(const float *box32, const float *bounds, uint32_t nt, uint32_t *keys, uint32_t *vals,
                   uint32_t *zero, uint32_t nzero, hipStream_t s)
{
    if (nt == 0)
        return;
    const uint32_t threads = nt > nzero ? nt : nzero;
    hipLaunchKernelGGL(k_morton, dim3((threads + 255) / 256), dim3(256), 0, s, box32, bounds, nt, keys, vals, zero,
                       nzero);
}

}  // namespace sbg
