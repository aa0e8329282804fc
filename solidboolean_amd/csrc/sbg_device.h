// Device-side geometry for the solidboolean hot path on gfx950.
//
// Every FP64 routine here rounds exactly where the reference's C/C++ does:
// the library is built with -ffp-contract=off (and the pragma below), FP64
// '/' and sqrt are IEEE correctly rounded on CDNA, and operand order and
// association follow the cited reference lines term by term.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace sbg {

struct V3 {
    double x, y, z;
};

__device__ __forceinline__ V3 vsub(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

// CROSS (thirdparty/GuigueDevillers03/tri_tri_intersect.c:73-76)
__device__ __forceinline__ V3 vcross(const V3 &a, const V3 &b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// DOT (tri_tri_intersect.c:78), left to right
__device__ __forceinline__ double vdot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Vector3::normal (src/vector3.h:155-176) with Double::isZero (src/double.h:31-34)
__device__ __forceinline__ V3 tri_normal(const V3 &a, const V3 &b, const V3 &c)
{
    double baX = b.x - a.x, baY = b.y - a.y, baZ = b.z - a.z;
    double caX = c.x - a.x, caY = c.y - a.y, caZ = c.z - a.z;
    double crossX = baY * caZ - baZ * caY;
    double crossY = baZ * caX - baX * caZ;
    double crossZ = baX * caY - baY * caX;
    double length2 = crossX * crossX + crossY * crossY + crossZ * crossZ;
    double length = __builtin_sqrt(length2);
    if (__builtin_fabs(length) <= 2.220446049250313080847e-16)
        return {0.0, 0.0, 0.0};
    return {crossX / length, crossY / length, crossZ / length};
}

// One intersection-segment endpoint (tri_tri_intersect.c:297-301 and the seven
// sibling blocks of CONSTRUCT_INTERSECTION): v1 = base - onplane,
// v2 = base - other, alpha = DOT(v1,N)/DOT(v2,N), out = base - alpha * v2.
__device__ __forceinline__ V3 edge_plane(const V3 &base, const V3 &onplane, const V3 &other, const V3 &N)
{
    V3 v1 = vsub(base, onplane);
    V3 v2 = vsub(base, other);
    double alpha = vdot(v1, N) / vdot(v2, N);
    V3 s = {alpha * v2.x, alpha * v2.y, alpha * v2.z};
    return vsub(base, s);
}

// CONSTRUCT_INTERSECTION (tri_tri_intersect.c:285-356).
__device__ __forceinline__ int gd_construct(const V3 &p1, const V3 &q1, const V3 &r1, const V3 &p2,
                                            const V3 &q2, const V3 &r2, const V3 &N1, const V3 &N2,
                                            V3 &src, V3 &tgt)
{
    V3 v1 = vsub(q1, p1);
    V3 v2 = vsub(r2, p1);
    V3 N = vcross(v1, v2);
    V3 v = vsub(p2, p1);
    if (vdot(v, N) > 0.0) {
        v1 = vsub(r1, p1);
        N = vcross(v1, v2);
        if (vdot(v, N) <= 0.0) {
            v2 = vsub(q2, p1);
            N = vcross(v1, v2);
            if (vdot(v, N) > 0.0) {
                src = edge_plane(p1, p2, r1, N2);
                tgt = edge_plane(p2, p1, r2, N1);
            } else {
                src = edge_plane(p2, p1, q2, N1);
                tgt = edge_plane(p2, p1, r2, N1);
            }
            return 1;
        }
        return 0;
    }
    v2 = vsub(q2, p1);
    N = vcross(v1, v2);
    if (vdot(v, N) < 0.0)
        return 0;
    v1 = vsub(r1, p1);
    N = vcross(v1, v2);
    if (vdot(v, N) >= 0.0) {
        src = edge_plane(p1, p2, r1, N2);
        tgt = edge_plane(p1, p2, q1, N2);
    } else {
        src = edge_plane(p2, p1, q2, N1);
        tgt = edge_plane(p1, p2, q1, N2);
    }
    return 1;
}

// TRI_TRI_INTER_3D (tri_tri_intersect.c:360-385): 1 hit, 0 none, -1 coplanar.
__device__ __forceinline__ int gd_inter3d(const V3 &p1, const V3 &q1, const V3 &r1, const V3 &p2,
                                          const V3 &q2, const V3 &r2, double dp2, double dq2,
                                          double dr2, const V3 &N1, const V3 &N2, V3 &s, V3 &t)
{
    if (dp2 > 0.0) {
        if (dq2 > 0.0)
            return gd_construct(p1, r1, q1, r2, p2, q2, N1, N2, s, t);
        if (dr2 > 0.0)
            return gd_construct(p1, r1, q1, q2, r2, p2, N1, N2, s, t);
        return gd_construct(p1, q1, r1, p2, q2, r2, N1, N2, s, t);
    }
    if (dp2 < 0.0) {
        if (dq2 < 0.0)
            return gd_construct(p1, q1, r1, r2, p2, q2, N1, N2, s, t);
        if (dr2 < 0.0)
            return gd_construct(p1, q1, r1, q2, r2, p2, N1, N2, s, t);
        return gd_construct(p1, r1, q1, p2, q2, r2, N1, N2, s, t);
    }
    if (dq2 < 0.0) {
        if (dr2 >= 0.0)
            return gd_construct(p1, r1, q1, q2, r2, p2, N1, N2, s, t);
        return gd_construct(p1, q1, r1, p2, q2, r2, N1, N2, s, t);
    }
    if (dq2 > 0.0) {
        if (dr2 > 0.0)
            return gd_construct(p1, r1, q1, p2, q2, r2, N1, N2, s, t);
        return gd_construct(p1, q1, r1, q2, r2, p2, N1, N2, s, t);
    }
    if (dr2 > 0.0)
        return gd_construct(p1, q1, r1, r2, p2, q2, N1, N2, s, t);
    if (dr2 < 0.0)
        return gd_construct(p1, r1, q1, r2, p2, q2, N1, N2, s, t);
    return -1;
}

// Exit kinds recorded per pair (for the exit histogram).
enum GdExit : uint8_t {
    GD_REJECT_PLANE1 = 0,
    GD_REJECT_PLANE2 = 1,
    GD_REJECT_INTERVAL = 2,
    GD_HIT = 3,
    GD_COPLANAR = 4,
};

// tri_tri_intersection_test_3d (tri_tri_intersect.c:395-472) as used by
// intersectTwoFaces (src/solidboolean.cpp:103-122). Returns a GdExit.
__device__ __forceinline__ int gd_fp64(const V3 &p1, const V3 &q1, const V3 &r1, const V3 &p2,
                                       const V3 &q2, const V3 &r2, V3 &src, V3 &tgt)
{
    V3 N2 = vcross(vsub(p2, r2), vsub(q2, r2));
    double dp1 = vdot(vsub(p1, r2), N2);
    double dq1 = vdot(vsub(q1, r2), N2);
    double dr1 = vdot(vsub(r1, r2), N2);
    if (((dp1 * dq1) > 0.0) && ((dp1 * dr1) > 0.0))
        return GD_REJECT_PLANE1;
    V3 N1 = vcross(vsub(q1, p1), vsub(r1, p1));
    double dp2 = vdot(vsub(p2, r1), N1);
    double dq2 = vdot(vsub(q2, r1), N1);
    double dr2 = vdot(vsub(r2, r1), N1);
    if (((dp2 * dq2) > 0.0) && ((dp2 * dr2) > 0.0))
        return GD_REJECT_PLANE2;
    int r;
    if (dp1 > 0.0) {
        if (dq1 > 0.0)
            r = gd_inter3d(r1, p1, q1, p2, r2, q2, dp2, dr2, dq2, N1, N2, src, tgt);
        else if (dr1 > 0.0)
            r = gd_inter3d(q1, r1, p1, p2, r2, q2, dp2, dr2, dq2, N1, N2, src, tgt);
        else
            r = gd_inter3d(p1, q1, r1, p2, q2, r2, dp2, dq2, dr2, N1, N2, src, tgt);
    } else if (dp1 < 0.0) {
        if (dq1 < 0.0)
            r = gd_inter3d(r1, p1, q1, p2, q2, r2, dp2, dq2, dr2, N1, N2, src, tgt);
        else if (dr1 < 0.0)
            r = gd_inter3d(q1, r1, p1, p2, q2, r2, dp2, dq2, dr2, N1, N2, src, tgt);
        else
            r = gd_inter3d(p1, q1, r1, p2, r2, q2, dp2, dr2, dq2, N1, N2, src, tgt);
    } else {
        if (dq1 < 0.0) {
            if (dr1 >= 0.0)
                r = gd_inter3d(q1, r1, p1, p2, r2, q2, dp2, dr2, dq2, N1, N2, src, tgt);
            else
                r = gd_inter3d(p1, q1, r1, p2, q2, r2, dp2, dq2, dr2, N1, N2, src, tgt);
        } else if (dq1 > 0.0) {
            if (dr1 > 0.0)
                r = gd_inter3d(p1, q1, r1, p2, r2, q2, dp2, dr2, dq2, N1, N2, src, tgt);
            else
                r = gd_inter3d(q1, r1, p1, p2, q2, r2, dp2, dq2, dr2, N1, N2, src, tgt);
        } else {
            if (dr1 > 0.0)
                r = gd_inter3d(r1, p1, q1, p2, q2, r2, dp2, dq2, dr2, N1, N2, src, tgt);
            else if (dr1 < 0.0)
                r = gd_inter3d(r1, p1, q1, p2, r2, q2, dp2, dr2, dq2, N1, N2, src, tgt);
            else
                r = -1;
        }
    }
    if (r < 0)
        return GD_COPLANAR;
    return r ? GD_HIT : GD_REJECT_INTERVAL;
}

// ---------------------------------------------------------------------------
// FP32 filter geometry.  Every GD predicate has the shape
//   d = DOT(a, CROSS(b, c)),  a, b, c differences of input points.
// Evaluated from FP32-rounded coordinates (u = 2^-24), each difference carries
// an input-rounding error |alpha| <= eta = 2uM (M bounds every |coordinate|)
// plus its own rounding; expanding (A+alpha)(B+beta)(C+kappa)(1+theta_8) over
// the 6 monomials of the triple product gives, with P = sum |a_i|(|b_j c_k| +
// |b_k c_j|), Q = sum_i (|b_j c_k| + |b_k c_j|) + |a_i|(|b_j|+|c_k|+|b_k|+|c_j|)
// and S1 = sum of all nine |components| (all from the FP32 differences):
//   |d32 - d_exact| <= eta Q (1+5u) + 4 eta^2 S1 (1+2u) + 42 eta^3 + 8u P (1+11u)
// and the FP64 reference's own error is below 8 * 2^-53 P (1 + tiny).  The
// kernel uses 1.5 eta Q + 5 eta^2 S1 + 50 eta^3 + 12 u P + 1e-36, which also
// absorbs the rounding of P, Q, S1 and E themselves, FP32 gradual underflow
// and the 2uM(1+2u) rounding of eta; a sign is certified when |d32| exceeds it.
// ---------------------------------------------------------------------------
struct F3 {
    float x, y, z;
};

__device__ __forceinline__ F3 fsub(const F3 &a, const F3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }

// Triangle box as AxisAlignedBoudingBox::update builds it
// (src/axisalignedboundingbox.h:31-41, initial state :108-117).
struct Box64 {
    V3 mn, mx;
};
__device__ __forceinline__ void box_update(Box64 &b, const V3 &p)
{
    if (p.x > b.mx.x) b.mx.x = p.x;
    if (p.x < b.mn.x) b.mn.x = p.x;
    if (p.y > b.mx.y) b.mx.y = p.y;
    if (p.y < b.mn.y) b.mn.y = p.y;
    if (p.z > b.mx.z) b.mx.z = p.z;
    if (p.z < b.mn.z) b.mn.z = p.z;
}
__device__ __forceinline__ Box64 tri_box64(const V3 &a, const V3 &b, const V3 &c)
{
    const double M = 1.7976931348623157e308;
    Box64 r = {{M, M, M}, {-M, -M, -M}};
    box_update(r, a);
    box_update(r, b);
    box_update(r, c);
    return r;
}
// closed-interval overlap (src/axisalignedboundingbox.h:95-105)
__device__ __forceinline__ bool box64_overlap(const Box64 &a, const Box64 &b)
{
    return a.mn.x <= b.mx.x && a.mx.x >= b.mn.x && a.mn.y <= b.mx.y && a.mx.y >= b.mn.y &&
           a.mn.z <= b.mx.z && a.mx.z >= b.mn.z;
}

// Outward-rounded FP32 box of an FP64 interval.
__device__ __forceinline__ float f_down(double x) { return __double2float_rd(x); }
__device__ __forceinline__ float f_up(double x) { return __double2float_ru(x); }

// ---- XCD-aware block order ---------------------------------------------------
// The chip dispatches workgroups to its 8 XCDs, each with its own L2, in
// round-robin order of the linear block index.  xcd_block maps the hardware
// index so that each XCD works through one contiguous range of logical blocks;
// neighbouring blocks, which read the same data, then share an L2.  The grid's
// x dimension must be a multiple of 8 (xcd_grid); logical blocks past the real
// count must return.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t hw, uint32_t gridBlocks)
{
    return (hw % kXcds) * (gridBlocks / kXcds) + hw / kXcds;
}

// ---- Morton codes ----------------------------------------------------------
// spread the low 10 bits of x to every third bit
__device__ __forceinline__ uint32_t spread3(uint32_t x)
{
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// centre of triangle i's FP32 leaf box (8 floats: lo xyz, pad, hi xyz, pad)
__device__ __forceinline__ float3 box_centre(const float *__restrict__ box32, uint32_t i)
{
    const float4 lo = reinterpret_cast<const float4 *>(box32 + (size_t)8 * i)[0];
    const float4 hi = reinterpret_cast<const float4 *>(box32 + (size_t)8 * i)[1];
    return make_float3(0.5f * lo.x + 0.5f * hi.x, 0.5f * lo.y + 0.5f * hi.y, 0.5f * lo.z + 0.5f * hi.z);
}

// 30-bit Morton code of c within bounds (lo xyz, hi xyz), 10 bits per axis
__device__ __forceinline__ uint32_t morton30(float3 c, const float *bounds)
{
    const float cc[3] = {c.x, c.y, c.z};
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float ext = bounds[3 + k] - bounds[k];
        float t = ext > 0.0f ? (cc[k] - bounds[k]) / ext : 0.0f;
        t = fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f);
        q[k] = (uint32_t)t;
    }
    return (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
}

// PositionKey (src/positionkey.cpp:25-37): (long)(x * 100000), truncation.
struct Key3 {
    long long x, y, z;
};
__device__ __forceinline__ Key3 position_key(const V3 &p)
{
    return {(long long)(p.x * 100000.0), (long long)(p.y * 100000.0), (long long)(p.z * 100000.0)};
}

}  // namespace sbg
