// FP32 certificate for one ray-parity candidate (reference: isPointInMesh,
// src/solidboolean.cpp:48-92, Vector3::intersectSegmentAndPlane
// src/vector3.h:264-280, Vector3::normal :155-176).
//
// The reference tests a triangle V0 V1 V2 against the ray from p along axis k
// in FP64: plane parameter s = num / d (num = -n.(p - V0), d = n.(e - p)),
// hit point H = p + s (e - p), three normals n_i = normal(H, V_i, V_i+1) and
// "hit" iff n0.n1 > 0 and n0.n2 > 0, then dedup of hits by PositionKey.
// This decides, from the 48-B FP32 record alone, what that FP64 sequence
// answers, or says it cannot.
//
// Exact-arithmetic picture.  With (k, u, w) the axes in cyclic order and
// O_i = orient2d_uw(p, V_i, V_i+1) (the k-component of (V_i - p) x
// (V_i+1 - p)), the line p + t e_k meets the triangle's plane at H* with
// H*_u = p_u, H*_w = p_w, and cross(V_i - H*, V_i+1 - H*) = (O_i / N_k) N
// (N the unit normal), so n_i.n_j has the sign of O_i O_j: inside iff the
// three O_i share one strict sign.  The distance along the ray is
// t* = H*_k - p_k = (O_1 A_0 + O_2 A_1 + O_0 A_2) / (O_0 + O_1 + O_2),
// A_i = V_i,k - p_k (barycentric weights O_i+1 / sum O).
//
// The FP64 sequence agrees with that picture once the quantities clear
// margins that bound its own rounding:
// * kappa = |e1||e2| / |N~| <= E^2 / |sum O| (E bounds an edge length, N~ the
//   unnormalised normal, |N~| >= |N~_k| = |sum O|) bounds the stored unit
//   normal's componentwise error by (32 kappa + 8) u (u = 2^-53), and
//   1 / |N_k| <= kappa.  The certificate needs kappa <= 2^16.
// * d = n.(e - p) = n_k DBL_MAX (1 +- 2u) (e_k = DBL_MAX for |p_k| < 2^970,
//   and |n_k| <= 1 keeps it finite), so s is about t* / DBL_MAX: a
//   subnormal, quantised in steps that put 8.9e-16 of absolute error on
//   s (e - p)_k.  Altogether |H_k - (p_k + t*)| <= Delta with
//   Delta = (kappa (260 kappa + 96) + 14) u M + 1e-15 (M bounds every
//   |coordinate|, M < 2^20), and |s (e - p)_u| < 3.1e-310, so H_u = p_u,
//   H_w = p_w, or both key to 0: every hit of one ray has p's off-axis
//   PositionKey components.
// * With H off the plane by at most Delta along k, each computed cross
//   product differs from (O_i / N_k) N by at most Delta E + 16 u (E + Delta)^2
//   (its own rounding), and normalising and the dot add at most 11 u in
//   relative terms; so n_i.n_j has the exact sign of O_i O_j, and no normal is
//   zeroed (|cross| <= DBL_EPSILON), once |O_i|, |O_j| >= F with
//   F = 16 (Delta E + 16 u (E + Delta)^2 + DBL_EPSILON).
// * s >= 0 (plane ahead) needs sign(num) = sign(d), certain once
//   |t*| > Z = 6.1 kappa (32 kappa + 13) u M; s < 0 (behind) likewise.
//
// FP32 evaluation.  Coordinates are rounded to FP32 (error <= eta = 2^-24 M
// each); a = fl(V - p) then carries |a - A| <= 2 eta + 1.0000001 2^-24 |a|,
// and O_i from a 2x2 determinant is off by at most
// 4.0000002 2^-24 P + 2.0000002 eta S + 8 eta^2 (P = |a_u b_w| + |a_w b_u|,
// S = |a_u| + |a_w| + |b_u| + |b_w|); the code uses 5, 2.5 and 9, which also
// absorb the rounding of P and S themselves, plus 1e-36 for FP32 underflow.
// t* comes from the same O_i with its own propagated bound.
//
// Verdicts: kRayMiss (a certain miss), kRayHit (a certain hit at
// t in [t - et, t + et], t measured from p along +k), kRayUnsure.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace sbg {

enum RayVerdict : int { kRayMiss = 0, kRayHit = 1, kRayUnsure = 2 };

struct RayCert {
    int verdict;
    float t, et;
};

constexpr float kE32 = 0x1p-24f;      // FP32 unit roundoff
constexpr float kU64 = 0x1p-53f;      // FP64 unit roundoff, as a float
constexpr float kDblEps = 2.220446049250313080847e-16f * 1.0000001f;  // DBL_EPSILON, rounded up
constexpr float kMaxMag = 0x1p20f;    // coordinates the certificate accepts
constexpr float kMaxKappa = 0x1p16f;  // triangle shape / obliquity it accepts

// Rounding toward +inf / -inf of an FP64 value to FP32, for host and device.
__host__ __device__ __forceinline__ float fl_up(double x)
{
    float f = (float)x;
    if ((double)f < x)
        f = nextafterf(f, __builtin_inff());
    return f;
}
__host__ __device__ __forceinline__ float fl_down(double x)
{
    float f = (float)x;
    if ((double)f > x)
        f = nextafterf(f, -__builtin_inff());
    return f;
}

// The 48-B FP32 record of triangle (a, b, c): the three vertices rounded to
// nearest, the first pad an upper bound of the largest box extent, the second
// an upper bound of the largest |coordinate|.
__host__ __device__ __forceinline__ void make_tri32(const double a[3], const double b[3], const double c[3],
                                                    float4 out[3])
{
    double mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = a[k] < b[k] ? a[k] : b[k];
        mn[k] = c[k] < mn[k] ? c[k] : mn[k];
        mx[k] = a[k] > b[k] ? a[k] : b[k];
        mx[k] = c[k] > mx[k] ? c[k] : mx[k];
    }
    double ext = mx[0] - mn[0], mag = 0.0;
    for (int k = 1; k < 3; ++k)
        ext = (mx[k] - mn[k]) > ext ? (mx[k] - mn[k]) : ext;
    for (int k = 0; k < 3; ++k) {
        const double lo = mn[k] < 0.0 ? -mn[k] : mn[k], hi = mx[k] < 0.0 ? -mx[k] : mx[k];
        mag = lo > mag ? lo : mag;
        mag = hi > mag ? hi : mag;
    }
    out[0] = make_float4((float)a[0], (float)a[1], (float)a[2], fl_up(ext * (1.0 + 0x1p-40)));
    out[1] = make_float4((float)b[0], (float)b[1], (float)b[2], fl_up(mag));
    out[2] = make_float4((float)c[0], (float)c[1], (float)c[2], 0.0f);
}

// component k of (x, y, z), by selects (no indexed register array)
__host__ __device__ __forceinline__ float comp_of(const float4 &v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// p32: p rounded to nearest (w unused); pm: an upper bound of p's largest
// |coordinate|; k: the ray's axis; r: the record.
__host__ __device__ __forceinline__ RayCert ray_certificate(const float4 &p32, float pm, int k, const float4 r[3])
{
    RayCert out = {kRayUnsure, 0.0f, 0.0f};
    const int u = k == 2 ? 0 : k + 1, w = k == 0 ? 2 : k - 1;
    const float M = pm > r[1].w ? pm : r[1].w;
    if (!(M < kMaxMag))
        return out;
    const float eta = M * kE32;  // input rounding bound per coordinate (exact: a power-of-2 scaling)
    const float ext = r[0].w;
    const float E = 1.7320511f * 1.000001f * ext;  // edge length bound: sqrt(3) ext
    // differences V_i - p, on the three axes
    float ak[3], au[3], aw[3];
    const float pk = comp_of(p32, k), pu = comp_of(p32, u), pw = comp_of(p32, w);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ak[i] = comp_of(r[i], k) - pk;
        au[i] = comp_of(r[i], u) - pu;
        aw[i] = comp_of(r[i], w) - pw;
    }
    float O[3], eO[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = i == 2 ? 0 : i + 1;
        const float x1 = au[i] * aw[j], x2 = aw[i] * au[j];
        O[i] = x1 - x2;
        const float P = __builtin_fabsf(x1) + __builtin_fabsf(x2);
        const float S = __builtin_fabsf(au[i]) + __builtin_fabsf(aw[i]) + __builtin_fabsf(au[j]) +
                        __builtin_fabsf(aw[j]);
        eO[i] = 5.0f * kE32 * P + 2.5f * eta * S + 9.0f * eta * eta + 1e-36f;
    }
    // |sum O| from below gives kappa from above
    const float sumO = O[0] + O[1] + O[2];
    const float absO = __builtin_fabsf(O[0]) + __builtin_fabsf(O[1]) + __builtin_fabsf(O[2]);
    const float eSum = (eO[0] + eO[1] + eO[2]) * 1.0001f + 2.1f * kE32 * absO;
    const float sumLo = __builtin_fabsf(sumO) - eSum;
    if (!(sumLo > 0.0f))
        return out;
    const float kappa = E * E / sumLo * 1.0001f;
    if (!(kappa <= kMaxKappa))
        return out;
    const float Delta = (kappa * (260.0f * kappa + 96.0f) + 14.0f) * kU64 * M * 1.0001f + 1e-15f;
    const float Ed = E + Delta;
    const float F = 16.0f * (Delta * E + 16.0f * kU64 * Ed * Ed + kDblEps) * 1.0001f;
    // per O_i: +1 / -1 when its sign is certain and |O_i| >= F, 0 otherwise
    int sg[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float lo = __builtin_fabsf(O[i]) - eO[i] * 1.0001f;
        sg[i] = lo >= F ? (O[i] > 0.0f ? 1 : -1) : 0;
    }
    // the reference's test: n0.n1 > 0 and n0.n2 > 0
    if ((sg[0] && sg[1] && sg[0] != sg[1]) || (sg[0] && sg[2] && sg[0] != sg[2])) {
        out.verdict = kRayMiss;
        return out;
    }
    if (!(sg[0] && sg[0] == sg[1] && sg[0] == sg[2]))
        return out;
    // inside: t from the barycentric weights (all O_i of one sign, so no
    // cancellation in the sum)
    float eN = 0.0f, absN = 0.0f, T = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int j = i == 2 ? 0 : i + 1;  // weight of vertex i is O_j
        const float ra = 2.0f * eta + 1.0000002f * kE32 * __builtin_fabsf(ak[i]);
        eN += __builtin_fabsf(O[j]) * ra + (__builtin_fabsf(ak[i]) + ra) * eO[j];
        absN += __builtin_fabsf(O[j] * ak[i]);
        const float ta = __builtin_fabsf(ak[i]) + ra;
        T = ta > T ? ta : T;
    }
    const float num = O[1] * ak[0] + O[2] * ak[1] + O[0] * ak[2];
    eN = eN * 1.0001f + 3.1f * kE32 * absN;
    const float den = __builtin_fabsf(sumO);
    if (!(den > 2.0f * eSum))
        return out;
    const float t = num / sumO;
    const float et = ((eN + T * eSum) / (den - eSum) * 1.0001f + 3.0f * kE32 * __builtin_fabsf(t)) * 1.0001f + Delta;
    const float Z = 6.1f * kappa * (32.0f * kappa + 13.0f) * kU64 * M * 1.0001f;
    if (t - et > Z) {
        out.verdict = kRayHit;
        out.t = t;
        out.et = et;
    } else if (t + et < -Z) {
        out.verdict = kRayMiss;
    }
    return out;
}

// Certain hits of one ray, CAP of them in registers: their distances t from
// the ray's origin and error bounds.  A hit not certainly apart from one
// already held, or one more than CAP, sets `full` (the ray is then run in
// FP64).  New hits enter at slot 0 and shift the others up, so every slot
// index is a compile-time constant.
template <int CAP>
struct HitSpans {
    float t[CAP], e[CAP];
    int n = 0;
    bool full = false;
    __host__ __device__ __forceinline__ void add(float ht, float he);
};

// Two certain hits of one ray key to different PositionKeys on the ray's axis
// when their FP64 hit points are more than 2e-5 apart (the truncated
// products differ by more than 2, zero's double-width cell included).
__host__ __device__ __forceinline__ bool hits_apart(float ta, float ea, float tb, float eb)
{
    const float gap = __builtin_fabsf(ta - tb) * (1.0f - 2.0f * kE32);
    return gap - ea - eb > 2.0001e-5f + 4e-15f * kMaxMag;
}

template <int CAP>
__host__ __device__ __forceinline__ void HitSpans<CAP>::add(float ht, float he)
{
#pragma unroll
    for (int j = 0; j < CAP; ++j)
        if (j < n && !hits_apart(ht, he, t[j], e[j])) {
            full = true;
            return;
        }
    if (n == CAP) {
        full = true;
        return;
    }
#pragma unroll
    for (int j = CAP - 1; j > 0; --j) {
        t[j] = t[j - 1];
        e[j] = e[j - 1];
    }
    t[0] = ht;
    e[0] = he;
    ++n;
}

}  // namespace sbg
