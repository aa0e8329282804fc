// Inside/outside classification by ray parity (reference: isPointInMesh,
// src/solidboolean.cpp:48-92, its three ray directions :31-35, and the vote of
// decideGroupSide :482-510).
//
// One lane per (query point, axis), axis-major, and for triangle queries in
// the querying mesh's Morton order, so neighbouring lanes trace parallel rays
// from neighbouring points.  The ray is the segment p -> p + axis with
// axis = (DBL_MAX, eps, eps) and permutations; its FP64 box is rounded outward
// to FP32 for a first cut; every triangle whose exact FP64 box closed-overlaps
// the segment's box (the reference's candidate set, :55-63)
// gets the FP64 plane test (src/vector3.h:264-280) and the three-normal inside
// test (:78-86) in the reference's operation order.  Hits are deduplicated by
// PositionKey (src/positionkey.cpp:25-37); the lane's answer is the parity of
// the distinct keys (:89).
//
// The candidates come from a per-axis grid over the target's projected
// triangle boxes instead of a tree walk: a ray's candidates are the entries of
// the one cell holding its origin, a superset of the reference's set that the
// exact FP64 box test then cuts down to it.  k_rays_grid scans a cell of up
// to 48 entries with up to 3 distinct keys in registers; a ray with a crowded
// cell or more hits goes to k_rays_wave, where one wave splits the cell and
// counts distinct keys in LDS (up to 256 hits).  A ray past that raises
// `overflow` and the call fails: nothing is guessed.
#include "sbg_device.h"
#include "sbg_internal.h"
#include "sbg_rayfilter.h"

#include <cmath>

#pragma clang fp contract(off)

namespace sbg {

constexpr int kFastKeys = 3;     // distinct keys a fast lane keeps
constexpr int kFastCands = 48;   // cell entries a fast lane scans alone
constexpr int kCodeGroup = 8;    // entry codes a fast lane loads at once (4-8 beat 12 and 16 by 2-3 %)
constexpr int kWaveHits = 256;   // hits a wave gathers for one deferred ray
constexpr int kWaveLoads = 2;    // cell entries a wave lane tests per round

__global__ __launch_bounds__(256) void k_mark_intersected(const uint64_t *__restrict__ pairs,
                                                          const uint8_t *__restrict__ status, uint32_t n, int side,
                                                          uint8_t *__restrict__ flags)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || status[i] != 1)
        return;
    const uint64_t pr = pairs[i];
    flags[side == 0 ? (uint32_t)(pr >> 32) : (uint32_t)pr] = 1;
}

// (v0 + v1 + v2) / 3.0 (src/solidboolean.cpp:497-499)
__global__ __launch_bounds__(256) void k_centroids(const double *__restrict__ tri64, const uint8_t *__restrict__ flags,
                                                   uint32_t nt, double *__restrict__ pts)
{
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nt || (flags && flags[t]))
        return;
    const double *r = tri64 + (size_t)kTri64Stride * t;
    pts[3 * (size_t)t] = (r[0] + r[4] + r[8]) / 3.0;
    pts[3 * (size_t)t + 1] = (r[1] + r[5] + r[9]) / 3.0;
    pts[3 * (size_t)t + 2] = (r[2] + r[6] + r[10]) / 3.0;
}

// The same centroids in visiting order: query q (triangle order[q]) gets
// (x, y, z, tag) at q4 + 4q, with tag = the target mesh (target[t], 0 without
// a target list), kTagSkip for a triangle flagged as intersected (with flags;
// without, the queries do not depend on the intersections and the rays read
// the flags themselves), kTagNone for one without a target mesh.  A ray lane
// then reads its point without first waiting for order[q].
constexpr double kTagSkip = -1.0, kTagNone = -2.0;

__global__ __launch_bounds__(256) void k_centroids_ordered(const double *__restrict__ tri64,
                                                           const uint8_t *__restrict__ flags,
                                                           const uint32_t *__restrict__ order,
                                                           const uint32_t *__restrict__ target, uint32_t nt,
                                                           double *__restrict__ q4)
{
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nt)
        return;
    const uint32_t t = order[q];
    double2 *out = reinterpret_cast<double2 *>(q4) + 2 * (size_t)q;
    if (flags && flags[t]) {
        out[1] = make_double2(0.0, kTagSkip);
        return;
    }
    const uint32_t m = target ? target[t] : 0u;
    const double *r = tri64 + (size_t)kTri64Stride * t;
    out[0] = make_double2((r[0] + r[4] + r[8]) / 3.0, (r[1] + r[5] + r[9]) / 3.0);
    out[1] = make_double2((r[2] + r[6] + r[10]) / 3.0, m == ~0u ? kTagNone : (double)m);
}

// A ray is the segment from p to e = p + dir (src/solidboolean.cpp:60-66):
// dir is DBL_MAX along the axis and DBL_EPSILON along the other two, so
// e >= p in every component and the segment's box (:67-69) is exactly [p, e].
// Only p is kept; e is recomputed where it is needed.
struct Ray {
    V3 p;
    uint32_t axis;
    __device__ __forceinline__ V3 end() const
    {
        const double eps = 2.220446049250313080847e-16, big = 1.7976931348623157e308;
        return {p.x + (axis == 0 ? big : eps), p.y + (axis == 1 ? big : eps), p.z + (axis == 2 ? big : eps)};
    }
};

__device__ __forceinline__ Ray ray_from(const V3 &p, uint32_t axis)
{
    Ray r;
    r.p = p;
    r.axis = axis;
    return r;
}

// Lane g = axis * n + q: its ray, output index i and target mesh m.  Q4: the
// points are (x, y, z, tag) records in visiting order; otherwise xyz triples
// by output index, with optional order / gs.target.  Either way an optional
// skip flag per output index.  Returns 0 for a skipped query (no output), 1
// for one without a target (outside), 2 to cast.
template <bool Q4>
__device__ __forceinline__ int load_query(const GridSet &gs, const double *__restrict__ pts,
                                          const uint32_t *__restrict__ order, const uint8_t *__restrict__ skip,
                                          uint32_t q, uint32_t axis, uint32_t &i, uint32_t &m, Ray &r)
{
    if (Q4) {
        const double2 *x = reinterpret_cast<const double2 *>(pts) + 2 * (size_t)q;
        const double2 a = x[0], b = x[1];
        i = order[q];
        if (b.y == kTagSkip || (skip && skip[i]))
            return 0;
        if (b.y == kTagNone)
            return 1;
        m = (uint32_t)b.y;
        r = ray_from({a.x, a.y, b.x}, axis);
        return 2;
    }
    i = order ? order[q] : q;
    if (skip && skip[i])
        return 0;
    m = gs.target ? gs.target[i] : 0u;
    if (m == ~0u)
        return 1;
    r = ray_from({pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]}, axis);
    return 2;
}

// Distinct PositionKeys of one ray, CAP of them in registers.  Along an axis
// ray the hit points share (nearly always) the key's two off-axis components,
// so only the on-axis component is kept per key, as 32 bits.  A key with a
// component outside the 32-bit range (|coordinate| above about 21474), one
// whose off-axis components differ from the first key's, or one more than
// CAP, sets `full` and the ray is finished by the wave path, which compares
// whole 64-bit keys.  New keys enter at slot 0 and shift the others up, so
// every slot index is a compile-time constant and the set never lands in
// scratch memory.
template <int CAP>
struct AxisKeys {
    int a[CAP];
    int bu, bv;
    int n = 0;
    bool full = false;
    __device__ __forceinline__ static bool fits(long long k) { return k >= -2147483647ll - 1 && k <= 2147483647ll; }
    __device__ __forceinline__ void add(const Key3 &key, int axis)
    {
        const long long la = axis == 0 ? key.x : (axis == 1 ? key.y : key.z);
        const long long lu = axis == 0 ? key.y : key.x;
        const long long lv = axis == 2 ? key.y : key.z;
        if (!fits(la) || !fits(lu) || !fits(lv)) {
            full = true;
            return;
        }
        const int ka = (int)la, ku = (int)lu, kv = (int)lv;
        if (n == 0) {
            bu = ku;
            bv = kv;
        } else if (ku != bu || kv != bv) {
            full = true;
            return;
        }
#pragma unroll
        for (int j = 0; j < CAP; ++j)
            if (j < n && a[j] == ka)
                return;
        if (n == CAP) {
            full = true;
            return;
        }
#pragma unroll
        for (int j = CAP - 1; j > 0; --j)
            a[j] = a[j - 1];
        a[0] = ka;
        ++n;
    }
};

// a triangle's 96-B FP64 record: vertices, with the FP64 normal in the pad slots
struct TriRec {
    double2 r0, r1, r2, r3, r4, r5;
};

__device__ __forceinline__ TriRec load_rec(const double *__restrict__ tri64, uint32_t t)
{
    const double2 *x = reinterpret_cast<const double2 *>(tri64 + (size_t)kTri64Stride * t);
    return {x[0], x[1], x[2], x[3], x[4], x[5]};
}

// candidate triangle (its record): FP64 box test, plane hit, three-normal
// inside test; on a hit, `key` is the hit point's PositionKey
__device__ __forceinline__ bool ray_hit_rec(const TriRec &x, const Ray &r, Key3 &key)
{
    const double2 r0 = x.r0, r1 = x.r1, r2 = x.r2, r3 = x.r3, r4 = x.r4, r5 = x.r5;
    const V3 v0 = {r0.x, r0.y, r1.x}, v1 = {r2.x, r2.y, r3.x}, v2 = {r4.x, r4.y, r5.x};
    const V3 e = r.end();
    if (!box64_overlap(tri_box64(v0, v1, v2), Box64{r.p, e}))
        return false;
    const V3 u = vsub(e, r.p);
    const V3 n = {r1.y, r3.y, r5.y};  // the record's pad slots hold the FP64 normal
    const V3 w = vsub(r.p, v0);
    const double d = n.x * u.x + n.y * u.y + n.z * u.z;
    const double num = (-n.x) * w.x + (-n.y) * w.y + (-n.z) * w.z;
    if (__builtin_fabs(d) <= 2.220446049250313080847e-16)
        return false;
    const double s = num / d;
    if (s < 0 || s > 1 || __builtin_isnan(s) || __builtin_isinf(s))
        return false;
    const V3 hit = {r.p.x + s * u.x, r.p.y + s * u.y, r.p.z + s * u.z};
    const V3 n0 = tri_normal(hit, v0, v1);
    const V3 n1 = tri_normal(hit, v1, v2);
    const V3 n2 = tri_normal(hit, v2, v0);
    if (!(vdot(n0, n1) > 0 && vdot(n0, n2) > 0))
        return false;
    key = position_key(hit);
    return true;
}

__device__ __forceinline__ bool ray_hit(uint32_t t, const double *__restrict__ tri64, const Ray &r, Key3 &key)
{
    return ray_hit_rec(load_rec(tri64, t), r, key);
}

// ---- per-axis projection grids of the target meshes ------------------------
// For a ray along `axis`, (u, v) are the other two coordinates.  A target
// mesh's root box is split into G x G cells in (u, v); cell(x) is monotone in
// x, so a triangle filed under every cell its slack-widened (u, v) box spans
// is filed under the cell of every ray origin whose segment box it can
// overlap: the segment covers [p.u, p.u + eps] and the slack
// (1e-15 (1 + |lo|)) exceeds eps plus the rounding of lo - slack.  Several
// meshes (a batch) share one cell array: mesh m owns 3 G_m^2 cells from its
// first cell on.
//
// An entry is two 4-byte code words (one array) and the triangle index (a
// second array, read only for entries that pass the code test).  The codes are
// the triangle's FP64 box bounds lo_u, hi_u, lo_v, hi_v relative to the
// entry's cell (12 bits each, in 1/4096 of a cell, clamped to the cell) and
// hi_a relative to the mesh box along the axis (16 bits, in 1/65536 of its
// extent).  Every code is one monotone function of its coordinate, and a ray's
// own codes come from the same function in the same cell, so lo_u <= e.u
// implies code(lo_u) <= code(e.u) and so on: the code test passes every entry
// whose FP64 box overlaps the ray's segment box, and most others fail it
// without touching the FP64 record.  The resolution matters where faces lie
// nearly along the rays (edge-on, or close to a query face's plane): coarser
// codes pass whole layers of such faces to the FP64 test.
struct GridFrame {
    double lu, lv, la, iu, iv, ia;
};

__device__ __forceinline__ int axis_u(int axis) { return axis == 0 ? 1 : 0; }
__device__ __forceinline__ int axis_v(int axis) { return axis == 2 ? 1 : 2; }

constexpr double kCellSteps = 4096.0;  // u, v codes per cell
constexpr uint32_t kCellMax = 4095u;
constexpr double kAxisSteps = 65536.0;  // axis codes over the mesh extent
constexpr uint32_t kAxisMax = 65535u;

__device__ __forceinline__ GridFrame grid_frame(const float *__restrict__ box, int axis, uint32_t G)
{
    const int u = axis_u(axis), v = axis_v(axis);
    GridFrame f;
    f.lu = (double)box[u];
    f.lv = (double)box[v];
    f.la = (double)box[axis];
    const double su = (double)box[3 + u] - f.lu, sv = (double)box[3 + v] - f.lv;
    const double sa = (double)box[3 + axis] - f.la;
    f.iu = su > 0.0 ? (double)G / su : 0.0;
    f.iv = sv > 0.0 ? (double)G / sv : 0.0;
    f.ia = sa > 0.0 ? kAxisSteps / sa : 0.0;
    return f;
}

// x in cell units from lo (inv >= 0, so monotone in x)
__device__ __forceinline__ double cell_coord(double x, double lo, double inv) { return (x - lo) * inv; }

__device__ __forceinline__ uint32_t cell_of(double t, uint32_t G)
{
    if (!(t >= 0.0))
        return 0;
    return t < (double)G ? (uint32_t)t : G - 1u;
}

__device__ __forceinline__ uint32_t grid_cell(double x, double lo, double inv, uint32_t G)
{
    return cell_of(cell_coord(x, lo, inv), G);
}

// monotone code of y in [0, max] (NaN and negatives give 0)
__device__ __forceinline__ uint32_t code_of(double y, uint32_t max)
{
    if (!(y >= 0.0))
        return 0;
    return y < (double)max ? (uint32_t)y : max;
}

// code of cell coordinate t inside cell c
__device__ __forceinline__ uint32_t cell_code(double t, uint32_t c)
{
    return code_of((t - (double)c) * kCellSteps, kCellMax);
}

// code words: x = lo_u | hi_u << 12 | (hi_a & 255) << 24,
//             y = lo_v | hi_v << 12 | (hi_a >> 8) << 24
__device__ __forceinline__ uint2 pack_code(uint32_t lu, uint32_t hu, uint32_t lv, uint32_t hv, uint32_t ha)
{
    return make_uint2(lu | hu << 12 | (ha & 255u) << 24, lv | hv << 12 | (ha >> 8) << 24);
}

__device__ __forceinline__ double comp(const V3 &a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// (G, first cell) of mesh m
__device__ __forceinline__ uint2 grid_desc(const GridSet &gs, uint32_t m)
{
    return gs.desc ? gs.desc[m] : make_uint2(gs.G0, 0u);
}

struct GridSpan {
    uint32_t u0, u1, v0, v1;  // cells spanned (slack-widened low ends)
    double lu, hu, lv, hv;    // the FP64 box bounds in cell coordinates
    uint32_t ha;              // code of the box's upper bound along the axis
};

__device__ __forceinline__ GridSpan grid_span(const Box64 &b, int axis, const GridFrame &f, uint32_t G)
{
    const int u = axis_u(axis), v = axis_v(axis);
    const double lu = comp(b.mn, u), hu = comp(b.mx, u), lv = comp(b.mn, v), hv = comp(b.mx, v);
    GridSpan s;
    s.u0 = grid_cell(lu - 1e-15 * (1.0 + __builtin_fabs(lu)), f.lu, f.iu, G);
    s.v0 = grid_cell(lv - 1e-15 * (1.0 + __builtin_fabs(lv)), f.lv, f.iv, G);
    s.lu = cell_coord(lu, f.lu, f.iu);
    s.hu = cell_coord(hu, f.lu, f.iu);
    s.lv = cell_coord(lv, f.lv, f.iv);
    s.hv = cell_coord(hv, f.lv, f.iv);
    s.u1 = cell_of(s.hu, G);
    s.v1 = cell_of(s.hv, G);
    s.ha = code_of(cell_coord(comp(b.mx, axis), f.la, f.ia), kAxisMax);
    return s;
}

// the entry of span s in cell (cu, cv)
__device__ __forceinline__ uint2 span_code(const GridSpan &s, uint32_t cu, uint32_t cv)
{
    return pack_code(cell_code(s.lu, cu), cell_code(s.hu, cu), cell_code(s.lv, cv), cell_code(s.hv, cv), s.ha);
}

// Grid build: a workgroup takes tiles of T consecutive triangles
// of the visiting order (Morton order, so they are neighbours), files each under its
// cells on all three axes (one record load per triangle), and first counts its
// entries per cell in an LDS hash table; only then does it touch the global
// cell counters, once per distinct cell.  Device-scope atomics are the cost
// here (they are coherent across the chip's separate L2s), and neighbouring
// triangles share most cells, so this removes most of them.  A cell that finds
// no slot within kProbe probes falls back to per-entry global atomics; both
// paths only ever advance a cell's counter, so they mix freely.
// A build capped beside the single-pair search takes tiles of 512 triangles
// with an LDS table of 2,048 slots (less LDS held beside that search's
// kernels); the batch's uncapped build takes tiles of 1,024 with 8,192 slots
// (fewer cells spilling to global atomics).  One triangle per lane.
constexpr int kGridSlotBitsBeside = 11, kGridTrisBeside = 512;
constexpr int kGridSlotBitsWide = 13, kGridTrisWide = 1024;
constexpr int kProbe = 32;

// slot holding `key` in the LDS table (inserting it), or -1.  Slots only fill,
// so a key that once found no slot never finds one later in the same pass.  A
// plain read comes first: most entries find their cell already there, and
// reads of one slot by many lanes do not serialise the way atomics do.
template <int B>
__device__ __forceinline__ int table_slot(uint32_t *keys, uint32_t key)
{
    uint32_t h = (key * 2654435761u) >> (32 - B);
    for (int probe = 0; probe < kProbe; ++probe) {
        uint32_t cur = __hip_atomic_load(&keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == 0u)
            cur = atomicCAS(&keys[h], 0u, key);
        if (cur == 0u || cur == key)
            return (int)h;
        h = (h + 1) & ((1u << B) - 1);
    }
    return -1;
}

// slot of `key` in a table that no longer changes, or -1 when it is not there
template <int B>
__device__ __forceinline__ int table_find(const uint32_t *keys, uint32_t key)
{
    uint32_t h = (key * 2654435761u) >> (32 - B);
    for (int probe = 0; probe < kProbe; ++probe) {
        const uint32_t cur = keys[h];
        if (cur == key)
            return (int)h;
        if (cur == 0u)
            return -1;
        h = (h + 1) & ((1u << B) - 1);
    }
    return -1;
}

// one triangle's spans on the three axes
struct TriSpans {
    GridSpan s[3];
    uint32_t t, G;
    size_t base;  // counter index of the mesh's first cell, + 1
    bool valid;
};

// The spans come from the triangle's FP32 leaf box, rounded outward from its
// FP64 box: it contains the FP64 box, and every cell and code function is
// monotone, so the cells it is filed under and its codes cover what the FP64
// box would give (the ray's code test stays a superset of the FP64 box test),
// for a 32-B read instead of the 96-B FP64 record.
__device__ __forceinline__ TriSpans tri_spans(const float *__restrict__ box32, uint32_t nt, const GridSet &gs,
                                              uint32_t pos)
{
    TriSpans x;
    x.valid = pos < nt;
    if (!x.valid)
        return x;
    x.t = gs.order ? gs.order[pos] : pos;
    const uint32_t m = gs.meshOf ? gs.meshOf[x.t] : 0u;
    const uint2 d = grid_desc(gs, m);
    x.G = d.x;
    x.base = (size_t)d.y + 1;
    const float4 *r = reinterpret_cast<const float4 *>(box32 + (size_t)kBox32Stride * x.t);
    const float4 lo = r[0], hi = r[1];
    const Box64 b = {{lo.x, lo.y, lo.z}, {hi.x, hi.y, hi.z}};
    const float *bx = gs.boxes + 6 * (size_t)m;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis)
        x.s[axis] = grid_span(b, axis, grid_frame(bx, axis, x.G), x.G);
    return x;
}

__device__ __forceinline__ uint32_t cell_key(const TriSpans &x, int axis, uint32_t cu, uint32_t cv)
{
    return (uint32_t)(x.base + (size_t)axis * x.G * x.G + (size_t)cu * x.G + cv);
}

// The tiles of T triangles a workgroup of the grid kernels takes: a
// contiguous run, and runs in XCD-aware order, so neighbouring tiles (which
// share cells) stay in one L2.  The launch may cap its workgroups: a build
// beside the latency-bound pair search leaves that search's small launches
// free slots on every CU.
template <int T>
__device__ __forceinline__ uint2 grid_tiles(uint32_t nt)
{
    const uint32_t ntiles = (nt + T - 1) / T;
    const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    const uint32_t first = xcd_block(blockIdx.x, gridDim.x) * per;
    return make_uint2(first < ntiles ? first : ntiles, first + per < ntiles ? first + per : ntiles);
}

// cells: ncells + 1 u32, zero on entry; entries of cell c are counted in cells[c + 1]
template <int B, int T>
__global__ __launch_bounds__(T) void k_grid_count(const float *__restrict__ box32, uint32_t nt,
                                                               GridSet gs, uint32_t *__restrict__ cells)
{
    constexpr int kSlots = 1 << B;
    __shared__ uint32_t s_key[kSlots], s_cnt[kSlots];
    const uint2 tiles = grid_tiles<T>(nt);
    for (uint32_t tile = tiles.x; tile < tiles.y; ++tile) {
        for (int k = threadIdx.x; k < kSlots; k += T) {
            s_key[k] = 0;
            s_cnt[k] = 0;
        }
        __syncthreads();
        const TriSpans x = tri_spans(box32, nt, gs, tile * T + threadIdx.x);
        if (x.valid) {
#pragma unroll
            for (int axis = 0; axis < 3; ++axis)
                for (uint32_t cu = x.s[axis].u0; cu <= x.s[axis].u1; ++cu)
                    for (uint32_t cv = x.s[axis].v0; cv <= x.s[axis].v1; ++cv) {
                        const uint32_t key = cell_key(x, axis, cu, cv);
                        const int sl = table_slot<B>(s_key, key);
                        if (sl >= 0)
                            atomicAdd(&s_cnt[sl], 1u);
                        else
                            atomicAdd(&cells[key], 1u);
                    }
        }
        __syncthreads();
        for (int k = threadIdx.x; k < kSlots; k += T)
            if (s_key[k])
                atomicAdd(&cells[s_key[k]], s_cnt[k]);
        __syncthreads();  // the table is cleared for the next tile
    }
}

// After the exclusive scan, cells[c + 1] is the start of cell c; per tile the
// workgroup reserves one range per distinct cell and hands out slots from
// LDS, so afterwards cell c spans [cells[c], cells[c + 1]).
template <int B, int T>
__global__ __launch_bounds__(T) void k_grid_fill(const float *__restrict__ box32, uint32_t nt,
                                                              GridSet gs, uint32_t *__restrict__ cells,
                                                              uint2 *__restrict__ gcode, uint32_t *__restrict__ gtri)
{
    constexpr int kSlots = 1 << B;
    __shared__ uint32_t s_key[kSlots], s_cnt[kSlots];
    const uint2 tiles = grid_tiles<T>(nt);
    for (uint32_t tile = tiles.x; tile < tiles.y; ++tile) {
        for (int k = threadIdx.x; k < kSlots; k += T) {
            s_key[k] = 0;
            s_cnt[k] = 0;
        }
        __syncthreads();
        const TriSpans x = tri_spans(box32, nt, gs, tile * T + threadIdx.x);
        if (x.valid) {
#pragma unroll
            for (int axis = 0; axis < 3; ++axis)
                for (uint32_t cu = x.s[axis].u0; cu <= x.s[axis].u1; ++cu)
                    for (uint32_t cv = x.s[axis].v0; cv <= x.s[axis].v1; ++cv) {
                        const int sl = table_slot<B>(s_key, cell_key(x, axis, cu, cv));
                        if (sl >= 0)
                            atomicAdd(&s_cnt[sl], 1u);
                    }
        }
        __syncthreads();
        for (int k = threadIdx.x; k < kSlots; k += T)
            if (s_key[k])
                s_cnt[k] = atomicAdd(&cells[s_key[k]], s_cnt[k]);  // the range's first slot
        __syncthreads();
        if (x.valid) {
#pragma unroll
            for (int axis = 0; axis < 3; ++axis)
                for (uint32_t cu = x.s[axis].u0; cu <= x.s[axis].u1; ++cu)
                    for (uint32_t cv = x.s[axis].v0; cv <= x.s[axis].v1; ++cv) {
                        const uint32_t key = cell_key(x, axis, cu, cv);
                        const int sl = table_find<B>(s_key, key);
                        const uint32_t slot = sl >= 0 ? atomicAdd(&s_cnt[sl], 1u) : atomicAdd(&cells[key], 1u);
                        if (slot < gs.cap) {
                            gcode[slot] = span_code(x.s[axis], cu, cv);
                            gtri[slot] = x.t;
                        }
                    }
        }
        __syncthreads();  // the table is cleared for the next tile
    }
}

// A ray's cell in its target mesh: the candidate range and the ray's codes
// (lo ends from p, hi ends from e) in that cell.  A ray whose FP32 outward box
// misses the mesh's FP32 outward root box (which contains every triangle's
// FP64 box) has no candidate at all: without this, rays from outside the
// grid's square would be clamped into border cells, which hold every face
// seen edge-on.
struct RayCell {
    uint2 range;
    uint32_t u, v;  // code of p | code of e << 16, along u and along v
    uint32_t pa;    // code of p along the axis
};

__device__ __forceinline__ RayCell ray_cell(const GridSet &gs, uint32_t m, const Ray &r, int axis)
{
    RayCell c;
    c.range = make_uint2(0u, 0u);
    const float *bx = gs.boxes + 6 * (size_t)m;
    const V3 e = r.end();
    if (!(bx[0] <= f_up(e.x) && bx[3] >= f_down(r.p.x) && bx[1] <= f_up(e.y) && bx[4] >= f_down(r.p.y) &&
          bx[2] <= f_up(e.z) && bx[5] >= f_down(r.p.z)))
        return c;
    const uint2 d = grid_desc(gs, m);
    const uint32_t G = d.x;
    const GridFrame f = grid_frame(bx, axis, G);
    const int u = axis_u(axis), v = axis_v(axis);
    const double tu = cell_coord(comp(r.p, u), f.lu, f.iu), tv = cell_coord(comp(r.p, v), f.lv, f.iv);
    const uint32_t cu = cell_of(tu, G), cv = cell_of(tv, G);
    c.u = cell_code(tu, cu) | cell_code(cell_coord(comp(e, u), f.lu, f.iu), cu) << 16;
    c.v = cell_code(tv, cv) | cell_code(cell_coord(comp(e, v), f.lv, f.iv), cv) << 16;
    c.pa = code_of(cell_coord(comp(r.p, axis), f.la, f.ia), kAxisMax);
    const size_t k = (size_t)d.y + (size_t)axis * G * G + (size_t)cu * G + cv;
    c.range = make_uint2(gs.cells[k], gs.cells[k + 1]);
    return c;
}

// code test of an entry against the ray (a superset of the FP64 box test)
__device__ __forceinline__ bool code_may_hit(const uint2 &e, const RayCell &c)
{
    const uint32_t ha = e.x >> 24 | (e.y >> 24) << 8;
    return (e.x & kCellMax) <= (c.u >> 16) && ((e.x >> 12) & kCellMax) >= (c.u & 0xffffu) &&
           (e.y & kCellMax) <= (c.v >> 16) && ((e.y >> 12) & kCellMax) >= (c.v & 0xffffu) && ha >= c.pa;
}

// One lane per (axis, query).  A lane whose cell holds
// at most gs.fastCands entries (48 when unset) scans them with up to
// kFastKeys distinct keys in registers; any other lane (a crowded cell, or
// more distinct hits) hands (axis, query) to k_rays_wave.  Query i's rays go
// to mesh target[i] (mesh 0 without a target list); a query without a target
// mesh (~0u) counts as outside.
//
// The wave's lanes trace rays from neighbouring points, so they share few
// cells.  The wave first stages each distinct cell's entries (codes and
// triangles) in LDS with coalesced loads, once per cell, up to kStage
// entries; a lane then reads its cell from LDS (a cell that did not fit is
// read from global memory as before).
//
// FILTER: each candidate passing the code test is decided from its 48-B FP32
// record by the certificate of sbg_rayfilter.h (a certain miss, or a certain
// hit at a distance known to within a bound) instead of the 96-B FP64 record
// and the FP64 test.  Certain hits farther apart than PositionKey's cell are
// distinct keys; a ray with an uncertain candidate, or with two hits that
// close, goes to k_rays_wave, which runs it in FP64 like any deferred ray.
constexpr int kStage = 384;  // cell entries a wave stages in LDS
constexpr int kStageRounds = kStage / 64;

// Code test of entries [k0, k0 + 64) of a cell of len entries, as a bit mask,
// in groups of kCodeGroup independent loads (indices clamped into the cell,
// the excess masked off).  The whole chunk is tested before any candidate's
// record is loaded, so a lane's record loads are not held behind its code
// loads; the verdict does not depend on the order the hits arrive in.
template <typename CodeAt>
__device__ __forceinline__ unsigned long long pass_mask(const CodeAt &code_at, const RayCell &c, uint32_t k0,
                                                        uint32_t len)
{
    const uint32_t kn = len - k0 < 64u ? len - k0 : 64u;
    unsigned long long pass = 0;
    for (uint32_t j0 = 0; j0 < kn; j0 += kCodeGroup) {
#pragma unroll
        for (int j = 0; j < kCodeGroup; ++j) {
            const uint32_t jj = j0 + j < kn ? j0 + j : kn - 1;
            if (code_may_hit(code_at(k0 + jj), c) && j0 + j < kn)
                pass |= 1ull << (j0 + j);
        }
    }
    return pass;
}

template <bool Q4, bool LIST, bool FILTER>
__global__ __launch_bounds__(256) void k_rays_grid(GridSet gs, const double *__restrict__ tri64,
                                                   const double *__restrict__ pts,
                                                   const uint32_t *__restrict__ order, const uint8_t *__restrict__ skip,
                                                   uint32_t n, uint32_t lanes, const uint32_t *__restrict__ laneList,
                                                   const uint32_t *__restrict__ laneCount,
                                                   uint8_t *__restrict__ bits, uint32_t *__restrict__ ovfList,
                                                   uint32_t *ovfCount, uint32_t *overflow)
{
    // each of the workgroup's four waves stages into its own LDS region
    __shared__ uint2 s_codeW[4][kStage];
    __shared__ uint32_t s_triW[4][kStage];
    __shared__ uint32_t s_baseW[4][64], s_startW[4][64];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint2 *s_code = s_codeW[wave];
    uint32_t *s_tri = s_triW[wave], *s_base = s_baseW[wave], *s_start = s_startW[wave];
    const uint32_t total = LIST ? *laneCount : lanes;
    // every exit before the staging below is taken by the whole workgroup
    if (blockIdx.x * 256 >= total)
        return;
    if (gs.cells[gs.ncells] > gs.cap) {  // grid truncated: the host rebuilds it larger
        if (threadIdx.x == 0)
            atomicOr(overflow, 2u);
        return;
    }
    const uint32_t idx = blockIdx.x * 256 + threadIdx.x;
    uint32_t g = 0, axis = 0, i = 0, m = 0;
    Ray r;
    int kind = 0;
    if (idx < total) {
        g = LIST ? laneList[idx] : idx;
        axis = g / n;
        const uint32_t q = g - axis * n;
        kind = load_query<Q4>(gs, pts, order, skip, q, axis, i, m, r);
        if (kind == 1)
            bits[(size_t)axis * n + i] = 0;
    }
    RayCell c;
    c.range = make_uint2(0u, 0u);
    c.u = c.v = c.pa = 0u;
    if (kind == 2)
        c = ray_cell(gs, m, r, (int)axis);
    const uint32_t len = c.range.y - c.range.x;
    bool defer = kind == 2 && len > (gs.fastCands ? gs.fastCands : (uint32_t)kFastCands);
    const bool scan = kind == 2 && !defer && len > 0;

    // distinct cells of the scanning lanes, in order of their first lane:
    // slot table (LDS offset, first entry), and each lane's LDS offset
    int myBase = -1;
    uint32_t nslots = 0, used = 0;
    for (unsigned long long pending = __ballot(scan); pending;) {
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t lx = __shfl(c.range.x, leader), llen = __shfl(len, leader);
        const bool mine = scan && c.range.x == lx;  // non-empty cells never share a first entry
        pending &= ~__ballot(mine);
        if (used + llen > (uint32_t)kStage)
            continue;
        if (mine)
            myBase = (int)used;
        if (lane == 0) {
            s_base[nslots] = used;
            s_start[nslots] = lx;
        }
        ++nslots;
        used += llen;
    }
    __syncthreads();  // the slot tables (every wave of the workgroup passes both barriers)
    if (nslots) {
        uint2 code[kStageRounds];
        uint32_t tri[kStageRounds];
#pragma unroll
        for (int rr = 0; rr < kStageRounds; ++rr) {
            const uint32_t t = lane + 64u * rr;
            code[rr] = make_uint2(0u, 0u);
            tri[rr] = 0u;
            if (t < used) {
                // the last slot starting at or before t
                uint32_t lo = 0, hi = nslots - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (s_base[mid] <= t)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                const uint32_t e = s_start[lo] + (t - s_base[lo]);
                code[rr] = gs.gcode[e];
                tri[rr] = gs.gtri[e];
            }
        }
#pragma unroll
        for (int rr = 0; rr < kStageRounds; ++rr) {
            const uint32_t t = lane + 64u * rr;
            if (t < used) {
                s_code[t] = code[rr];
                s_tri[t] = tri[rr];
            }
        }
    }
    __syncthreads();  // the staged entries
    if (kind != 2)
        return;
    // entry k (0 <= k < len) of this lane's cell.  The per-lane choice
    // between LDS and global memory compiles to flat loads, one instruction
    // for a wave that mixes staged and unstaged cells; separate LDS and
    // global paths (divergent, both run) made the kernel 10 % slower.
    auto code_at = [&](uint32_t k) -> uint2 { return myBase >= 0 ? s_code[myBase + k] : gs.gcode[c.range.x + k]; };
    auto tri_at = [&](uint32_t k) -> uint32_t { return myBase >= 0 ? s_tri[myBase + k] : gs.gtri[c.range.x + k]; };
    uint32_t parity = 0;
    if (FILTER) {
        const float4 p32 = make_float4((float)r.p.x, (float)r.p.y, (float)r.p.z, 0.0f);
        const double pmax = fmax(fmax(__builtin_fabs(r.p.x), __builtin_fabs(r.p.y)), __builtin_fabs(r.p.z));
        const float pm = f_up(pmax);
        HitSpans<kFastKeys> hits;
        for (uint32_t k0 = 0; !defer && k0 < len; k0 += 64) {
            unsigned long long pass = pass_mask(code_at, c, k0, len);
            while (pass && !defer) {
                const uint32_t k = k0 + (uint32_t)__builtin_ctzll(pass);
                pass &= pass - 1;
                const float4 *rec = reinterpret_cast<const float4 *>(gs.tri32 + (size_t)kTri32Stride * tri_at(k));
                const float4 rr[3] = {rec[0], rec[1], rec[2]};
                const RayCert cert = ray_certificate(p32, pm, (int)axis, rr);
                if (cert.verdict == kRayUnsure) {
                    defer = true;
                } else if (cert.verdict == kRayHit) {
                    hits.add(cert.t, cert.et);
                    defer = hits.full;
                }
            }
        }
        parity = hits.n & 1;
    } else {
        AxisKeys<kFastKeys> keys;
        for (uint32_t k0 = 0; !defer && k0 < len; k0 += 64) {
            unsigned long long pass = pass_mask(code_at, c, k0, len);
            while (pass && !defer) {
                const uint32_t k = k0 + (uint32_t)__builtin_ctzll(pass);
                pass &= pass - 1;
                Key3 key;
                if (ray_hit(tri_at(k), tri64, r, key)) {
                    keys.add(key, (int)axis);
                    defer = keys.full;
                }
            }
        }
        parity = keys.n & 1;
    }
    if (!defer) {
        bits[(size_t)axis * n + i] = (uint8_t)parity;
        return;
    }
    // slot < lanes of this launch: every lane appends at most once.  The
    // record carries what the wave path would otherwise recompute.
    const uint32_t slot = atomicAdd(ovfCount, 1u);
    uint4 *rec = reinterpret_cast<uint4 *>(ovfList) + 2 * (size_t)slot;
    rec[0] = make_uint4(g, c.range.x, c.range.y, i);
    rec[1] = make_uint4(c.u, c.v, c.pa, 0u);
}

// One wave per deferred ray (record of 8 u32 written by k_rays_grid: lane
// g = axis * n + q, cell range, output index, the ray's codes): the 64 lanes split
// the cell's entries, hit keys are gathered in LDS (up to kWaveHits) and the
// distinct ones counted; more hits than that raise `overflow`.
template <bool Q4>
__global__ __launch_bounds__(64) void k_rays_wave(GridSet gs, const double *__restrict__ tri64,
                                                  const double *__restrict__ pts,
                                                  const uint32_t *__restrict__ order, uint32_t n,
                                                  uint8_t *__restrict__ bits, const uint32_t *__restrict__ list,
                                                  const uint32_t *__restrict__ count, uint32_t *overflow)
{
    __shared__ long long s_key[3][kWaveHits];
    const int lane = threadIdx.x;
    const unsigned long long below = (1ull << lane) - 1ull;
    const uint32_t total = *count;
    for (uint32_t k = blockIdx.x; k < total; k += gridDim.x) {
        // the fast path's record of the ray: lane, cell range, output index, codes
        const uint4 *rec = reinterpret_cast<const uint4 *>(list) + 2 * (size_t)k;
        const uint4 r0 = rec[0], r1 = rec[1];
        const uint32_t g = r0.x, i = r0.w;
        const uint32_t axis = g / n;
        const uint32_t q = g - axis * n;
        RayCell c;
        c.range = make_uint2(r0.y, r0.z);
        c.u = r1.x;
        c.v = r1.y;
        c.pa = r1.z;
        Ray r;
        if (Q4) {
            const double2 *x = reinterpret_cast<const double2 *>(pts) + 2 * (size_t)q;
            const double2 a = x[0], b = x[1];
            r = ray_from({a.x, a.y, b.x}, axis);
        } else {
            r = ray_from({pts[3 * (size_t)i], pts[3 * (size_t)i + 1], pts[3 * (size_t)i + 2]}, axis);
        }
        // kWaveLoads entries per lane per round, each stage's loads issued
        // together: codes, then the triangles passing, then their records
        uint32_t nh = 0;
        for (uint32_t k0 = c.range.x; k0 < c.range.y; k0 += 64 * kWaveLoads) {
            bool pass[kWaveLoads];
#pragma unroll
            for (int j = 0; j < kWaveLoads; ++j) {
                const uint32_t e = k0 + lane + 64 * j;
                pass[j] = e < c.range.y && code_may_hit(gs.gcode[e], c);
            }
            uint32_t t[kWaveLoads];
#pragma unroll
            for (int j = 0; j < kWaveLoads; ++j)
                t[j] = pass[j] ? gs.gtri[k0 + lane + 64 * j] : 0u;
            TriRec rec[kWaveLoads];
#pragma unroll
            for (int j = 0; j < kWaveLoads; ++j)
                if (pass[j])
                    rec[j] = load_rec(tri64, t[j]);
#pragma unroll
            for (int j = 0; j < kWaveLoads; ++j) {
                Key3 key;
                const bool hit = pass[j] && ray_hit_rec(rec[j], r, key);
                const unsigned long long mask = __ballot(hit);
                const uint32_t pos = nh + (uint32_t)__popcll(mask & below);
                if (hit && pos < (uint32_t)kWaveHits) {
                    s_key[0][pos] = key.x;
                    s_key[1][pos] = key.y;
                    s_key[2][pos] = key.z;
                }
                nh += (uint32_t)__popcll(mask);
            }
        }
        __syncthreads();
        if (nh > (uint32_t)kWaveHits) {
            if (lane == 0)
                atomicOr(overflow, 1u);
        } else {
            // a key counts where it first occurs
            uint32_t distinct = 0;
            for (uint32_t a = 0; a < nh; a += 64) {
                const uint32_t x = a + lane;
                bool first = x < nh;
                for (uint32_t b = 0; first && b < x; ++b)
                    if (s_key[0][b] == s_key[0][x] && s_key[1][b] == s_key[1][x] && s_key[2][b] == s_key[2][x])
                        first = false;
                distinct += (uint32_t)__popcll(__ballot(first));
            }
            if (lane == 0)
                bits[(size_t)axis * n + i] = (uint8_t)(distinct & 1);
        }
        __syncthreads();
    }
}

// The vote needs the z ray only where the x and y rays disagree: lists the
// lanes 2n + q of those queries (one atomic per wave).
template <bool Q4>
__global__ __launch_bounds__(256) void k_third_axis(const double *__restrict__ pts,
                                                    const uint32_t *__restrict__ order,
                                                    const uint8_t *__restrict__ skip, uint32_t n,
                                                    const uint8_t *__restrict__ bits, uint32_t *__restrict__ list,
                                                    uint32_t *count)
{
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    bool need = false;
    if (q < n) {
        const uint32_t i = order ? order[q] : q;
        const bool skipped = (Q4 && reinterpret_cast<const double2 *>(pts)[2 * (size_t)q + 1].y == kTagSkip) ||
                             (skip && skip[i]);
        need = !skipped && bits[i] != bits[(size_t)n + i];
    }
    const unsigned long long mask = __ballot(need);
    if (!mask)
        return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader)
        base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (need)
        list[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = 2 * n + q;
}

// Work of one k_rays_grid launch over `lanes` lanes, for measurement:
// work[0] rays cast, [1] grid entries scanned by the fast path, [2] entries
// passing the code test there, [3] rays deferred to the wave path.
template <bool Q4>
__global__ __launch_bounds__(256) void k_ray_work(GridSet gs, const double *__restrict__ pts,
                                                  const uint32_t *__restrict__ order,
                                                  const uint8_t *__restrict__ skip, uint32_t n, uint32_t lanes,
                                                  unsigned long long *work)
{
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    unsigned long long cast = 0, scanned = 0, passed = 0, deferred = 0;
    if (g < lanes) {
        const uint32_t axis = g / n, q = g - axis * n;
        uint32_t i, m = 0;
        Ray r;
        if (load_query<Q4>(gs, pts, order, skip, q, axis, i, m, r) == 2) {
            cast = 1;
            const RayCell c = ray_cell(gs, m, r, (int)axis);
            if (c.range.y - c.range.x > (gs.fastCands ? gs.fastCands : (uint32_t)kFastCands)) {
                deferred = 1;
            } else {
                scanned = c.range.y - c.range.x;
                for (uint32_t k = c.range.x; k < c.range.y; ++k)
                    passed += code_may_hit(gs.gcode[k], c) ? 1 : 0;
            }
        }
    }
    unsigned long long v[4] = {cast, scanned, passed, deferred};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        unsigned long long x = v[j];
        for (int off = 32; off > 0; off >>= 1)
            x += __shfl_down(x, off);
        if ((threadIdx.x & 63) == 0 && x)
            atomicAdd(work + j, x);
    }
}

// decideGroupSide's vote (src/solidboolean.cpp:492-508): inside iff
// (float)inside / 3 > 0.5, i.e. at least 2 of the 3 rays.  The z bit is read
// only where the x and y bits differ (it is not cast elsewhere), which gives
// the same majority.
__global__ __launch_bounds__(256) void k_vote(const uint8_t *__restrict__ bits, const uint8_t *__restrict__ skip,
                                              uint32_t n, uint8_t *__restrict__ out)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    if (skip && skip[i]) {
        out[i] = 2;
        return;
    }
    const int bx = bits[i], by = bits[n + i];
    const int inside = bx == by ? 2 * bx : bx + by + bits[2 * (size_t)n + i];
    out[i] = ((float)inside / 3.0f > 0.5f) ? 1 : 0;
}

void launch_mark_intersected(const uint64_t *pairs, const uint8_t *status, uint32_t n, int side,
                             uint8_t *flags, hipStream_t s)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_mark_intersected, dim3((n + 255) / 256), dim3(256), 0, s, pairs, status, n, side, flags);
}

void launch_centroids(const double *tri64, const uint8_t *flags, uint32_t nt, double *pts, hipStream_t s)
{
    if (nt == 0)
        return;
    hipLaunchKernelGGL(k_centroids, dim3((nt + 255) / 256), dim3(256), 0, s, tri64, flags, nt, pts);
}

// cells per axis side: G^2 = nt * density / 16, G in [1, 1024]
uint32_t grid_dim(uint32_t nt, int density)
{
    uint32_t G = (uint32_t)std::ceil(std::sqrt((double)nt * density / 16.0));
    return G < 1u ? 1u : (G > 1024u ? 1024u : G);
}

size_t grid_cells(uint32_t G) { return 3 * (size_t)G * G + 1; }


template <int B, int T>
static void grid_build(const float *box32, uint32_t nt, const GridSet &gs, uint32_t *cells, uint32_t *scanTemp,
                       uint2 *gcode, uint32_t *gtri, uint32_t maxBlocks, hipStream_t s)
{
    const uint32_t ntiles = (nt + T - 1) / T;
    const dim3 grid(xcd_grid(maxBlocks && ntiles > maxBlocks ? maxBlocks : ntiles));
    hipLaunchKernelGGL((k_grid_count<B, T>), grid, dim3(T), 0, s, box32, nt, gs, cells);
    launch_scan_large(cells, (size_t)gs.ncells + 1, scanTemp, s);
    hipLaunchKernelGGL((k_grid_fill<B, T>), grid, dim3(T), 0, s, box32, nt, gs, cells, gcode, gtri);
}

void launch_grid_build(const float *box32, uint32_t nt, const GridSet &gs, uint32_t *cells, uint32_t *scanTemp,
                       uint2 *gcode, uint32_t *gtri, uint32_t maxBlocks, hipStream_t s)
{
    const size_t m = (size_t)gs.ncells + 1;
    launch_fill_u32(cells, 0u, m, s);
    if (nt == 0)
        return;
    if (maxBlocks)
        grid_build<kGridSlotBitsBeside, kGridTrisBeside>(box32, nt, gs, cells, scanTemp, gcode, gtri, maxBlocks, s);
    else
        grid_build<kGridSlotBitsWide, kGridTrisWide>(box32, nt, gs, cells, scanTemp, gcode, gtri, 0, s);
}

// Rays of n queries.  allAxes: every query casts x, y and z; otherwise x and
// y, then z only where they disagree (enough for the vote; the per-axis bits
// of z are then valid only there).  lists: ray_list_u32(n) u32 (deferred-ray
// records of 8 u32 for up to 3n rays, then n z lanes); counters: 4 u32, zero
// on entry ([0] overflow bits, [1] deferred x/y rays, [2] z lanes, [3]
// deferred z rays).
template <bool Q4>
static void rays(const GridSet &gs, const double *tri64, const double *pts, const uint32_t *order,
                 const uint8_t *skip, uint32_t n, bool allAxes, uint8_t *axisBits, uint32_t *lists,
                 uint32_t *counters, hipStream_t s)
{
    if (n == 0)
        return;
    uint32_t *overflow = counters;
    const uint32_t lanes = allAxes ? 3u * n : 2u * n;
    uint32_t *deferred = lists, *zDeferred = lists + 8 * 2 * (size_t)n, *zLanes = lists + 8 * 3 * (size_t)n;
    if (gs.tri32)
        hipLaunchKernelGGL((k_rays_grid<Q4, false, true>), dim3((lanes + 255) / 256), dim3(256), 0, s, gs, tri64, pts,
                           order, skip, n, lanes, nullptr, nullptr, axisBits, deferred, counters + 1, overflow);
    else
        hipLaunchKernelGGL((k_rays_grid<Q4, false, false>), dim3((lanes + 255) / 256), dim3(256), 0, s, gs, tri64, pts,
                           order, skip, n, lanes, nullptr, nullptr, axisBits, deferred, counters + 1, overflow);
    hipLaunchKernelGGL(k_rays_wave<Q4>, dim3(8192), dim3(64), 0, s, gs, tri64, pts, order, n, axisBits, deferred,
                       counters + 1, overflow);
    if (allAxes)
        return;
    hipLaunchKernelGGL(k_third_axis<Q4>, dim3((n + 255) / 256), dim3(256), 0, s, pts, order, skip, n, axisBits,
                       zLanes, counters + 2);
    if (gs.tri32)
        hipLaunchKernelGGL((k_rays_grid<Q4, true, true>), dim3((n + 255) / 256), dim3(256), 0, s, gs, tri64, pts, order,
                           skip, n, n, zLanes, counters + 2, axisBits, zDeferred, counters + 3, overflow);
    else
        hipLaunchKernelGGL((k_rays_grid<Q4, true, false>), dim3((n + 255) / 256), dim3(256), 0, s, gs, tri64, pts, order,
                           skip, n, n, zLanes, counters + 2, axisBits, zDeferred, counters + 3, overflow);
    hipLaunchKernelGGL(k_rays_wave<Q4>, dim3(8192), dim3(64), 0, s, gs, tri64, pts, order, n, axisBits, zDeferred,
                       counters + 3, overflow);
}

size_t ray_list_u32(size_t n) { return 8 * 3 * n + n; }

void launch_rays(const GridSet &gs, const double *tri64, const double *pts, const uint32_t *order,
                 const uint8_t *skip, uint32_t n, bool allAxes, uint8_t *axisBits, uint32_t *lists,
                 uint32_t *counters, hipStream_t s)
{
    rays<false>(gs, tri64, pts, order, skip, n, allAxes, axisBits, lists, counters, s);
}

void launch_centroids_ordered(const double *tri64, const uint8_t *flags, const uint32_t *order,
                              const uint32_t *target, uint32_t nt, double *q4, hipStream_t s)
{
    if (nt == 0)
        return;
    hipLaunchKernelGGL(k_centroids_ordered, dim3((nt + 255) / 256), dim3(256), 0, s, tri64, flags, order, target,
                       nt, q4);
}

void launch_rays_xy_ordered(const GridSet &gs, const double *tri64, const double *q4, const uint32_t *order,
                            uint32_t n, uint8_t *axisBits, uint32_t *lists, uint32_t *counters, hipStream_t s)
{
    if (n == 0)
        return;
    const uint32_t lanes = 2u * n;
    if (gs.tri32)
        hipLaunchKernelGGL((k_rays_grid<true, false, true>), dim3((lanes + 255) / 256), dim3(256), 0, s, gs, tri64, q4,
                           order, nullptr, n, lanes, nullptr, nullptr, axisBits, lists, counters + 1, counters);
    else
        hipLaunchKernelGGL((k_rays_grid<true, false, false>), dim3((lanes + 255) / 256), dim3(256), 0, s, gs, tri64, q4,
                           order, nullptr, n, lanes, nullptr, nullptr, axisBits, lists, counters + 1, counters);
}

void launch_rays_wave_ordered(const GridSet &gs, const double *tri64, const double *q4, const uint32_t *order,
                              uint32_t n, uint8_t *axisBits, const uint32_t *lists, uint32_t *counters, hipStream_t s)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_rays_wave<true>, dim3(8192), dim3(64), 0, s, gs, tri64, q4, order, n, axisBits, lists,
                       counters + 1, counters);
}

void launch_ray_work_ordered(const GridSet &gs, const double *q4, const uint32_t *order, uint32_t n,
                             unsigned long long *work, hipStream_t s)
{
    if (n == 0)
        return;
    const uint32_t lanes = 2u * n;
    hipLaunchKernelGGL(k_ray_work<true>, dim3((lanes + 255) / 256), dim3(256), 0, s, gs, q4, order, nullptr, n, lanes,
                       work);
}

void launch_rays_ordered(const GridSet &gs, const double *tri64, const double *q4, const uint32_t *order,
                         const uint8_t *skip, uint32_t n, bool allAxes, uint8_t *axisBits, uint32_t *lists,
                         uint32_t *counters, hipStream_t s)
{
    rays<true>(gs, tri64, q4, order, skip, n, allAxes, axisBits, lists, counters, s);
}

// per-axis bits [axis][n] -> per-point triples [n][3]
__global__ __launch_bounds__(256) void k_interleave_axes(const uint8_t *__restrict__ bits, uint32_t n,
                                                         uint8_t *__restrict__ out)
{
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    out[3 * (size_t)i] = bits[i];
    out[3 * (size_t)i + 1] = bits[(size_t)n + i];
    out[3 * (size_t)i + 2] = bits[2 * (size_t)n + i];
}

void launch_interleave_axes(const uint8_t *axisBits, uint32_t n, uint8_t *out, hipStream_t s)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_interleave_axes, dim3((n + 255) / 256), dim3(256), 0, s, axisBits, n, out);
}

void launch_vote(const uint8_t *axisBits, const uint8_t *skip, uint32_t n, uint8_t *out, hipStream_t s)
{
    if (n == 0)
        return;
    hipLaunchKernelGGL(k_vote, dim3((n + 255) / 256), dim3(256), 0, s, axisBits, skip, n, out);
}

}  // namespace sbg
