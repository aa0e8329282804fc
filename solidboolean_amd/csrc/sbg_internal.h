// Internal host-side declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace sbg {

constexpr int kMortonBits = 24;  // single-mesh Morton keys (top bits of the 30-bit code)

// ---- device data layouts (HBM) -------------------------------------------
// Triangle record, FP64: 3 vertices x (x, y, z, pad) = 96 B, 16-B aligned.
constexpr int kTri64Stride = 12;  // doubles
// Triangle record, FP32: 3 vertices x (x, y, z, pad) = 48 B.
constexpr int kTri32Stride = 12;  // floats
// Leaf box, FP32 outward-rounded: (minx, miny, minz, 0, maxx, maxy, maxz, 0) = 32 B.
constexpr int kBox32Stride = 8;  // floats

// BVH2 internal node, 64 B: both children's FP32 boxes + child references.
// A child reference with the top bit set is a leaf whose low 31 bits are the
// triangle index; otherwise it is an internal node index.
struct alignas(16) BNode {
    float4 a;  // c0.min.x c0.min.y c0.min.z c0.max.x
    float4 b;  // c0.max.y c0.max.z c1.min.x c1.min.y
    float4 c;  // c1.min.z c1.max.x c1.max.y c1.max.z
    int4 d;    // c0 ref, c1 ref, parent, unused
};
constexpr uint32_t kLeafBit = 0x80000000u;

// blocks rounded up to a multiple of the XCD count (see xcd_block)
inline unsigned xcd_grid(uint64_t blocks) { return (unsigned)((blocks + 7) / 8 * 8); }

// ---- launchers (each .hip file owns its kernels) --------------------------
// k_prepare.hip
void launch_records(const double *V, const uint32_t *T, uint32_t nt, double *tri64, float *tri32, float *box32,
                    float *blockBounds, int nblocks, hipStream_t s);
void launch_normals64(const double *tri64, uint32_t nt, double *out, hipStream_t s);
int records_blocks(uint32_t nt);
void launch_boxes64(const double *tri64, uint32_t nt, double *out, hipStream_t s);
void launch_bounds_reduce(const float *blockBounds, int nblocks, float *bounds, hipStream_t s);
// also clears zero[0, nzero) (the hierarchy's scratch, one launch fewer)
void launch_morton(const float *box32, const float *bounds, uint32_t nt, uint32_t *keys, uint32_t *vals,
                   uint32_t *zero, uint32_t nzero, hipStream_t s);

// k_sort.hip: stable LSD radix sort; keys (and optional u32 values) ping-pong
// between the two buffers; returns 1 when the result ended in the alt buffers.
struct SortTemp {
    uint32_t *hist = nullptr;  // 256 * nblocks
    size_t histCap = 0;
};
int sort_blocks(size_t n);
void launch_exclusive_scan(uint32_t *data, size_t m, hipStream_t s);
size_t sort_hist_entries(size_t n);
// exclusive scan of m u32 in place; temp holds scan_large_temp(m) u32
size_t scan_large_temp(size_t m);
void launch_scan_large(uint32_t *data, size_t m, uint32_t *temp, hipStream_t s);
int radix_sort_u32(uint32_t *keys, uint32_t *vals, uint32_t *keysAlt, uint32_t *valsAlt, size_t n,
                   int beginBit, int endBit, uint32_t *hist, hipStream_t s);
// keys sorted by their high word: each run of equal high words sorted by the
// low word; a run longer than maxRun raises *flag and is left unsorted.
// count (may be null, here and below): only the first min(n, *count) keys,
// a length known on the device, are read and moved
void launch_sort_runs(uint64_t *keys, size_t n, const uint32_t *count, uint32_t maxRun, uint32_t *flag,
                      hipStream_t s);
int radix_sort_u64(uint64_t *keys, uint64_t *keysAlt, size_t n, const int *bitRanges, int nRanges,
                   uint32_t *hist, hipStream_t s, const uint32_t *count = nullptr);

// k_lbvh.hip
// work: 2n + 1 u32 of scratch shared by the two (top-node count, the leaf
// positions where top nodes' ranges start and end, top-node list);
// refit scratch: refit_scratch_floats(n) floats (per-tile prefix / suffix box
// unions and the tile sparse table)
size_t refit_scratch_floats(uint32_t n);
// launch_hierarchy expects the first hierarchy_zeroed_u32(n) u32 of work to
// be zero (launch_morton clears them); launch_hierarchy64 clears them itself
uint32_t hierarchy_zeroed_u32(uint32_t n);
void launch_hierarchy(const uint32_t *codes, const uint32_t *ids, uint32_t n, BNode *nodes,
                      uint32_t *leafParent, uint32_t *work, hipStream_t s);
void launch_hierarchy64(const uint64_t *codes, const uint32_t *ids, uint32_t n, BNode *nodes,
                        uint32_t *leafParent, uint32_t *work, hipStream_t s);
void launch_refit(const uint32_t *ids, const float *box32, uint32_t n, BNode *nodes,
                  const uint32_t *leafParent, uint32_t *work, float *scratch, float *rootBox, hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t value, size_t n, hipStream_t s);
void launch_copy16(const void *src, void *dst, size_t n16, hipStream_t s);
void launch_copy_u32(const uint32_t *src, uint32_t *dst, size_t n, hipStream_t s);

// k_traverse.hip
struct TraverseArgs {
    const BNode *nodesA;
    const float *boxA;
    const double *tri64A;
    const BNode *nodesB;
    const float *boxB;
    const double *tri64B;
};
void launch_expand(const TraverseArgs &t, const uint2 *in, const uint32_t *inCount, uint2 *out,
                   uint32_t *outCount, uint32_t outCap, uint64_t *cand, uint32_t *candCount,
                   uint32_t candCap, uint32_t *overflow, int grid, hipStream_t s);
// persistent depth-first search from the frontier seeds[0, *seedCount); *head
// zero on entry; spill holds waves * spillPerWave pairs; overflow bit 2: a
// wave's spill region filled up, bit 3: iteration cap (the host retries)
void launch_traverse_dfs(const TraverseArgs &t, const uint2 *seeds, const uint32_t *seedCount, uint32_t seedCap,
                         uint32_t *head, uint2 *spill, uint32_t spillPerWave, int waves, uint64_t *cand,
                         uint32_t *candCount, uint32_t candCap, uint32_t *overflow, hipStream_t s);
// the single-pair search after launch_seed: levels in one workgroup while a
// level holds at most topMax pairs, `extra` grid-wide levels, then the
// depth-first stage (levels up to lastLevel breadth-first)
void launch_search_levels(const TraverseArgs &t, uint2 *front0, uint2 *front1, uint32_t *cnt, uint32_t fcap,
                          uint32_t topMax, int lastLevel, int extra, uint32_t *levelWord, uint32_t *head,
                          uint2 *spill, uint32_t spillPerWave, int waves, uint64_t *cand, uint32_t *candCount,
                          uint32_t candCap, uint32_t *overflow, int expandGrid, hipStream_t s);
void launch_seed(const TraverseArgs &t, uint32_t rootA, uint32_t rootB, const float *rootBoxA,
                 const float *rootBoxB, uint2 *frontier, uint32_t *count, uint64_t *cand,
                 uint32_t *candCount, hipStream_t s);

// k_tritri.hip
struct GdStats {
    unsigned long long exits[5];   // plane-1, plane-2, interval, hit, coplanar
    unsigned long long fallback;   // pairs re-run in FP64 because a sign was uncertain
    unsigned long long filtered;   // pairs decided by the FP32 filter alone
};
// work: tritri_work_u32(n) u32 of scratch (undecided-pair slices, per-chunk
// counts and statistics)
size_t tritri_work_u32(uint32_t n);
// count (may be null): device-side number of pairs (at most n), read at run time
void launch_tritri(const double *tri64A, const double *tri64B, const float *tri32A, const float *tri32B,
                   const uint64_t *pairs, uint32_t n, const uint32_t *count, int precision, uint8_t *status,
                   double *seg, uint32_t *work, GdStats *stats, hipStream_t s);
void launch_tritri_raw(const double *a9, const double *b9, uint32_t n, uint8_t *status, double *seg,
                       hipStream_t s);
// compaction of hits (status == 1) into (pair, seg) lists, in pair order
void launch_compact_hits(const uint64_t *pairs, const uint8_t *status, const double *seg, uint32_t n,
                         uint32_t *blockCounts, uint64_t *hitPairs, double *hitSeg, uint32_t *nHits,
                         hipStream_t s);
int compact_blocks(uint32_t n);

// k_classify.hip
void launch_mark_intersected(const uint64_t *pairs, const uint8_t *status, uint32_t n, int side,
                             uint8_t *flags, hipStream_t s);
void launch_centroids(const double *tri64, const uint8_t *flags, uint32_t nt, double *pts, hipStream_t s);
// Per-axis projection grids of one or more target meshes sharing one cell
// array (ncells cells + 1 total) and one entry array (cap entries).
struct GridSet {
    const uint32_t *meshOf;  // triangle -> mesh (null: a single mesh, id 0)
    const uint32_t *target;  // query -> target mesh, ~0u for none (null: mesh 0)
    const uint2 *desc;       // mesh -> (G, first cell) (null: (G0, 0))
    const float *boxes;      // mesh -> FP32 root box (lo xyz, hi xyz)
    const uint32_t *cells;
    const uint2 *gcode;      // per entry: two box code words
    const uint32_t *gtri;    // per entry: the triangle
    const uint32_t *order;  // grid build's triangle visiting order, e.g. Morton (null: index order)
    // FP32 records of the target meshes: rays decide their candidates under
    // the FP32 certificate and run in FP64 only when one is uncertain (null:
    // every candidate in FP64)
    const float *tri32;
    uint32_t G0, cap, ncells;
    uint32_t fastCands;  // cell entries one lane scans before its ray goes to the wave kernel (0: 48)
};
uint32_t grid_dim(uint32_t nt, int density);
size_t grid_cells(uint32_t G);  // 3 G^2 + 1
// maxBlocks: workgroups at most, each taking a contiguous run of tiles of 512
// triangles (0: one workgroup per tile of 1,024 triangles, with a larger LDS
// table)
void launch_grid_build(const float *box32, uint32_t nt, const GridSet &gs, uint32_t *cells, uint32_t *scanTemp,
                       uint2 *gcode, uint32_t *gtri, uint32_t maxBlocks, hipStream_t s);
// the single-pair build's cap: it runs beside the pair search
constexpr uint32_t kGridBesideSearch = 256;
// order (may be null): query q uses point order[q].  allAxes: cast x, y and z
// for every query; otherwise x and y, and z only where they disagree (the
// vote's answer is the same; z's bits are then valid only there).  lists:
// ray_list_u32(n) u32 of scratch; counters: 4 u32, zero on entry, counters[0] gets the
// overflow bits (bit 0: a ray has too many distinct hits, bit 1: the grid was
// truncated, cells[ncells] > cap).  tri64 records carry the FP64 normal in
// their pad slots.
size_t ray_list_u32(size_t n);
// per-axis bits [axis][n] -> [n][3]
void launch_interleave_axes(const uint8_t *axisBits, uint32_t n, uint8_t *out, hipStream_t s);
void launch_rays(const GridSet &gs, const double *tri64, const double *pts, const uint32_t *order,
                 const uint8_t *skip, uint32_t n, bool allAxes, uint8_t *axisBits, uint32_t *lists,
                 uint32_t *counters, hipStream_t s);
// Visiting-order queries: q4 holds 4 nt doubles, (x, y, z, tag) of triangle
// order[q] at 4q (tag: target mesh, or skip where flags is given / no
// target); the rays write axisBits by triangle like launch_rays and leave out
// the queries tagged skip or flagged in skip (by triangle; may be null).
void launch_centroids_ordered(const double *tri64, const uint8_t *flags, const uint32_t *order,
                              const uint32_t *target, uint32_t nt, double *q4, hipStream_t s);
void launch_rays_ordered(const GridSet &gs, const double *tri64, const double *q4, const uint32_t *order,
                         const uint8_t *skip, uint32_t n, bool allAxes, uint8_t *axisBits, uint32_t *lists,
                         uint32_t *counters, hipStream_t s);
// measurement: the x/y fast ray kernel alone (lists: ray_list_u32(n) u32, counters: 4 u32
// zero on entry), and the work counts of one such launch (work: 4 u64, zero on
// entry: rays cast, entries scanned, entries passing the code test, deferred)
void launch_rays_xy_ordered(const GridSet &gs, const double *tri64, const double *q4, const uint32_t *order,
                            uint32_t n, uint8_t *axisBits, uint32_t *lists, uint32_t *counters, hipStream_t s);
// the wave kernel alone over the rays launch_rays_xy_ordered deferred (same
// lists and counters)
void launch_rays_wave_ordered(const GridSet &gs, const double *tri64, const double *q4, const uint32_t *order,
                              uint32_t n, uint8_t *axisBits, const uint32_t *lists, uint32_t *counters, hipStream_t s);
void launch_ray_work_ordered(const GridSet &gs, const double *q4, const uint32_t *order, uint32_t n,
                             unsigned long long *work, hipStream_t s);
void launch_vote(const uint8_t *axisBits, const uint8_t *skip, uint32_t n, uint8_t *out, hipStream_t s);

// k_components.hip: triangle components across opposite half-edges
size_t component_table_slots(uint32_t nt);
void launch_components(const uint32_t *T, uint32_t nt, const uint8_t *blocked, unsigned long long *keys,
                       uint32_t *owners, uint32_t *parent, uint32_t *labels, hipStream_t s);
void launch_half_edges(const uint32_t *T, uint32_t nt, const unsigned long long *queries, uint32_t nq,
                       unsigned long long *keys, uint32_t *owners, uint32_t *opposite, uint32_t *queryOwner,
                       uint32_t *firstRepeat, hipStream_t s);

// k_adjoint.hip
void launch_adjoint(const double *xyz, uint32_t nin, const uint32_t *src, uint32_t nd, const uint32_t *faces,
                    uint32_t nf, double weight, const double *gradDerived, double *all, double *vol, double *gall,
                    hipStream_t s);

// k_batch.hip
void launch_batch_bounds_init(int *mb, uint32_t nmeshes, hipStream_t s);
// waves: 6 * ceil(nt / 64) ints of scratch
void launch_batch_bounds(const float *box32, const uint32_t *meshOf, const uint32_t *triBegin, uint32_t nmeshes,
                         uint32_t nt, int *mb, int *waves, hipStream_t s);
void launch_batch_keys(const float *box32, const uint32_t *meshOf, const uint32_t *triBegin, const int *mb,
                       uint32_t nt, int localBits, int mortonBits, uint64_t *keys, hipStream_t s);
void launch_batch_codes(const uint64_t *keys, uint32_t n, int localBits, int mortonBits, const uint32_t *triBegin,
                        uint64_t *codes, uint32_t *ids, hipStream_t s);
void launch_mesh_roots(const BNode *nodes, const uint32_t *ids, const uint32_t *triBegin, uint32_t nmeshes,
                       uint32_t n, const float *box32, const float *rootBox, uint32_t *meshRoot, float *meshBox,
                       hipStream_t s);
void launch_batch_seed(const TraverseArgs &t, const uint32_t *meshRoot, const float *meshBox, uint32_t npairs,
                       uint2 *frontier, uint32_t *frontCount, uint32_t frontCap, uint64_t *cand, uint32_t *candCount,
                       uint32_t candCap, uint32_t *overflow, hipStream_t s);
void launch_mark_both(const uint64_t *pairs, const uint8_t *status, const uint32_t *count, uint32_t cap,
                      uint8_t *flags, hipStream_t s);

}  // namespace sbg
