/*
 * sbgpu.h - C ABI of the MI355X (gfx950) hot path of a closed-mesh boolean:
 * LBVH build, tree-vs-tree candidate pairs, Guigue-Devillers triangle-triangle
 * test with segment construction, and ray-parity inside/outside classification.
 *
 * Plain pointers and sizes only.  Every function returns an int status
 * (SBG_OK = 0, negative on error; sbg_last_error() has the message) and never
 * throws.  All device work of one context is ordered on that context's HIP
 * stream; functions that hand results back to the host synchronise it.  One
 * context (and the meshes / pair lists made from it) must be used by one host
 * thread at a time; distinct contexts are independent.
 *
 * Reference interfaces replaced (huxingyi/solidboolean):
 *   SolidMesh::setVertices / setTriangles / prepare   src/solidmesh.h:31-34, src/solidmesh.cpp:42-76
 *   SolidMesh::triangleNormals / triangleAxisAlignedBoundingBoxes  src/solidmesh.h:46-59
 *   SolidBoolean::searchPotentialIntersectedPairs      src/solidboolean.cpp:94-101
 *   SolidBoolean::intersectTwoFaces (+ the pair loop)  src/solidboolean.cpp:103-122, 315-341
 *   tri_tri_intersection_test_3d                       thirdparty/GuigueDevillers03/tri_tri_intersect.h:46-49
 *   SolidBoolean::isPointInMesh / decideGroupSide      src/solidboolean.cpp:48-92, 482-510
 */
#ifndef SBGPU_H
#define SBGPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBG_OK 0
#define SBG_ERR_INVALID (-1)   /* bad argument (null handle, index out of range, ...) */
#define SBG_ERR_HIP (-2)       /* a HIP runtime call failed */
#define SBG_ERR_OOM (-3)       /* device allocation failed */
#define SBG_ERR_OVERFLOW (-4)  /* a bounded device list overflowed (see message) */
#define SBG_ERR_NO_DEVICE (-5) /* no usable GPU */
#define SBG_ERR_STATE (-6)     /* call out of order (e.g. results before sbg_tri_tri) */

#define SBG_PRECISION_FP64 0          /* the reference's FP64 operation sequence */
#define SBG_PRECISION_FP32_FILTERED 1 /* FP32 sign certificates, FP64 where uncertain; same results */

typedef struct sbg_context_s *sbg_context;
typedef struct sbg_mesh_s *sbg_mesh;
typedef struct sbg_pairs_s *sbg_pairs;

typedef struct {
    uint64_t tested;         /* candidate pairs tested */
    uint64_t exit_plane1;    /* rejected by T1 vs plane of T2 (tri_tri_intersect.c:421) */
    uint64_t exit_plane2;    /* rejected by T2 vs plane of T1 (:438) */
    uint64_t exit_interval;  /* rejected inside CONSTRUCT_INTERSECTION (:322, :328) */
    uint64_t hits;           /* intersecting, non-coplanar: a segment was built */
    uint64_t coplanar;       /* coplanar (intersectTwoFaces returns false, solidboolean.cpp:119) */
    uint64_t fallback_pairs; /* FP32-filtered mode: pairs re-run in FP64 for an uncertain sign */
    uint64_t filtered_pairs; /* FP32-filtered mode: pairs decided by FP32 certificates alone */
} sbg_tri_tri_stats;

/* Stage timers (milliseconds, HIP events on the context stream). */
enum {
    SBG_STAGE_RECORDS = 0, /* triangle records, normals, leaf boxes, Morton keys */
    SBG_STAGE_SORT,        /* Morton radix sort */
    SBG_STAGE_BVH,         /* Karras hierarchy + refit */
    SBG_STAGE_TRAVERSE,    /* tree-vs-tree candidate search */
    SBG_STAGE_PAIR_SORT,   /* canonical (a, b) order of the candidates */
    SBG_STAGE_TRI_TRI,     /* Guigue-Devillers kernel */
    SBG_STAGE_COMPACT,     /* hit compaction */
    SBG_STAGE_CLASSIFY,    /* ray parity + vote */
    SBG_STAGE_GRID,        /* per-axis ray grid of a target mesh */
    SBG_STAGE_COUNT
};

/* Last error message of the calling thread ("" if none). */
const char *sbg_last_error(void);

/* Number of visible HIP devices. */
int sbg_device_count(int *count);

/* A context binds one device and one HIP stream (default priority). */
int sbg_context_create(int device, sbg_context *out);
int sbg_context_destroy(sbg_context ctx);
int sbg_context_synchronize(sbg_context ctx);
/* The context's hipStream_t, for callers that order their own work after it. */
int sbg_context_stream(sbg_context ctx, void **hip_stream);

/* Mesh from host arrays (copied to HBM): xyz is nverts x 3 FP64, tris is
 * ntris x 3 vertex indices.  Replaces SolidMesh::setVertices/setTriangles
 * (src/solidmesh.h:31-32).  Indices are validated; ntris < 2^31. */
int sbg_mesh_create(sbg_context ctx, const double *xyz, size_t nverts, const uint32_t *tris, size_t ntris,
                    sbg_mesh *out);
/* The same with 64-bit vertex indices, as SolidMesh::setTriangles takes them
 * (std::vector<std::vector<size_t>>, src/solidmesh.h:34); narrowed to the
 * device's 32 bits while staged, an index >= nverts (or >= 2^32) is an error.
 * Both entry points stage the copies through page-locked memory on a few
 * host threads and return once the data is in HBM. */
int sbg_mesh_create64(sbg_context ctx, const double *xyz, size_t nverts, const uint64_t *tris, size_t ntris,
                      sbg_mesh *out);
/* Page-locked host memory, for result arrays: the classification entry
 * points copy their votes straight into page-locked output arrays by DMA
 * (any page-locked memory qualifies; other memory gets a staged copy).  No
 * counterpart in the reference. */
int sbg_host_alloc(size_t bytes, void **out);
int sbg_host_free(void *p);
/* Device SolidMesh::prepare (src/solidmesh.cpp:42-76): FP64 normals
 * (src/vector3.h:155-176), per-triangle boxes and the LBVH.  Asynchronous. */
int sbg_mesh_prepare(sbg_mesh mesh);
int sbg_mesh_destroy(sbg_mesh mesh);
int sbg_mesh_counts(sbg_mesh mesh, size_t *nverts, size_t *ntris);
/* SolidMesh::triangleNormals (src/solidmesh.h:46-49): ntris x 3 FP64. */
int sbg_mesh_normals(sbg_mesh mesh, double *out);
/* SolidMesh::triangleAxisAlignedBoundingBoxes (src/solidmesh.h:56-59):
 * ntris x 6 FP64 = min xyz, max xyz (exact, as AxisAlignedBoudingBox::update). */
int sbg_mesh_boxes(sbg_mesh mesh, double *out);
/* The LBVH for inspection: ntris-1 nodes of 64 B (child boxes, child refs,
 * parent) into nodes_out (may be NULL), the root reference and FP32 root box. */
int sbg_mesh_bvh(sbg_mesh mesh, void *nodes_out, uint32_t *root_ref, float *root_box6);

/* SolidBoolean::searchPotentialIntersectedPairs (src/solidboolean.cpp:94-101):
 * every (a in A, b in B) whose FP64 triangle boxes closed-overlap
 * (src/axisalignedboundingbox.h:95-105), sorted by (a, b). */
int sbg_find_pairs(sbg_mesh a, sbg_mesh b, sbg_pairs *out);
/* A pair list from host memory: ab = n x (a, b). */
int sbg_pairs_create(sbg_context ctx, const uint32_t *ab, size_t n, sbg_pairs *out);
int sbg_pairs_count(sbg_pairs pairs, size_t *n);
int sbg_pairs_copy(sbg_pairs pairs, uint32_t *ab);
int sbg_pairs_destroy(sbg_pairs pairs);

/* intersectTwoFaces over every pair (src/solidboolean.cpp:103-122): stores
 * per-pair status (0 none, 1 intersecting segment, 2 coplanar) and segments in
 * the pair list.  stats may be NULL (then the call does not synchronise). */
int sbg_tri_tri(sbg_mesh a, sbg_mesh b, sbg_pairs pairs, int precision, sbg_tri_tri_stats *stats);
/* Per-pair status (n bytes) and segments (n x 6 FP64: source, target;
 * meaningful where status == 1).  Either pointer may be NULL. */
int sbg_pairs_results(sbg_pairs pairs, uint8_t *status, double *seg);
/* Intersecting pairs only, in pair order.  With ab == NULL only *nhits is set. */
int sbg_pairs_hits(sbg_pairs pairs, size_t *nhits, uint32_t *ab, double *seg);

/* tri_tri_intersection_test_3d as intersectTwoFaces uses it, on n host
 * triangle pairs (a9, b9: n x 9 FP64 = p, q, r).  status as above; seg n x 6. */
int sbg_tri_tri_batch(sbg_context ctx, const double *a9, const double *b9, size_t n, uint8_t *status,
                      double *seg);

/* isPointInMesh along the three axes of src/solidboolean.cpp:31-35 for n host
 * points against `target`, and decideGroupSide's vote (:492-508).  axis_inside
 * (n x 3, may be NULL) gets each ray's parity; inside gets the vote. */
int sbg_classify_points(sbg_mesh target, const double *pts, size_t n, uint8_t *axis_inside, uint8_t *inside);

/* sbg_classify_points with no device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * results as sbg_classify_points for n <= 65536 points (larger n fails with
 * SBG_ERR_INVALID).  Every buffer its kernels use beyond the caller's inputs
 * and outputs lives in one fixed device array of about 8.4 MB per GPU, held
 * until the process ends; it keeps the last call's contents.  Calls that
 * overlap on one GPU (from two threads or contexts) race on that array.  The
 * target's ray grid must already exist: classify against the target once
 * with an allocating call first (SBG_ERR_STATE otherwise). */
int sbg_classify_points_noalloc(sbg_mesh target, const double *pts, size_t n, uint8_t *axis_inside,
                                uint8_t *inside);

/* Every triangle of `self` that no intersecting pair of `pairs` touches is
 * classified by its centroid (v0 + v1 + v2) / 3.0 against `other`
 * (src/solidboolean.cpp:497-508).  side = 0 when self is the pairs' first mesh,
 * 1 when it is the second.  out (ntris bytes): 0 outside, 1 inside,
 * 2 intersected (not classified).  Requires sbg_tri_tri on `pairs`. */
int sbg_classify_triangles(sbg_mesh self, sbg_mesh other, sbg_pairs pairs, int side, uint8_t *out);

/* Both meshes of one boolean at once (the vote of src/solidboolean.cpp:497-508
 * for every uncut triangle of each): out_a for a against b (pairs side 0) and
 * out_b for b against a (side 1), exactly as two sbg_classify_triangles calls
 * give them.  Both sides are queued before one wait, so the host's launches
 * for the second overlap the first's kernels. */
int sbg_classify_pair(sbg_mesh a, sbg_mesh b, sbg_pairs pairs, uint8_t *out_a, uint8_t *out_b);

/* Face groups of src/solidboolean.cpp:167-239 for triangles that no
 * intersection loop touches: connected components across opposite half-edges
 * (triangle t's edge i -> j leads to the triangle owning j -> i), over the
 * triangles with blocked[t] == 0 (blocked may be NULL).  labels[t] (ntris) =
 * the smallest triangle index of t's component, which is where the
 * reference's leftover pass (:230-238) starts that group, or 0xffffffff for a
 * blocked triangle.  A directed edge shared by two unblocked triangles fails
 * with SBG_ERR_INVALID, as addUnintersectedTriangles (:250-285) fails on a
 * repeated half-edge.  Needs no sbg_mesh_prepare. */
int sbg_components(sbg_mesh m, const uint8_t *blocked, uint32_t *labels);

/* The half-edge map of src/solidboolean.cpp:250-286 and :386-396 (directed
 * edge (from << 32 | to) -> the triangle that has it), built over a host
 * triangle list in index order, answering what buildFaceGroups (:167-239)
 * asks of it.  opposite[3 t + i] (ntris * 3) = the triangle that has the
 * reverse of t's edge from corner i to corner i + 1 (the earliest one), or
 * 0xffffffff; query_owner[q] (nqueries) = the triangle that has directed edge
 * queries[q], or 0xffffffff; *first_repeat = the first triangle whose
 * insertion would meet a directed edge already present (where
 * addUnintersectedTriangles, :271-282, fails), or 0xffffffff.  Vertex indices
 * must be below 0xffffffff. */
int sbg_half_edges(sbg_context ctx, const uint32_t *tris, size_t ntris, const uint64_t *queries, size_t nqueries,
                   uint32_t *opposite, uint32_t *query_owner, uint32_t *first_repeat);

/* Stage timers: when enabled, every stage is bracketed by HIP events.
 * sbg_context_stage_ms synchronises and returns the milliseconds accumulated
 * per stage (SBG_STAGE_COUNT doubles) since the last reset. */
int sbg_context_set_timing(sbg_context ctx, int enable);
int sbg_context_stage_ms(sbg_context ctx, double *ms, int reset);

/* Runs the tri-tri kernel `iters` times back to back on the context stream
 * between two HIP events; *avg_ms = elapsed / iters. */
int sbg_bench_tri_tri(sbg_mesh a, sbg_mesh b, sbg_pairs pairs, int precision, int iters, double *avg_ms);

/* Measurement aid (no counterpart in the reference): sets up the rays of
 * sbg_classify_triangles(self, other, pairs, side) and runs the fast x/y ray
 * kernel alone `iters` times (in the context's ray precision), HIP events
 * around each launch on the context stream; *avg_ms = mean launch time.  work
 * (5 values, may be null) gets the work of one launch: rays cast, grid
 * entries scanned by the fast path, entries passing the code test, rays whose
 * cell is too crowded for the fast path, and the rays the timed kernel handed
 * to the FP64 wave path (crowded cells, more distinct hits than it holds,
 * and in the FP32-filtered precision uncertain certificates). */
int sbg_bench_rays(sbg_mesh self, sbg_mesh other, sbg_pairs pairs, int side, int iters, double *avg_ms,
                   uint64_t *work);

/* Measurement aid (no counterpart in the reference): the same set-up and one
 * launch of the x/y ray kernel, then the FP64 wave kernel alone `iters` times
 * over the rays that launch deferred (HIP events around each launch on the
 * context stream); *avg_ms = mean launch time, *rays (may be null) = the rays
 * it handles. */
int sbg_bench_rays_wave(sbg_mesh self, sbg_mesh other, sbg_pairs pairs, int side, int iters, double *avg_ms,
                        uint64_t *rays);

/* Testing aids (no counterpart in the reference), run on the host, no device
 * needed: the FP32 certificate the ray kernels apply to one candidate of
 * isPointInMesh (src/solidboolean.cpp:64-87), for n candidates: point pts
 * (n x 3 FP64), axis 0/1/2, triangle tris9 (n x 9 FP64).  verdict: 0 a
 * certain miss, 1 a certain hit whose FP64 hit point lies within et of
 * p + t along the axis, 2 not certain (the kernels then run the ray in FP64).
 * sbg_ray_hits_apart: 1 where two certain hits (t, et) of one ray are
 * certain to have different PositionKeys. */
int sbg_ray_certify(const double *pts, const int32_t *axis, const double *tris9, size_t n, int32_t *verdict,
                    float *t, float *et);
int sbg_ray_hits_apart(const float *ta, const float *ea, const float *tb, const float *eb, size_t n,
                       uint8_t *out);

/* Measurement aid (no counterpart in the reference): a plain device-to-device
 * copy kernel of `bytes` bytes run `iters` times, for calibrating the HBM
 * traffic counters (FETCH_SIZE / WRITE_SIZE) against a known byte count. */
int sbg_calibrate_copy(sbg_context ctx, size_t bytes, int iters, double *avg_ms);

/* ---- tuning of the single-pair path -----------------------------------------
 * Knobs that change speed, never results.  sbg_autotune_pair times the chain
 * prepare both meshes -> find_pairs -> tri_tri (FP32-filtered) -> classify
 * both meshes over a coordinate search of the knobs and leaves the fastest
 * set (it re-prepares both meshes). */
typedef struct sbg_tuning {
    int bfs_levels;   /* breadth-first levels before the persistent depth-first search (0..64) */
    int dfs_waves;    /* one-wave workgroups of the depth-first search (64..16384) */
    int grid_density; /* ray-grid cells per axis plane = triangles * grid_density / 16 (1..64) */
    int depth_first;  /* 0: breadth-first levels only, with a host check every 8 levels */
    int early_grids;  /* 2 (default): sbg_mesh_prepare queues the mesh's ray grid on a third
                         stream after the prepare, so it builds beside the other mesh's prepare
                         and the pair search; 1: sbg_find_pairs queues both meshes' grids on the
                         second stream, beside the search; 0: each grid is built when a
                         classification first needs it */
    int ray_precision; /* SBG_PRECISION_FP32_FILTERED (default): each ray candidate is decided from
                          the FP32 record under a certificate, and a ray with an uncertain candidate
                          runs in FP64; SBG_PRECISION_FP64: every candidate in FP64.  Same results. */
    int ray_cell_cap;  /* grid entries one lane scans for its ray (16..4096, default 64); a ray whose
                          cell holds more goes to the kernel that gives each ray a whole wave.  The
                          batch's rays use it too (as it is when a batch run is queued). */
} sbg_tuning;
int sbg_context_set_tuning(sbg_context ctx, const sbg_tuning *t);
int sbg_context_get_tuning(sbg_context ctx, sbg_tuning *t);
int sbg_autotune_pair(sbg_mesh a, sbg_mesh b, int reps, sbg_tuning *best, double *best_ms);

/* ---- topology-fixed adjoint ------------------------------------------------
 * No reference counterpart (the reference computes no derivatives).  nin
 * input vertices xyz_in; nd derived vertices, derived d being the point where
 * the line through inputs src[5d], src[5d + 1] meets the plane through inputs
 * src[5d + 2 .. 5d + 4] (how every intersection-segment endpoint is made,
 * thirdparty/GuigueDevillers03/tri_tri_intersect.c:297-356); faces (nf x 3)
 * index [inputs | derived] and should form a closed, consistently oriented
 * surface (e.g. a union).  Writes the derived positions (nd x 3) and the
 * faces' signed volume, and, when grad_in is non-null,
 *   grad_in = vol_weight * d(volume)/d(xyz_in) + J^T grad_derived,
 * J being d(derived)/d(xyz_in); grad_derived may be null.  Outputs may be null. */
int sbg_adjoint(sbg_context ctx, const double *xyz_in, size_t nin, const uint32_t *src, size_t nd,
                const uint32_t *faces, size_t nf, double vol_weight, const double *grad_derived, double *derived_xyz,
                double *volume, double *grad_in);

/* ---- batched booleans ------------------------------------------------------
 * Many independent mesh pairs in one structure.  There are nmeshes = 2 *
 * npairs meshes in concatenated arrays: mesh m owns vertices
 * [vert_begin[m], vert_begin[m + 1]) of xyz and triangles
 * [tri_begin[m], tri_begin[m + 1]) of tris, whose indices are local to the
 * mesh's own vertices; pair k is (mesh 2k, mesh 2k + 1).  One run is, for
 * every pair, the hot path of one SolidBoolean::combine
 * (src/solidboolean.cpp:94-122 candidate search + Guigue-Devillers, :482-510
 * ray classification of every non-intersected triangle against the partner
 * mesh), all pairs sharing every launch.  The reference runs pairs one by one
 * and has no batched call. */
typedef struct sbg_batch_s *sbg_batch;

typedef struct sbg_batch_tuning {
    int bfs_levels;   /* breadth-first levels before the persistent depth-first search (0..32) */
    int dfs_waves;    /* one-wave workgroups of the depth-first search (64..16384) */
    int grid_density; /* ray-grid cells per axis plane = triangles * grid_density / 16 (1..64) */
    int use_graph;    /* nonzero: once sized, each run replays one captured hipGraph */
} sbg_batch_tuning;

int sbg_batch_create(sbg_context ctx, const double *xyz, const uint32_t *tris, const uint64_t *vert_begin,
                     const uint64_t *tri_begin, size_t nmeshes, sbg_batch *out);
/* new vertex coordinates (same count and topology); the captured graph stays valid */
int sbg_batch_set_vertices(sbg_batch b, const double *xyz);
int sbg_batch_set_tuning(sbg_batch b, const sbg_batch_tuning *t);
int sbg_batch_get_tuning(sbg_batch b, sbg_batch_tuning *t);
/* one run at SBG_PRECISION_*; returns when its results are ready */
int sbg_batch_run(sbg_batch b, int precision);
/* device time of the last run's work (HIP events), in ms */
int sbg_batch_last_ms(sbg_batch b, double *ms);
/* Measurement aid: the Guigue-Devillers kernel alone over the last run's
 * candidates, `iters` launches between two HIP events; *avg_ms per launch;
 * stats (may be null) gets the exit and fallback counts of one launch. */
int sbg_batch_bench_tri_tri(sbg_batch b, int precision, int iters, double *avg_ms, sbg_tri_tri_stats *stats);
int sbg_batch_counts(sbg_batch b, size_t *candidates, size_t *hits);
/* candidates of the last run sorted by (a, b) as global triangle indices
 * (2 per candidate), with status (1 hit, 2 coplanar, 0 none) and, for hits,
 * the segment (6 doubles per candidate); any output may be null */
int sbg_batch_candidates(sbg_batch b, uint32_t *ab, uint8_t *status, double *seg);
/* per triangle: 0 outside / 1 inside its pair partner (3-ray vote), 2 intersected */
int sbg_batch_sides(sbg_batch b, uint8_t *side);
/* coordinate search over the tuning knobs, `reps` timed runs per setting; leaves the best set */
int sbg_batch_autotune(sbg_batch b, int precision, int reps, sbg_batch_tuning *best, double *best_ms);
/* Testing aid (no counterpart in the reference): sets the depth-first search's
 * spill region per wave (0..32768 node pairs) and drops the captured graph.  A
 * run whose search fills the region grows it and repeats, captured runs
 * included, so results never depend on it. */
int sbg_batch_set_spill(sbg_batch b, uint32_t per_wave);
int sbg_batch_destroy(sbg_batch b);

#ifdef __cplusplus
}
#endif
#endif /* SBGPU_H */
