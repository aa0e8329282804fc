/*
 * ORACLE - test infrastructure only.
 *
 * A plain-C, single-threaded, FP64 restatement of the reference's hot path
 * (huxingyi/solidboolean), compiled with -ffp-contract=off so every rounding
 * happens where the reference's does.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it, and only as the checker: the product
 * (solidboolean_amd, libsbgpu.so) never links, loads or calls it.
 *
 * Parity pins (see tests/test_oracle.py):
 *   orc_normal / orc_prepare     : golden normals + boxes from the compiled
 *                                  reference; closed forms on axis-aligned cases
 *   orc_tree_* / orc_find_pairs  : golden DFS-ordered pair lists; brute-force
 *                                  all-pairs closed box overlap
 *   orc_tri_tri                  : golden per-pair (status, segment) vectors;
 *                                  hand-worked axis-aligned intersections;
 *                                  segment endpoints lie on both planes
 *   orc_point_in_mesh / classify : golden ray answers; |p| < r closed form on
 *                                  spheres away from the surface
 */
#ifndef SB_ORACLE_H
#define SB_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Vector3::normal (src/vector3.h:155-176). */
void orc_normal(const double a[3], const double b[3], const double c[3], double out[3]);

/* SolidMesh::prepare's per-triangle part (src/solidmesh.cpp:47-62):
 * normals[nt*3]; boxes[nt*9] = min xyz, max xyz, center xyz. */
void orc_prepare(const double *V, const int64_t *T, size_t nt, double *normals, double *boxes);

/* Reference-shaped top-down tree (src/axisalignedboundingboxtree.cpp:27-141). */
typedef struct orc_mesh orc_mesh;
orc_mesh *orc_mesh_new(const double *V, size_t nv, const int64_t *T, size_t nt);
void orc_mesh_free(orc_mesh *m);
/* nodes, leaves, max depth, max leaf size */
void orc_mesh_tree_stats(const orc_mesh *m, int64_t out[4]);
const double *orc_mesh_normals(const orc_mesh *m);
const double *orc_mesh_boxes(const orc_mesh *m);

/* Candidate pairs in DFS order (src/axisalignedboundingboxtree.h:54-95 via
 * src/solidboolean.cpp:94-101). Returns the count; writes min(count, cap). */
size_t orc_find_pairs(const orc_mesh *a, const orc_mesh *b, int64_t *out, size_t cap);

/* Guigue-Devillers (thirdparty/GuigueDevillers03/tri_tri_intersect.c:285-472).
 * Returns 1 on a non-coplanar intersection (seg = source, target), 0 otherwise;
 * *coplanar = 1 in the coplanar case. */
int orc_tri_tri(const double p1[3], const double q1[3], const double r1[3],
                const double p2[3], const double q2[3], const double r2[3],
                int *coplanar, double source[3], double target[3]);

/* intersectTwoFaces over a pair list (src/solidboolean.cpp:103-122).
 * status: 0 none, 1 hit, 2 coplanar; seg[n*6] written for hits, zero otherwise. */
void orc_tri_tri_pairs(const orc_mesh *a, const orc_mesh *b, const int64_t *pairs, size_t n,
                       uint8_t *status, double *seg);

/* isPointInMesh (src/solidboolean.cpp:48-92) along axis 0/1/2 (:31-35). */
int orc_point_in_mesh(const orc_mesh *m, const double p[3], int axis);

/* One candidate of isPointInMesh (src/solidboolean.cpp:55-87) on its own: the
 * closed overlap of the triangle's box with the segment p -> p + axis
 * direction, the plane hit (src/vector3.h:264-280, n = the triangle's
 * Vector3::normal) and the three-normal inside test.  Returns 1 on a hit with
 * the FP64 hit point in hit. */
int orc_ray_candidate(const double p[3], int axis, const double v0[3], const double v1[3], const double v2[3],
                      const double n[3], double hit[3]);

/* orc_ray_candidate over n candidates (tris9: n x 9 vertices); hits n x 3. */
void orc_ray_candidates(size_t n, const double *pts, const int32_t *axis, const double *tris9, uint8_t *hit,
                        double *hits);

/* decideGroupSide's vote (src/solidboolean.cpp:492-508): 1 iff inside along
 * at least 2 of the 3 axes. */
int orc_classify_point(const orc_mesh *m, const double p[3]);

/* For every triangle of `self` not flagged in `intersected`: the vote of its
 * centroid (v0 + v1 + v2) / 3.0 (src/solidboolean.cpp:497-499) against
 * `other`. out: 0 outside, 1 inside, 2 intersected (not classified). */
void orc_classify_triangles(const orc_mesh *self, const orc_mesh *other,
                            const uint8_t *intersected, uint8_t *out);

/* PositionKey (src/positionkey.cpp:25-37). */
void orc_position_key(const double p[3], int64_t out[3]);

#ifdef __cplusplus
}
#endif
#endif
