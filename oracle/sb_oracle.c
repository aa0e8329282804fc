/*
 * ORACLE - test infrastructure only (see sb_oracle.h).  Plain C99, FP64,
 * built with -ffp-contract=off: every a*b+c below rounds twice, as the
 * reference's does on x86-64 without -march.  Expressions are written with the
 * same operand order and association as the reference lines they cite.
 */
#include "sb_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Planted mistakes for the oracle's own tests (tests/test_oracle_mutations.py
 * builds one mutant per value and checks that the golden / closed-form checks
 * reject it).  0, the default and the only value any other build uses, is the
 * faithful restatement. */
#ifndef ORC_MUTATION
#define ORC_MUTATION 0
#endif
#if ORC_MUTATION == 1 /* open instead of closed box interval */
#define ORC_LE(a, b) ((a) < (b))
#else
#define ORC_LE(a, b) ((a) <= (b))
#endif
#if ORC_MUTATION == 2 /* split ties go left */
#define ORC_SPLIT_LT(c, s) ((c) <= (s))
#else
#define ORC_SPLIT_LT(c, s) ((c) < (s))
#endif
#if ORC_MUTATION == 3 /* transposed operands of the endpoint parameter */
#define ORC_ALPHA(v1, v2, N) (dot3(v2, N) / dot3(v1, N))
#else
#define ORC_ALPHA(v1, v2, N) (dot3(v1, N) / dot3(v2, N))
#endif
#if ORC_MUTATION == 4 /* >= instead of > in CONSTRUCT_INTERSECTION */
#define ORC_GT0(x) ((x) >= 0.0)
#else
#define ORC_GT0(x) ((x) > 0.0)
#endif
#if ORC_MUTATION == 5 /* PositionKey grid off by 10x */
#define ORC_KEY_FACTOR 10000L
#else
#define ORC_KEY_FACTOR 100000L
#endif
#if ORC_MUTATION == 6 /* dropped term in the ray-plane dot product */
#define ORC_PLANE_DOT(n, u) (n[0] * u[0] + n[1] * u[1])
#else
#define ORC_PLANE_DOT(n, u) (n[0] * u[0] + n[1] * u[1] + n[2] * u[2])
#endif
#if ORC_MUTATION == 7 /* >= instead of > in the three-normal inside test */
#define ORC_INSIDE_GT(x) ((x) >= 0)
#else
#define ORC_INSIDE_GT(x) ((x) > 0)
#endif
#if ORC_MUTATION == 8 /* one inside ray out of three carries the vote */
#define ORC_VOTE 0.3
#else
#define ORC_VOTE 0.5
#endif

/* ------------------------------------------------------------------------- */
/* small vector helpers (thirdparty/GuigueDevillers03/tri_tri_intersect.c:73-89,
 * src/vector3.h:140-153)                                                      */
/* ------------------------------------------------------------------------- */
static void sub3(double d[3], const double a[3], const double b[3])
{
    d[0] = a[0] - b[0];
    d[1] = a[1] - b[1];
    d[2] = a[2] - b[2];
}

static void cross3(double d[3], const double a[3], const double b[3])
{
    d[0] = a[1] * b[2] - a[2] * b[1];
    d[1] = a[2] * b[0] - a[0] * b[2];
    d[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* Double::isZero (src/double.h:31-34): absolute threshold DBL_EPSILON. */
static int is_zero(double x) { return fabs(x) <= DBL_EPSILON; }

/* src/vector3.h:155-176 */
void orc_normal(const double a[3], const double b[3], const double c[3], double out[3])
{
    double baX = b[0] - a[0], baY = b[1] - a[1], baZ = b[2] - a[2];
    double caX = c[0] - a[0], caY = c[1] - a[1], caZ = c[2] - a[2];
    double crossX = baY * caZ - baZ * caY;
    double crossY = baZ * caX - baX * caZ;
    double crossZ = baX * caY - baY * caX;
    double length2 = crossX * crossX + crossY * crossY + crossZ * crossZ;
    double length = sqrt(length2);
    if (is_zero(length)) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    out[0] = crossX / length;
    out[1] = crossY / length;
    out[2] = crossZ / length;
}

/* AxisAlignedBoudingBox::update / updateCenter (src/axisalignedboundingbox.h:31-53)
 * for one triangle, from the box's initial state (:108-117). */
static void triangle_box(const double *V, const int64_t *t, double box[9])
{
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double sum[3] = {0.0, 0.0, 0.0};
    for (int v = 0; v < 3; ++v) {
        const double *p = V + 3 * t[v];
        for (int i = 0; i < 3; ++i) {
            if (p[i] > mx[i])
                mx[i] = p[i];
            if (p[i] < mn[i])
                mn[i] = p[i];
            sum[i] += p[i];
        }
    }
    for (int i = 0; i < 3; ++i) {
        box[i] = mn[i];
        box[3 + i] = mx[i];
        box[6 + i] = sum[i] / (double)(float)3; /* m_sum /= (float)m_num */
    }
}

void orc_prepare(const double *V, const int64_t *T, size_t nt, double *normals, double *boxes)
{
    for (size_t i = 0; i < nt; ++i) {
        const int64_t *t = T + 3 * i;
        if (normals)
            orc_normal(V + 3 * t[0], V + 3 * t[1], V + 3 * t[2], normals + 3 * i);
        if (boxes)
            triangle_box(V, t, boxes + 9 * i);
    }
}

/* closed-interval overlap (src/axisalignedboundingbox.h:95-105) */
static int boxes_overlap(const double *amin, const double *amax, const double *bmin, const double *bmax)
{
    for (int i = 0; i < 3; ++i) {
        if (ORC_LE(amin[i], bmax[i]) && amax[i] >= bmin[i])
            continue;
        return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------------- */
/* reference-shaped tree (src/axisalignedboundingboxtree.cpp:25-141)           */
/* ------------------------------------------------------------------------- */
typedef struct {
    double mn[3], mx[3], center[3];
    int64_t start, count; /* range of the mesh's index permutation */
    int64_t left, right;  /* -1 for none */
} node_t;

struct orc_mesh {
    size_t nv, nt;
    double *V;
    int64_t *T;
    double *normals;
    double *boxes; /* nt * 9 */
    node_t *nodes;
    size_t nnodes, capnodes;
    int64_t *perm;    /* nt: node index lists live in ranges of this array */
    int64_t *scratch; /* 2 * nt + 2: m_boxIndicesOrderList */
};

#define LEAF_MAX 20 /* m_leafMaxNodeSize, axisalignedboundingboxtree.cpp:25 */

static int64_t new_node(orc_mesh *m)
{
    if (m->nnodes == m->capnodes) {
        m->capnodes = m->capnodes ? m->capnodes * 2 : 64;
        m->nodes = (node_t *)realloc(m->nodes, m->capnodes * sizeof(node_t));
    }
    node_t *n = &m->nodes[m->nnodes];
    for (int i = 0; i < 3; ++i) {
        n->mn[i] = DBL_MAX;
        n->mx[i] = -DBL_MAX;
        n->center[i] = 0.0;
    }
    n->start = n->count = 0;
    n->left = n->right = -1;
    return (int64_t)m->nnodes++;
}

static void box_update(double *mn, double *mx, const double *p)
{
    for (int i = 0; i < 3; ++i) {
        if (p[i] > mx[i])
            mx[i] = p[i];
        if (p[i] < mn[i])
            mn[i] = p[i];
    }
}

/* splitNode (axisalignedboundingboxtree.cpp:71-141), iterative on an explicit
 * stack with the same left-then-right order; a node's index list is the range
 * [start, start + count) of perm, in the reference's list order. */
static void split_all(orc_mesh *m, int64_t root)
{
    int64_t *stack = (int64_t *)malloc(sizeof(int64_t) * 256);
    size_t sp = 0, cap = 256;
    stack[sp++] = root;
    while (sp) {
        int64_t ni = stack[--sp];
        node_t nd = m->nodes[ni];
        size_t n = (size_t)nd.count;
        if (n <= LEAF_MAX)
            continue;
        /* longest axis of float-narrowed spans; std::max_element keeps the first
         * maximum (axisalignedboundingboxtree.cpp:79-84, .h:101) */
        float spans[3];
        for (int i = 0; i < 3; ++i)
            spans[i] = (float)(nd.mx[i] - nd.mn[i]);
        int axis = 0;
        for (int i = 1; i < 3; ++i)
            if (spans[axis] < spans[i])
                axis = i;
        double split = nd.center[axis];
        int64_t *buf = m->scratch;
        size_t leftOffset = n, rightOffset = n - 1, leftCount = 0, rightCount = 0;
        const int64_t *list = m->perm + nd.start;
        for (size_t i = 0; i < n; ++i) {
            int64_t bi = list[i];
            double c = m->boxes[9 * bi + 6 + axis];
            if (ORC_SPLIT_LT(c, split)) {
                buf[--leftOffset] = bi;
                ++leftCount;
            } else {
                buf[++rightOffset] = bi;
                ++rightCount;
            }
        }
        if (leftCount == 0) {
            leftCount = rightCount / 2;
            rightCount -= leftCount;
            leftOffset = rightOffset - n + 1;
        } else if (rightCount == 0) {
            rightCount = leftCount / 2;
            leftCount -= rightCount;
            rightOffset = leftOffset + n - 1;
        }
        size_t middle = leftOffset + leftCount - 1;
        int64_t li = new_node(m);
        int64_t ri = new_node(m);
        node_t *L = &m->nodes[li];
        node_t *R = &m->nodes[ri];
        int64_t *dst = m->perm + nd.start;
        size_t k = 0;
        L->start = nd.start;
        for (size_t i = leftOffset; i <= middle; ++i) {
            int64_t bi = buf[i];
            const double *bx = m->boxes + 9 * bi;
            box_update(L->mn, L->mx, bx);
            box_update(L->mn, L->mx, bx + 3);
            dst[k++] = bi;
            for (int d = 0; d < 3; ++d)
                L->center[d] += bx[6 + d];
        }
        L->count = (int64_t)k;
        R->start = nd.start + (int64_t)k;
        for (size_t i = middle + 1; i <= rightOffset; ++i) {
            int64_t bi = buf[i];
            const double *bx = m->boxes + 9 * bi;
            box_update(R->mn, R->mx, bx);
            box_update(R->mn, R->mx, bx + 3);
            dst[k++] = bi;
            for (int d = 0; d < 3; ++d)
                R->center[d] += bx[6 + d];
        }
        R->count = (int64_t)k - L->count;
        for (int d = 0; d < 3; ++d) {
            L->center[d] /= (double)(float)L->count;
            R->center[d] /= (double)(float)R->count;
        }
        m->nodes[ni].left = li;
        m->nodes[ni].right = ri;
        if (sp + 2 > cap) {
            cap *= 2;
            stack = (int64_t *)realloc(stack, sizeof(int64_t) * cap);
        }
        stack[sp++] = ri; /* left is split first */
        stack[sp++] = li;
    }
    free(stack);
}

orc_mesh *orc_mesh_new(const double *V, size_t nv, const int64_t *T, size_t nt)
{
    orc_mesh *m = (orc_mesh *)calloc(1, sizeof(orc_mesh));
    m->nv = nv;
    m->nt = nt;
    m->V = (double *)malloc(sizeof(double) * 3 * (nv ? nv : 1));
    memcpy(m->V, V, sizeof(double) * 3 * nv);
    m->T = (int64_t *)calloc(3 * (nt ? nt : 1), sizeof(int64_t));
    memcpy(m->T, T, sizeof(int64_t) * 3 * nt);
    m->normals = (double *)malloc(sizeof(double) * 3 * (nt ? nt : 1));
    m->boxes = (double *)malloc(sizeof(double) * 9 * (nt ? nt : 1));
    orc_prepare(m->V, m->T, nt, m->normals, m->boxes);
    m->perm = (int64_t *)malloc(sizeof(int64_t) * (nt ? nt : 1));
    m->scratch = (int64_t *)malloc(sizeof(int64_t) * (2 * nt + 2));
    for (size_t i = 0; i < nt; ++i)
        m->perm[i] = (int64_t)i;
    /* root: box of every triangle vertex, center = mean of box centers
     * (solidmesh.cpp:64-75, axisalignedboundingboxtree.cpp:27-45) */
    int64_t r = new_node(m);
    node_t *root = &m->nodes[r];
    for (size_t i = 0; i < nt; ++i)
        for (int v = 0; v < 3; ++v)
            box_update(root->mn, root->mx, m->V + 3 * m->T[3 * i + v]);
    root->start = 0;
    root->count = (int64_t)nt;
    if (nt) {
        for (size_t i = 0; i < nt; ++i)
            for (int d = 0; d < 3; ++d)
                root->center[d] += m->boxes[9 * i + 6 + d];
        for (int d = 0; d < 3; ++d)
            root->center[d] /= (double)(float)nt;
    }
    split_all(m, r);
    return m;
}

void orc_mesh_free(orc_mesh *m)
{
    if (!m)
        return;
    free(m->V);
    free(m->T);
    free(m->normals);
    free(m->boxes);
    free(m->nodes);
    free(m->perm);
    free(m->scratch);
    free(m);
}

const double *orc_mesh_normals(const orc_mesh *m) { return m->normals; }
const double *orc_mesh_boxes(const orc_mesh *m) { return m->boxes; }

static void stats_rec(const orc_mesh *m, int64_t ni, int64_t depth, int64_t *st)
{
    const node_t *n = &m->nodes[ni];
    st[0] += 1;
    if (n->left < 0 && n->right < 0) {
        st[1] += 1;
        if (depth > st[2])
            st[2] = depth;
        if (n->count > st[3])
            st[3] = n->count;
        return;
    }
    stats_rec(m, n->left, depth + 1, st);
    stats_rec(m, n->right, depth + 1, st);
}

void orc_mesh_tree_stats(const orc_mesh *m, int64_t out[4])
{
    out[0] = out[1] = out[2] = out[3] = 0;
    if (m->nnodes)
        stats_rec(m, 0, 0, out);
}

/* ------------------------------------------------------------------------- */
/* simultaneous DFS (src/axisalignedboundingboxtree.h:54-88)                   */
/* ------------------------------------------------------------------------- */
typedef struct {
    int64_t *out;
    size_t cap, n;
} pairbuf_t;

static void emit(pairbuf_t *pb, int64_t a, int64_t b)
{
    if (pb->n < pb->cap) {
        pb->out[2 * pb->n] = a;
        pb->out[2 * pb->n + 1] = b;
    }
    pb->n++;
}

static void test_nodes(const orc_mesh *A, int64_t ai, const orc_mesh *B, int64_t bi, pairbuf_t *pb)
{
    const node_t *f = &A->nodes[ai];
    const node_t *s = &B->nodes[bi];
    if (!boxes_overlap(f->mn, f->mx, s->mn, s->mx))
        return;
    int fleaf = f->left < 0 && f->right < 0;
    int sleaf = s->left < 0 && s->right < 0;
    if (fleaf) {
        if (sleaf) {
            for (int64_t i = 0; i < f->count; ++i) {
                int64_t a = A->perm[f->start + i];
                const double *ba = A->boxes + 9 * a;
                for (int64_t j = 0; j < s->count; ++j) {
                    int64_t b = B->perm[s->start + j];
                    const double *bb = B->boxes + 9 * b;
                    if (boxes_overlap(ba, ba + 3, bb, bb + 3))
                        emit(pb, a, b);
                }
            }
        } else {
            test_nodes(A, ai, B, s->left, pb);
            test_nodes(A, ai, B, s->right, pb);
        }
    } else {
        if (sleaf) {
            test_nodes(A, f->left, B, bi, pb);
            test_nodes(A, f->right, B, bi, pb);
        } else if (f->count < s->count) {
            test_nodes(A, ai, B, s->left, pb);
            test_nodes(A, ai, B, s->right, pb);
        } else {
            test_nodes(A, f->left, B, bi, pb);
            test_nodes(A, f->right, B, bi, pb);
        }
    }
}

size_t orc_find_pairs(const orc_mesh *a, const orc_mesh *b, int64_t *out, size_t cap)
{
    pairbuf_t pb = {out, out ? cap : 0, 0};
    if (a->nnodes && b->nnodes)
        test_nodes(a, 0, b, 0, &pb);
    return pb.n;
}

/* ------------------------------------------------------------------------- */
/* Guigue-Devillers with segment construction                                  */
/* (thirdparty/GuigueDevillers03/tri_tri_intersect.c:285-472)                  */
/* ------------------------------------------------------------------------- */

/* One endpoint: the edge (base, other) cut by the plane through `onplane` with
 * normal N:  v1 = base - onplane, v2 = base - other,
 * alpha = DOT(v1,N) / DOT(v2,N), out = base - alpha * v2   (:297-301 etc.) */
static void edge_plane(const double base[3], const double onplane[3], const double other[3],
                       const double N[3], double out[3])
{
    double v1[3], v2[3], s[3];
    sub3(v1, base, onplane);
    sub3(v2, base, other);
    double alpha = ORC_ALPHA(v1, v2, N);
    s[0] = alpha * v2[0];
    s[1] = alpha * v2[1];
    s[2] = alpha * v2[2];
    sub3(out, base, s);
}

/* CONSTRUCT_INTERSECTION (:285-356); N1, N2 are the original planes' normals. */
static int construct(const double *p1, const double *q1, const double *r1, const double *p2,
                     const double *q2, const double *r2, const double *N1, const double *N2,
                     double *src, double *tgt)
{
    double v1[3], v2[3], v[3], N[3];
    sub3(v1, q1, p1);
    sub3(v2, r2, p1);
    cross3(N, v1, v2);
    sub3(v, p2, p1);
    if (ORC_GT0(dot3(v, N))) {
        sub3(v1, r1, p1);
        cross3(N, v1, v2);
        if (dot3(v, N) <= 0.0) {
            sub3(v2, q2, p1);
            cross3(N, v1, v2);
            if (dot3(v, N) > 0.0) {
                edge_plane(p1, p2, r1, N2, src); /* :297-301 */
                edge_plane(p2, p1, r2, N1, tgt); /* :302-306 */
            } else {
                edge_plane(p2, p1, q2, N1, src); /* :309-313 */
                edge_plane(p2, p1, r2, N1, tgt); /* :314-318 */
            }
            return 1;
        }
        return 0;
    }
    sub3(v2, q2, p1);
    cross3(N, v1, v2);
    if (dot3(v, N) < 0.0)
        return 0;
    sub3(v1, r1, p1);
    cross3(N, v1, v2);
    if (dot3(v, N) >= 0.0) {
        edge_plane(p1, p2, r1, N2, src); /* :333-337 */
        edge_plane(p1, p2, q1, N2, tgt); /* :338-342 */
    } else {
        edge_plane(p2, p1, q2, N1, src); /* :345-349 */
        edge_plane(p1, p2, q1, N2, tgt); /* :350-354 */
    }
    return 1;
}

/* TRI_TRI_INTER_3D (:360-385). Returns 1 hit, 0 none, -1 coplanar. */
static int inter3d(const double *p1, const double *q1, const double *r1, const double *p2,
                   const double *q2, const double *r2, double dp2, double dq2, double dr2,
                   const double *N1, const double *N2, double *s, double *t)
{
    if (dp2 > 0.0) {
        if (dq2 > 0.0)
            return construct(p1, r1, q1, r2, p2, q2, N1, N2, s, t);
        if (dr2 > 0.0)
            return construct(p1, r1, q1, q2, r2, p2, N1, N2, s, t);
        return construct(p1, q1, r1, p2, q2, r2, N1, N2, s, t);
    }
    if (dp2 < 0.0) {
        if (dq2 < 0.0)
            return construct(p1, q1, r1, r2, p2, q2, N1, N2, s, t);
        if (dr2 < 0.0)
            return construct(p1, q1, r1, q2, r2, p2, N1, N2, s, t);
        return construct(p1, r1, q1, p2, q2, r2, N1, N2, s, t);
    }
    if (dq2 < 0.0) {
        if (dr2 >= 0.0)
            return construct(p1, r1, q1, q2, r2, p2, N1, N2, s, t);
        return construct(p1, q1, r1, p2, q2, r2, N1, N2, s, t);
    }
    if (dq2 > 0.0) {
        if (dr2 > 0.0)
            return construct(p1, r1, q1, p2, q2, r2, N1, N2, s, t);
        return construct(p1, q1, r1, q2, r2, p2, N1, N2, s, t);
    }
    if (dr2 > 0.0)
        return construct(p1, q1, r1, r2, p2, q2, N1, N2, s, t);
    if (dr2 < 0.0)
        return construct(p1, r1, q1, r2, p2, q2, N1, N2, s, t);
    return -1;
}

/* tri_tri_intersection_test_3d (:395-472). The coplanar branch's 2-D overlap
 * result is not computed: intersectTwoFaces discards it (solidboolean.cpp:119). */
int orc_tri_tri(const double p1[3], const double q1[3], const double r1[3],
                const double p2[3], const double q2[3], const double r2[3],
                int *coplanar, double source[3], double target[3])
{
    double v1[3], v2[3], N1[3], N2[3];
    double dp1, dq1, dr1, dp2, dq2, dr2;
    sub3(v1, p2, r2);
    sub3(v2, q2, r2);
    cross3(N2, v1, v2);
    sub3(v1, p1, r2);
    dp1 = dot3(v1, N2);
    sub3(v1, q1, r2);
    dq1 = dot3(v1, N2);
    sub3(v1, r1, r2);
    dr1 = dot3(v1, N2);
    if (((dp1 * dq1) > 0.0) && ((dp1 * dr1) > 0.0))
        return 0;
    sub3(v1, q1, p1);
    sub3(v2, r1, p1);
    cross3(N1, v1, v2);
    sub3(v1, p2, r1);
    dp2 = dot3(v1, N1);
    sub3(v1, q2, r1);
    dq2 = dot3(v1, N1);
    sub3(v1, r2, r1);
    dr2 = dot3(v1, N1);
    if (((dp2 * dq2) > 0.0) && ((dp2 * dr2) > 0.0))
        return 0;

    int r;
    if (dp1 > 0.0) {
        if (dq1 > 0.0)
            r = inter3d(r1, p1, q1, p2, r2, q2, dp2, dr2, dq2, N1, N2, source, target);
        else if (dr1 > 0.0)
            r = inter3d(q1, r1, p1, p2, r2, q2, dp2, dr2, dq2, N1, N2, source, target);
        else
            r = inter3d(p1, q1, r1, p2, q2, r2, dp2, dq2, dr2, N1, N2, source, target);
    } else if (dp1 < 0.0) {
        if (dq1 < 0.0)
            r = inter3d(r1, p1, q1, p2, q2, r2, dp2, dq2, dr2, N1, N2, source, target);
        else if (dr1 < 0.0)
            r = inter3d(q1, r1, p1, p2, q2, r2, dp2, dq2, dr2, N1, N2, source, target);
        else
            r = inter3d(p1, q1, r1, p2, r2, q2, dp2, dr2, dq2, N1, N2, source, target);
    } else {
        if (dq1 < 0.0) {
            if (dr1 >= 0.0)
                r = inter3d(q1, r1, p1, p2, r2, q2, dp2, dr2, dq2, N1, N2, source, target);
            else
                r = inter3d(p1, q1, r1, p2, q2, r2, dp2, dq2, dr2, N1, N2, source, target);
        } else if (dq1 > 0.0) {
            if (dr1 > 0.0)
                r = inter3d(p1, q1, r1, p2, r2, q2, dp2, dr2, dq2, N1, N2, source, target);
            else
                r = inter3d(q1, r1, p1, p2, q2, r2, dp2, dq2, dr2, N1, N2, source, target);
        } else {
            if (dr1 > 0.0)
                r = inter3d(r1, p1, q1, p2, q2, r2, dp2, dq2, dr2, N1, N2, source, target);
            else if (dr1 < 0.0)
                r = inter3d(r1, p1, q1, p2, r2, q2, dp2, dr2, dq2, N1, N2, source, target);
            else
                r = -1;
        }
    }
    if (r < 0) {
        *coplanar = 1;
        return 0;
    }
    return r;
}

void orc_tri_tri_pairs(const orc_mesh *a, const orc_mesh *b, const int64_t *pairs, size_t n,
                       uint8_t *status, double *seg)
{
    for (size_t i = 0; i < n; ++i) {
        const int64_t *ta = a->T + 3 * pairs[2 * i];
        const int64_t *tb = b->T + 3 * pairs[2 * i + 1];
        double s[3] = {0, 0, 0}, t[3] = {0, 0, 0};
        int cop = 0;
        int hit = orc_tri_tri(a->V + 3 * ta[0], a->V + 3 * ta[1], a->V + 3 * ta[2],
                              b->V + 3 * tb[0], b->V + 3 * tb[1], b->V + 3 * tb[2], &cop, s, t);
        status[i] = cop ? 2 : (hit ? 1 : 0);
        for (int k = 0; k < 3; ++k) {
            seg[6 * i + k] = (hit && !cop) ? s[k] : 0.0;
            seg[6 * i + 3 + k] = (hit && !cop) ? t[k] : 0.0;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* ray parity (src/solidboolean.cpp:31-35, 48-92; src/vector3.h:264-280)      */
/* ------------------------------------------------------------------------- */
void orc_position_key(const double p[3], int64_t out[3])
{
    for (int i = 0; i < 3; ++i)
        out[i] = (int64_t)(long)(p[i] * (double)ORC_KEY_FACTOR);
}

static int key_cmp(const void *x, const void *y)
{
    const int64_t *a = (const int64_t *)x;
    const int64_t *b = (const int64_t *)y;
    for (int i = 0; i < 3; ++i) {
        if (a[i] < b[i])
            return -1;
        if (a[i] > b[i])
            return 1;
    }
    return 0;
}

typedef struct {
    int64_t *idx;
    size_t n, cap;
} idxbuf_t;

static void collect_ray(const orc_mesh *m, int64_t ni, const double *rmn, const double *rmx, idxbuf_t *ib)
{
    const node_t *f = &m->nodes[ni];
    if (!boxes_overlap(f->mn, f->mx, rmn, rmx))
        return;
    if (f->left < 0 && f->right < 0) {
        for (int64_t i = 0; i < f->count; ++i) {
            int64_t a = m->perm[f->start + i];
            const double *ba = m->boxes + 9 * a;
            if (boxes_overlap(ba, ba + 3, rmn, rmx)) {
                if (ib->n == ib->cap) {
                    ib->cap = ib->cap ? 2 * ib->cap : 64;
                    ib->idx = (int64_t *)realloc(ib->idx, sizeof(int64_t) * ib->cap);
                }
                ib->idx[ib->n++] = a;
            }
        }
        return;
    }
    collect_ray(m, f->left, rmn, rmx, ib);
    collect_ray(m, f->right, rmn, rmx, ib);
}

/* One candidate of isPointInMesh's loop (src/solidboolean.cpp:64-87): the
 * plane hit of the segment p -> end (vector3.h:264-280) and the three-normal
 * inside test; 1 with the hit point in x on a hit. */
static int ray_test(const double p[3], const double end[3], const double u[3], const double *v0, const double *v1,
                    const double *v2, const double *n, double x[3])
{
    double w[3];
    sub3(w, p, v0);
    double d = ORC_PLANE_DOT(n, u);
    double num = (-n[0]) * w[0] + (-n[1]) * w[1] + (-n[2]) * w[2];
    (void)end;
    if (fabs(d) <= DBL_EPSILON)
        return 0;
    double s = num / d;
    if (s < 0 || s > 1 || isnan(s) || isinf(s))
        return 0;
    x[0] = p[0] + s * u[0];
    x[1] = p[1] + s * u[1];
    x[2] = p[2] + s * u[2];
    const double *tv[3] = {v0, v1, v2};
    double nn[3][3];
    for (int i = 0; i < 3; ++i)
        orc_normal(x, tv[i], tv[(i + 1) % 3], nn[i]);
    return ORC_INSIDE_GT(dot3(nn[0], nn[1])) && ORC_INSIDE_GT(dot3(nn[0], nn[2]));
}

static void ray_segment(const double p[3], int axis, double end[3], double u[3], double rmn[3], double rmx[3])
{
    double dir[3] = {DBL_EPSILON, DBL_EPSILON, DBL_EPSILON};
    dir[axis] = DBL_MAX;
    for (int k = 0; k < 3; ++k) {
        end[k] = p[k] + dir[k];
        rmn[k] = DBL_MAX;
        rmx[k] = -DBL_MAX;
    }
    box_update(rmn, rmx, p);
    box_update(rmn, rmx, end);
    sub3(u, end, p);
}

int orc_ray_candidate(const double p[3], int axis, const double v0[3], const double v1[3], const double v2[3],
                      const double n[3], double hit[3])
{
    double end[3], u[3], rmn[3], rmx[3], bmn[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double bmx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    ray_segment(p, axis, end, u, rmn, rmx);
    box_update(bmn, bmx, v0);
    box_update(bmn, bmx, v1);
    box_update(bmn, bmx, v2);
    if (!boxes_overlap(bmn, bmx, rmn, rmx))
        return 0;
    return ray_test(p, end, u, v0, v1, v2, n, hit);
}

void orc_ray_candidates(size_t n, const double *pts, const int32_t *axis, const double *tris9, uint8_t *hit,
                        double *hits)
{
    for (size_t i = 0; i < n; ++i) {
        const double *t = tris9 + 9 * i;
        double nrm[3];
        orc_normal(t, t + 3, t + 6, nrm);
        hits[3 * i] = hits[3 * i + 1] = hits[3 * i + 2] = 0.0;
        hit[i] = (uint8_t)orc_ray_candidate(pts + 3 * i, axis[i], t, t + 3, t + 6, nrm, hits + 3 * i);
    }
}

int orc_point_in_mesh(const orc_mesh *m, const double p[3], int axis)
{
    double end[3], u[3], rmn[3], rmx[3];
    ray_segment(p, axis, end, u, rmn, rmx);
    idxbuf_t ib = {NULL, 0, 0};
    if (m->nnodes)
        collect_ray(m, 0, rmn, rmx, &ib);
    int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * 3 * (ib.n ? ib.n : 1));
    size_t nk = 0;
    for (size_t c = 0; c < ib.n; ++c) {
        int64_t t = ib.idx[c];
        double x[3];
        if (ray_test(p, end, u, m->V + 3 * m->T[3 * t], m->V + 3 * m->T[3 * t + 1], m->V + 3 * m->T[3 * t + 2],
                     m->normals + 3 * t, x)) {
            orc_position_key(x, keys + 3 * nk);
            ++nk;
        }
    }
    size_t unique = 0;
    if (nk) {
        qsort(keys, nk, sizeof(int64_t) * 3, key_cmp);
        unique = 1;
        for (size_t i = 1; i < nk; ++i)
            if (key_cmp(keys + 3 * i, keys + 3 * (i - 1)) != 0)
                ++unique;
    }
    free(keys);
    free(ib.idx);
    return (int)(unique % 2 != 0);
}

int orc_classify_point(const orc_mesh *m, const double p[3])
{
    int inside = 0;
    for (int axis = 0; axis < 3; ++axis)
        inside += orc_point_in_mesh(m, p, axis);
    return (float)inside / 3 > ORC_VOTE;
}

void orc_classify_triangles(const orc_mesh *self, const orc_mesh *other,
                            const uint8_t *intersected, uint8_t *out)
{
    for (size_t t = 0; t < self->nt; ++t) {
        if (intersected && intersected[t]) {
            out[t] = 2;
            continue;
        }
        const double *a = self->V + 3 * self->T[3 * t];
        const double *b = self->V + 3 * self->T[3 * t + 1];
        const double *c = self->V + 3 * self->T[3 * t + 2];
        double p[3];
        for (int i = 0; i < 3; ++i)
            p[i] = (a[i] + b[i] + c[i]) / 3.0;
        out[t] = (uint8_t)orc_classify_point(other, p);
    }
}
