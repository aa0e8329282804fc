// Test infrastructure only: a C-ABI shim over the reference's own sources,
// compiled from /root/reference by oracle/Makefile into oracle/_ref/libsbref.so.
// It is used as a checker (golden-vector generation, CPU baseline timing) and is
// never linked into, or called by, the product path.
//
// Private members of SolidBoolean (isPointInMesh, groups, sides, intersected
// face sets) are reached by compiling this translation unit with `private`
// defined as `public`, after every standard / Eigen header is already included.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <limits>
#include <map>
#include <queue>
#include <set>
#include <sstream>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#include <Eigen/Dense>

#define private public
#include "solidboolean.h"
#undef private
#include "tri_tri_intersect.h"

#include <array>
#include "earcut.hpp"

#define TINYOBJLOADER_IMPLEMENTATION
#include "tiny_obj_loader.h"

namespace {

struct RefMesh {
    std::vector<Vector3> vertices;
    std::vector<std::vector<size_t>> triangles;
    SolidMesh mesh;
    double prepareSeconds = 0.0;
};

struct RefBoolean {
    RefMesh *a = nullptr;
    RefMesh *b = nullptr;
    SolidBoolean *sb = nullptr;
    std::vector<std::vector<size_t>> fetched[3];
    bool combined = false;
    bool ok = false;
    double combineSeconds = 0.0;
};

// The three ray directions of solidboolean.cpp:31-35 (file-static there).
const Vector3 kAxes[3] = {
    {std::numeric_limits<double>::max(), std::numeric_limits<double>::epsilon(), std::numeric_limits<double>::epsilon()},
    {std::numeric_limits<double>::epsilon(), std::numeric_limits<double>::max(), std::numeric_limits<double>::epsilon()},
    {std::numeric_limits<double>::epsilon(), std::numeric_limits<double>::epsilon(), std::numeric_limits<double>::max()},
};

double seconds(std::chrono::high_resolution_clock::time_point a,
               std::chrono::high_resolution_clock::time_point b)
{
    return std::chrono::duration<double>(b - a).count();
}

}  // namespace

extern "C" {

// Builds a SolidMesh from flat arrays and runs SolidMesh::prepare()
// (solidmesh.cpp:42-76).
void *sbref_mesh_new(const double *xyz, size_t nv, const int64_t *tri, size_t nt)
{
    RefMesh *m = new RefMesh;
    m->vertices.resize(nv);
    for (size_t i = 0; i < nv; ++i)
        m->vertices[i] = Vector3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    m->triangles.resize(nt);
    for (size_t i = 0; i < nt; ++i)
        m->triangles[i] = {(size_t)tri[3 * i], (size_t)tri[3 * i + 1], (size_t)tri[3 * i + 2]};
    m->mesh.setVertices(&m->vertices);
    m->mesh.setTriangles(&m->triangles);
    auto t0 = std::chrono::high_resolution_clock::now();
    m->mesh.prepare();
    auto t1 = std::chrono::high_resolution_clock::now();
    m->prepareSeconds = seconds(t0, t1);
    return m;
}

void sbref_mesh_free(void *p) { delete (RefMesh *)p; }

double sbref_mesh_prepare_seconds(void *p) { return ((RefMesh *)p)->prepareSeconds; }

// Per-triangle normals (vector3.h:155-176 via solidmesh.cpp:47-55), n x 3.
void sbref_mesh_normals(void *p, double *out)
{
    RefMesh *m = (RefMesh *)p;
    const auto &n = *m->mesh.triangleNormals();
    for (size_t i = 0; i < n.size(); ++i)
        for (int k = 0; k < 3; ++k)
            out[3 * i + k] = n[i][k];
}

// Per-triangle boxes (axisalignedboundingbox.h:31-53): n x 9 = min, max, center.
void sbref_mesh_boxes(void *p, double *out)
{
    RefMesh *m = (RefMesh *)p;
    const auto &bx = *m->mesh.triangleAxisAlignedBoundingBoxes();
    for (size_t i = 0; i < bx.size(); ++i)
        for (int k = 0; k < 3; ++k) {
            out[9 * i + k] = bx[i].lowerBound()[k];
            out[9 * i + 3 + k] = bx[i].upperBound()[k];
            out[9 * i + 6 + k] = bx[i].center()[k];
        }
}

static void treeStats(const AxisAlignedBoudingBoxTree::Node *n, size_t depth, int64_t *st)
{
    if (!n)
        return;
    st[0] += 1;
    if (n->isLeaf()) {
        st[1] += 1;
        if ((int64_t)depth > st[2])
            st[2] = depth;
        if ((int64_t)n->boxIndices.size() > st[3])
            st[3] = n->boxIndices.size();
        return;
    }
    treeStats(n->left, depth + 1, st);
    treeStats(n->right, depth + 1, st);
}

// nodes, leaves, max depth, max leaf size of the reference tree.
void sbref_mesh_tree_stats(void *p, int64_t *st)
{
    st[0] = st[1] = st[2] = st[3] = 0;
    treeStats(((RefMesh *)p)->mesh.axisAlignedBoundingBoxTree()->root(), 0, st);
}

// Candidate pairs in the reference's DFS order (solidboolean.cpp:94-101,
// axisalignedboundingboxtree.h:54-95). Returns the count; writes at most cap pairs.
size_t sbref_find_pairs(void *pa, void *pb, int64_t *out, size_t cap, double *secondsOut)
{
    RefMesh *a = (RefMesh *)pa;
    RefMesh *b = (RefMesh *)pb;
    std::vector<std::pair<size_t, size_t>> pairs;
    auto t0 = std::chrono::high_resolution_clock::now();
    const AxisAlignedBoudingBoxTree *lt = a->mesh.axisAlignedBoundingBoxTree();
    const AxisAlignedBoudingBoxTree *rt = b->mesh.axisAlignedBoundingBoxTree();
    lt->test(lt->root(), rt->root(), b->mesh.triangleAxisAlignedBoundingBoxes(), &pairs);
    auto t1 = std::chrono::high_resolution_clock::now();
    if (secondsOut)
        *secondsOut = seconds(t0, t1);
    for (size_t i = 0; i < pairs.size() && i < cap; ++i) {
        out[2 * i] = (int64_t)pairs[i].first;
        out[2 * i + 1] = (int64_t)pairs[i].second;
    }
    return pairs.size();
}

// The vendored Guigue-Devillers routine on one triangle pair
// (tri_tri_intersect.c:395-472). seg = source[3], target[3].
int sbref_tri_tri(const double *a9, const double *b9, int *coplanar, double *seg)
{
    double p1[3], q1[3], r1[3], p2[3], q2[3], r2[3];
    std::memcpy(p1, a9, 24); std::memcpy(q1, a9 + 3, 24); std::memcpy(r1, a9 + 6, 24);
    std::memcpy(p2, b9, 24); std::memcpy(q2, b9 + 3, 24); std::memcpy(r2, b9 + 6, 24);
    int cop = 0;
    int r = tri_tri_intersection_test_3d(p1, q1, r1, p2, q2, r2, &cop, seg, seg + 3);
    *coplanar = cop;
    return r;
}

// intersectTwoFaces (solidboolean.cpp:103-122) over a pair list.
// status: 0 = no intersection, 1 = hit, 2 = coplanar (returns false in the reference).
// raw: tri_tri_intersection_test_3d's own return value.
void sbref_tri_tri_pairs(void *pa, void *pb, const int64_t *pairs, size_t n,
                         uint8_t *status, uint8_t *raw, double *seg, double *secondsOut)
{
    RefMesh *a = (RefMesh *)pa;
    RefMesh *b = (RefMesh *)pb;
    SolidBoolean sb(&a->mesh, &b->mesh);
    auto t0 = std::chrono::high_resolution_clock::now();
    for (size_t i = 0; i < n; ++i) {
        std::pair<Vector3, Vector3> e;
        bool hit = sb.intersectTwoFaces((size_t)pairs[2 * i], (size_t)pairs[2 * i + 1], e);
        status[i] = hit ? 1 : 0;
        for (int k = 0; k < 3; ++k) {
            seg[6 * i + k] = e.first[k];
            seg[6 * i + 3 + k] = e.second[k];
        }
    }
    auto t1 = std::chrono::high_resolution_clock::now();
    if (secondsOut)
        *secondsOut = seconds(t0, t1);
    // coplanar flag / raw return, outside the timed loop
    for (size_t i = 0; i < n; ++i) {
        const auto &fa = a->triangles[(size_t)pairs[2 * i]];
        const auto &fb = b->triangles[(size_t)pairs[2 * i + 1]];
        double s[3], t[3];
        int cop = 0;
        int r = tri_tri_intersection_test_3d(
            (double *)a->vertices[fa[0]].constData(), (double *)a->vertices[fa[1]].constData(),
            (double *)a->vertices[fa[2]].constData(), (double *)b->vertices[fb[0]].constData(),
            (double *)b->vertices[fb[1]].constData(), (double *)b->vertices[fb[2]].constData(),
            &cop, s, t);
        if (raw)
            raw[i] = (uint8_t)r;
        if (cop && status[i] == 0)
            status[i] = 2;
    }
}

// isPointInMesh (solidboolean.cpp:48-92) for n query points along axis 0, 1 or 2
// (solidboolean.cpp:31-35), against `target` with its own tree.
void sbref_points_in_mesh(void *ptarget, const double *pts, size_t n, int axis, uint8_t *out,
                          double *secondsOut)
{
    RefMesh *t = (RefMesh *)ptarget;
    SolidBoolean sb(&t->mesh, &t->mesh);
    auto t0 = std::chrono::high_resolution_clock::now();
    for (size_t i = 0; i < n; ++i) {
        Vector3 p(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        out[i] = sb.isPointInMesh(p, &t->mesh, t->mesh.axisAlignedBoundingBoxTree(),
                                  kAxes[axis]) ? 1 : 0;
    }
    auto t1 = std::chrono::high_resolution_clock::now();
    if (secondsOut)
        *secondsOut = seconds(t0, t1);
}

// Full boolean (SolidBoolean::combine, solidboolean.cpp:288-480) + fetch.
void *sbref_boolean_new(void *pa, void *pb)
{
    RefBoolean *r = new RefBoolean;
    r->a = (RefMesh *)pa;
    r->b = (RefMesh *)pb;
    r->sb = new SolidBoolean(&r->a->mesh, &r->b->mesh);
    return r;
}

void sbref_boolean_free(void *p)
{
    RefBoolean *r = (RefBoolean *)p;
    delete r->sb;
    delete r;
}

// Returns combine()'s bool. Fetches union/diff/intersect (solidboolean.cpp:512-565)
// only when combine() succeeded.
int sbref_boolean_combine(void *p, int quiet)
{
    RefBoolean *r = (RefBoolean *)p;
    std::streambuf *old = nullptr;
    std::ostringstream sink;
    if (quiet)
        old = std::cout.rdbuf(sink.rdbuf());
    auto t0 = std::chrono::high_resolution_clock::now();
    r->ok = r->sb->combine();
    auto t1 = std::chrono::high_resolution_clock::now();
    r->combineSeconds = seconds(t0, t1);
    if (r->ok) {
        r->sb->fetchUnion(r->fetched[0]);
        r->sb->fetchDiff(r->fetched[1]);
        r->sb->fetchIntersect(r->fetched[2]);
    }
    if (quiet)
        std::cout.rdbuf(old);
    r->combined = true;
    return r->ok ? 1 : 0;
}

double sbref_boolean_seconds(void *p) { return ((RefBoolean *)p)->combineSeconds; }

// Stage durations in seconds, in the order of solidboolean.h:45-58 (7 stages).
void sbref_boolean_stage_seconds(void *p, double *out)
{
    SolidBoolean *s = ((RefBoolean *)p)->sb;
    out[0] = seconds(s->benchBegin_searchPotentialIntersectedPairs, s->benchEnd_searchPotentialIntersectedPairs);
    out[1] = seconds(s->benchBegin_processPotentialIntersectedPairs, s->benchEnd_processPotentialIntersectedPairs);
    out[2] = seconds(s->benchBegin_addUnintersectedTriangles, s->benchEnd_addUnintersectedTriangles);
    out[3] = seconds(s->benchBegin_reTriangulate, s->benchEnd_reTriangulate);
    out[4] = seconds(s->benchBegin_buildPolygonsFromEdges, s->benchEnd_buildPolygonsFromEdges);
    out[5] = seconds(s->benchBegin_buildFaceGroups, s->benchEnd_buildFaceGroups);
    out[6] = seconds(s->benchBegin_decideGroupSide, s->benchEnd_decideGroupSide);
}

size_t sbref_boolean_vertices(void *p, double *out, size_t cap)
{
    const auto &v = ((RefBoolean *)p)->sb->resultVertices();
    for (size_t i = 0; i < v.size() && i < cap; ++i)
        for (int k = 0; k < 3; ++k)
            out[3 * i + k] = v[i][k];
    return v.size();
}

// which: 0 union, 1 diff, 2 intersect.
size_t sbref_boolean_fetch(void *p, int which, int64_t *out, size_t cap)
{
    const auto &f = ((RefBoolean *)p)->fetched[which];
    for (size_t i = 0; i < f.size() && i < cap; ++i)
        for (int k = 0; k < 3; ++k)
            out[3 * i + k] = (int64_t)f[i][k];
    return f.size();
}

// Intersected face indices of mesh 0 (first) or 1 (second), sorted.
size_t sbref_boolean_intersected(void *p, int which, int64_t *out, size_t cap)
{
    SolidBoolean *s = ((RefBoolean *)p)->sb;
    const auto &set = which == 0 ? s->m_firstIntersectedFaces : s->m_secondIntersectedFaces;
    std::vector<size_t> v(set.begin(), set.end());
    std::sort(v.begin(), v.end());
    for (size_t i = 0; i < v.size() && i < cap; ++i)
        out[i] = (int64_t)v[i];
    return v.size();
}

// Groups of mesh 0/1: count; sizes[g]; first triangle (new-triangle index) of each
// group; sides[g] (cap-limited).
size_t sbref_boolean_groups(void *p, int which, int64_t *sizes, int64_t *firsts, uint8_t *sides,
                            size_t cap)
{
    SolidBoolean *s = ((RefBoolean *)p)->sb;
    const auto &g = which == 0 ? s->m_firstTriangleGroups : s->m_secondTriangleGroups;
    const auto &sd = which == 0 ? s->m_firstGroupSides : s->m_secondGroupSides;
    for (size_t i = 0; i < g.size() && i < cap; ++i) {
        if (sizes)
            sizes[i] = (int64_t)g[i].size();
        if (firsts)
            firsts[i] = g[i].empty() ? -1 : (int64_t)g[i][0];
        if (sides)
            sides[i] = (i < sd.size() && sd[i]) ? 1 : 0;
    }
    return g.size();
}

// Members (new-triangle indices) of every group of mesh 0/1, flattened in group
// order; sizes come from sbref_boolean_groups.
size_t sbref_boolean_group_members(void *p, int which, int64_t *out, size_t cap)
{
    SolidBoolean *s = ((RefBoolean *)p)->sb;
    const auto &g = which == 0 ? s->m_firstTriangleGroups : s->m_secondTriangleGroups;
    size_t k = 0;
    for (const auto &grp : g)
        for (size_t t : grp) {
            if (k < cap)
                out[k] = (int64_t)t;
            ++k;
        }
    return k;
}

// m_newTriangles (all, in creation order).
size_t sbref_boolean_new_triangles(void *p, int64_t *out, size_t cap)
{
    const auto &t = ((RefBoolean *)p)->sb->m_newTriangles;
    for (size_t i = 0; i < t.size() && i < cap; ++i)
        for (int k = 0; k < 3; ++k)
            out[3 * i + k] = (int64_t)t[i][k];
    return t.size();
}

// The vendored earcut (thirdparty/earcut.hpp) on rings given as xy pairs:
// ring k has sizes[k] points.  Returns the index count; writes <= cap indices.
size_t sbref_earcut(const double *xy, const int64_t *sizes, size_t nrings, int64_t *out, size_t cap)
{
    std::vector<std::vector<std::array<double, 2>>> rings(nrings);
    size_t k = 0;
    for (size_t r = 0; r < nrings; ++r)
        for (int64_t i = 0; i < sizes[r]; ++i, ++k)
            rings[r].push_back({xy[2 * k], xy[2 * k + 1]});
    std::vector<size_t> idx = mapbox::earcut<size_t>(rings);
    for (size_t i = 0; i < idx.size() && i < cap; ++i)
        out[i] = (int64_t)idx[i];
    return idx.size();
}

// OBJ loading exactly as the reference driver does (test/main.cpp:32-71, tinyobj
// with float coordinates and its own polygon triangulation).
int sbref_load_obj(const char *path, double *xyz, size_t capv, size_t *nv,
                   int64_t *tri, size_t capt, size_t *nt)
{
    tinyobj::attrib_t attributes;
    std::vector<tinyobj::shape_t> shapes;
    std::vector<tinyobj::material_t> materials;
    std::string warn, err;
    if (!tinyobj::LoadObj(&attributes, &shapes, &materials, &warn, &err, path))
        return 0;
    size_t v = attributes.vertices.size() / 3;
    *nv = v;
    for (size_t i = 0; i < v && i < capv; ++i)
        for (int k = 0; k < 3; ++k)
            xyz[3 * i + k] = (double)attributes.vertices[3 * i + k];
    size_t t = 0;
    for (const auto &shape : shapes)
        for (size_t i = 0; i + 2 < shape.mesh.indices.size(); i += 3) {
            if (t < capt)
                for (int k = 0; k < 3; ++k)
                    tri[3 * t + k] = shape.mesh.indices[i + k].vertex_index;
            ++t;
        }
    *nt = t;
    return 1;
}

}  // extern "C"
