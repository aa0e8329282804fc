// C entry points of the host library (libsolidboolean.so) for callers that
// cannot use the C++ classes directly (ctypes tests, other languages).
#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "earclip.h"
#include "objfile.h"
#include "referencetree.h"
#include "solidboolean.h"
#include "solidmesh.h"

namespace {

thread_local std::string g_hostErr;

struct HostMesh {
    std::vector<Vector3> vertices;
    std::vector<std::vector<size_t>> triangles;
    SolidMesh mesh;
};

struct HostBoolean {
    SolidBoolean sb;
    bool ok = false;
    std::vector<std::vector<size_t>> fetched[3];  // filled by the first sbh_boolean_fetch of each
    bool haveFetched[3] = {false, false, false};
    HostBoolean(const SolidMesh *a, const SolidMesh *b) : sb(a, b) {}
};

double secs(std::chrono::high_resolution_clock::time_point a, std::chrono::high_resolution_clock::time_point b)
{
    return std::chrono::duration<double>(b - a).count();
}

}  // namespace

extern "C" {

const char *sbh_last_error(void) { return g_hostErr.c_str(); }

// SolidMesh over copies of the given arrays, prepared on the GPU; NULL on error.
void *sbh_mesh_new(const double *xyz, size_t nv, const int64_t *tri, size_t nt)
{
    try {
        HostMesh *m = new HostMesh;
        m->vertices.resize(nv);
        for (size_t i = 0; i < nv; ++i)
            m->vertices[i] = Vector3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        m->triangles.resize(nt);
        for (size_t i = 0; i < nt; ++i)
            m->triangles[i] = {(size_t)tri[3 * i], (size_t)tri[3 * i + 1], (size_t)tri[3 * i + 2]};
        m->mesh.setVertices(&m->vertices);
        m->mesh.setTriangles(&m->triangles);
        m->mesh.prepare();
        return m;
    } catch (const std::exception &e) {
        g_hostErr = e.what();
        return nullptr;
    }
}

void sbh_mesh_free(void *p) { delete (HostMesh *)p; }

// SolidBoolean::combine() + fetch*; returns 1 / 0 as combine(), -1 on a GPU
// error.  flags: 1 SolidBoolean::setReferencePairOrder, 2 setRobustWeld.
int sbh_boolean_new_opts(void *a, void *b, unsigned flags, void **out)
{
    HostBoolean *h = new HostBoolean(&((HostMesh *)a)->mesh, &((HostMesh *)b)->mesh);
    *out = h;
    h->sb.setReferencePairOrder((flags & 1u) != 0);
    h->sb.setRobustWeld((flags & 2u) != 0);
    try {
        h->ok = h->sb.combine();
        return h->ok ? 1 : 0;
    } catch (const std::exception &e) {
        g_hostErr = e.what();
        return -1;
    }
}

int sbh_boolean_new(void *a, void *b, void **out) { return sbh_boolean_new_opts(a, b, 1u, out); }

void sbh_boolean_free(void *p) { delete (HostBoolean *)p; }

size_t sbh_boolean_vertices(void *p, double *out, size_t cap)
{
    const auto &v = ((HostBoolean *)p)->sb.resultVertices();
    for (size_t i = 0; i < v.size() && i < cap; ++i)
        for (int k = 0; k < 3; ++k)
            out[3 * i + k] = v[i][k];
    return v.size();
}

size_t sbh_boolean_fetch(void *p, int which, int64_t *out, size_t cap)
{
    HostBoolean *h = (HostBoolean *)p;
    if (which < 0 || which > 2)
        return 0;
    if (h->ok && !h->haveFetched[which]) {
        // fetch* appends, as the reference's do
        if (which == 0)
            h->sb.fetchUnion(h->fetched[0]);
        else if (which == 1)
            h->sb.fetchDiff(h->fetched[1]);
        else
            h->sb.fetchIntersect(h->fetched[2]);
        h->haveFetched[which] = true;
    }
    const auto &f = h->fetched[which];
    for (size_t i = 0; i < f.size() && i < cap; ++i)
        for (int k = 0; k < 3; ++k)
            out[3 * i + k] = (int64_t)f[i][k];
    return f.size();
}

size_t sbh_boolean_triangle_sides(void *p, int which, uint8_t *out, size_t cap)
{
    const auto s = ((HostBoolean *)p)->sb.triangleSides(which);
    for (size_t i = 0; i < s.size() && i < cap; ++i)
        out[i] = s[i];
    return s.size();
}

// 5 input-vertex indices per result vertex (SolidBoolean::vertexSources);
// returns the result vertex count, or (size_t)-1 if a source was not found
size_t sbh_boolean_vertex_sources(void *p, uint32_t *out, size_t cap)
{
    std::vector<uint32_t> src;
    if (!((HostBoolean *)p)->sb.vertexSources(src))
        return (size_t)-1;
    for (size_t i = 0; i < src.size() && i < 5 * cap; ++i)
        out[i] = src[i];
    return src.size() / 5;
}

// stage durations (s) in the order of the reference's bench time points
void sbh_boolean_stage_seconds(void *p, double *out)
{
    const SolidBoolean &s = ((HostBoolean *)p)->sb;
    out[0] = secs(s.benchBegin_searchPotentialIntersectedPairs, s.benchEnd_searchPotentialIntersectedPairs);
    out[1] = secs(s.benchBegin_processPotentialIntersectedPairs, s.benchEnd_processPotentialIntersectedPairs);
    out[2] = secs(s.benchBegin_addUnintersectedTriangles, s.benchEnd_addUnintersectedTriangles);
    out[3] = secs(s.benchBegin_reTriangulate, s.benchEnd_reTriangulate);
    out[4] = secs(s.benchBegin_buildPolygonsFromEdges, s.benchEnd_buildPolygonsFromEdges);
    out[5] = secs(s.benchBegin_buildFaceGroups, s.benchEnd_buildFaceGroups);
    out[6] = secs(s.benchBegin_decideGroupSide, s.benchEnd_decideGroupSide);
}

// Triangulates rings (outer first) given as concatenated xy pairs with
// sizes[k] points in ring k.  Returns the index count; writes <= cap indices.
size_t sbh_earclip(const double *xy, const int64_t *sizes, size_t nrings, int64_t *out, size_t cap)
{
    std::vector<sbhost::Ring> rings(nrings);
    size_t k = 0;
    for (size_t r = 0; r < nrings; ++r)
        for (int64_t i = 0; i < sizes[r]; ++i, ++k)
            rings[r].push_back({xy[2 * k], xy[2 * k + 1]});
    std::vector<size_t> idx = sbhost::earclip(rings);
    for (size_t i = 0; i < idx.size() && i < cap; ++i)
        out[i] = (int64_t)idx[i];
    return idx.size();
}

// Reads an OBJ file as the reference's test driver does (objfile.h).  Call
// with capv = capt = 0 for the counts, then with room for them (xyz: 3 nv
// doubles, tris: 3 nt indices).  1 on success; 0 on failure, with the message
// in sbh_last_error().
int sbh_obj_load(const char *path, double *xyz, size_t capv, size_t *nv, int64_t *tris, size_t capt, size_t *nt)
{
    try {
        std::vector<Vector3> v;
        std::vector<std::vector<size_t>> t;
        std::string err;
        if (!path || !sbobj::loadObj(path, v, t, &err)) {
            g_hostErr = path ? err : "null path";
            return 0;
        }
        if (nv)
            *nv = v.size();
        if (nt)
            *nt = t.size();
        for (size_t i = 0; i < v.size() && i < capv; ++i)
            for (int k = 0; k < 3; ++k)
                xyz[3 * i + k] = v[i].constData()[k];
        for (size_t i = 0; i < t.size() && i < capt; ++i)
            for (int k = 0; k < 3; ++k)
                tris[3 * i + k] = (int64_t)t[i][k];
        return 1;
    } catch (const std::exception &e) {
        g_hostErr = e.what();
        return 0;
    }
}

// Writes nv vertices and nf triangles as test/main.cpp:9-28 does.  1 / 0.
int sbh_obj_export(const char *path, const double *xyz, size_t nv, const int64_t *tris, size_t nf)
{
    try {
        std::vector<Vector3> v(nv);
        for (size_t i = 0; i < nv; ++i)
            v[i] = Vector3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        std::vector<std::vector<size_t>> f(nf);
        for (size_t i = 0; i < nf; ++i)
            f[i] = {(size_t)tris[3 * i], (size_t)tris[3 * i + 1], (size_t)tris[3 * i + 2]};
        if (!path || !sbobj::exportObject(path, v, f)) {
            g_hostErr = "cannot write the file";
            return 0;
        }
        return 1;
    } catch (const std::exception &e) {
        g_hostErr = e.what();
        return 0;
    }
}

// The permutation that puts the n pairs ab (a of mesh A, b of mesh B) into
// the order the reference's search reports them (referencetree.h).  Host only.
// 1 on success, 0 on failure.
int sbh_reference_pair_order(const double *xyzA, size_t nvA, const int64_t *triA, size_t ntA, const double *xyzB,
                             size_t nvB, const int64_t *triB, size_t ntB, const uint32_t *ab, size_t n,
                             uint64_t *perm)
{
    try {
        auto mesh = [](const double *xyz, size_t nv, const int64_t *tri, size_t nt, std::vector<Vector3> &v,
                       std::vector<std::vector<size_t>> &t) {
            v.resize(nv);
            for (size_t i = 0; i < nv; ++i)
                v[i] = Vector3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
            t.resize(nt);
            for (size_t i = 0; i < nt; ++i) {
                t[i] = {(size_t)tri[3 * i], (size_t)tri[3 * i + 1], (size_t)tri[3 * i + 2]};
                for (size_t k : t[i])
                    if (k >= nv)
                        throw std::out_of_range("triangle index out of range");
            }
        };
        std::vector<Vector3> va, vb;
        std::vector<std::vector<size_t>> ta, tb;
        mesh(xyzA, nvA, triA, ntA, va, ta);
        mesh(xyzB, nvB, triB, ntB, vb, tb);
        for (size_t i = 0; i < n; ++i)
            if (ab[2 * i] >= ntA || ab[2 * i + 1] >= ntB)
                throw std::out_of_range("pair index out of range");
        const sbhost::ReferenceTree A(va, ta), B(vb, tb);
        const std::vector<size_t> order = sbhost::referencePairOrder(A, B, ab, n);
        for (size_t i = 0; i < n; ++i)
            perm[i] = order[i];
        return 1;
    } catch (const std::exception &e) {
        g_hostErr = e.what();
        return 0;
    }
}

// The reference's own search on the host (no GPU): both meshes' trees built
// by AxisAlignedBoudingBoxTree's public constructor over their triangle boxes
// and all-vertex boxes (src/solidmesh.cpp:57-75), then test(root, root,
// boxes, pairs) (src/solidboolean.cpp:94-101).  Writes up to `cap` pairs
// (a, b) into ab and returns how many there are; (size_t)-1 on bad input.
size_t sbh_tree_pairs(const double *xyzA, size_t nvA, const int64_t *triA, size_t ntA, const double *xyzB,
                      size_t nvB, const int64_t *triB, size_t ntB, uint64_t *ab, size_t cap)
{
    try {
        struct Side {
            std::vector<AxisAlignedBoudingBox> boxes;
            std::vector<size_t> indices;
            AxisAlignedBoudingBox all;
        };
        auto side = [](const double *xyz, size_t nv, const int64_t *tri, size_t nt, Side &s) {
            s.boxes.resize(nt);
            s.indices.resize(nt);
            for (size_t i = 0; i < nt; ++i) {
                for (size_t k = 0; k < 3; ++k) {
                    const int64_t v = tri[3 * i + k];
                    if (v < 0 || (size_t)v >= nv)
                        throw std::out_of_range("triangle index out of range");
                    const Vector3 p(xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]);
                    s.boxes[i].update(p);
                    s.all.update(p);
                }
                s.boxes[i].updateCenter();
                s.indices[i] = i;
            }
            s.all.updateCenter();
        };
        Side a, b;
        side(xyzA, nvA, triA, ntA, a);
        side(xyzB, nvB, triB, ntB, b);
        const AxisAlignedBoudingBoxTree ta(&a.boxes, a.indices, a.all), tb(&b.boxes, b.indices, b.all);
        std::vector<std::pair<size_t, size_t>> pairs;
        ta.test(ta.root(), tb.root(), &b.boxes, &pairs);
        for (size_t i = 0; i < pairs.size() && i < cap; ++i) {
            ab[2 * i] = pairs[i].first;
            ab[2 * i + 1] = pairs[i].second;
        }
        return pairs.size();
    } catch (const std::exception &e) {
        g_hostErr = e.what();
        return (size_t)-1;
    }
}

}  // extern "C"
