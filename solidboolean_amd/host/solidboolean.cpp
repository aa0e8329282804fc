// SolidBoolean; see solidboolean.h.  Stage names and their order follow
// src/solidboolean.cpp:288-480.
#include "solidboolean.h"

#include <algorithm>
#include <cstring>

#include <deque>
#include <stdexcept>

#include "referencetree.h"
#include "retriangulator.h"
#include "sbhost_log.h"

using Clock = std::chrono::high_resolution_clock;

// ---------------------------------------------------------------------------
// HalfEdgeMap: linear probing, power-of-two capacity, load <= 1/2
// ---------------------------------------------------------------------------
static constexpr uint64_t kEmpty = ~0ull;

static inline size_t slot_of(uint64_t key, size_t mask)
{
    return (size_t)((key * 0x9E3779B97F4A7C15ull) >> 17) & mask;
}

void SolidBoolean::HalfEdgeMap::reserve(size_t n)
{
    size_t cap = 16;
    while (cap < 2 * n)
        cap <<= 1;
    if (cap <= m_slots.size())
        return;
    std::vector<Slot> slots(cap, Slot{kEmpty, 0});
    const size_t mask = cap - 1;
    for (const Slot &x : m_slots)
        if (x.key != kEmpty) {
            size_t s = slot_of(x.key, mask);
            while (slots[s].key != kEmpty)
                s = (s + 1) & mask;
            slots[s] = x;
        }
    m_slots.swap(slots);
    m_mask = mask;
}

void SolidBoolean::HalfEdgeMap::grow() { reserve(m_slots.empty() ? 8 : m_slots.size()); }

bool SolidBoolean::HalfEdgeMap::insert(uint64_t key, size_t value)
{
    if (2 * (m_count + 1) > m_slots.size())
        grow();
    size_t s = slot_of(key, m_mask);
    while (m_slots[s].key != kEmpty) {
        if (m_slots[s].key == key)
            return false;
        s = (s + 1) & m_mask;
    }
    m_slots[s] = Slot{key, value};
    ++m_count;
    return true;
}

const size_t *SolidBoolean::HalfEdgeMap::find(uint64_t key) const
{
    if (m_slots.empty())
        return nullptr;
    size_t s = slot_of(key, m_mask);
    while (m_slots[s].key != kEmpty) {
        if (m_slots[s].key == key)
            return &m_slots[s].value;
        s = (s + 1) & m_mask;
    }
    return nullptr;
}

void SolidBoolean::HalfEdgeMap::prefetch(uint64_t key) const
{
    if (!m_slots.empty())
        __builtin_prefetch(&m_slots[slot_of(key, m_mask)]);
}

// ---------------------------------------------------------------------------

SolidBoolean::SolidBoolean(const SolidMesh *firstMesh, const SolidMesh *secondMesh)
    : m_firstMesh(firstMesh), m_secondMesh(secondMesh)
{
}

SolidBoolean::~SolidBoolean() = default;

size_t SolidBoolean::addNewPoint(const Vector3 &position, int64_t origin)
{
    if (m_weld == Weld::Crossings && origin >= 0 && (size_t)origin < m_endpointKeys.size()) {
        auto r = m_newKeyMap.insert({m_endpointKeys[(size_t)origin], m_newVertices.size()});
        if (r.second) {
            m_newVertices.push_back(position);
            m_vertexSource.push_back(-origin - 1);
        }
        return r.first->second;
    }
    auto r = m_newPositionMap.insert({PositionKey(position), m_newVertices.size()});
    if (r.second) {
        m_newVertices.push_back(position);
        m_vertexSource.push_back(-origin - 1);
    }
    return r.first->second;
}

bool SolidBoolean::addUnintersectedTriangles(const SolidMesh *mesh, const std::vector<uint8_t> &used,
                                             HalfEdgeMap *halfEdges, size_t inputOffset)
{
    // (reference: src/solidboolean.cpp:250-286; a repeated half-edge stops the copy)
    const size_t oldVertexCount = m_newVertices.size();
    const auto &vertices = *mesh->vertices();
    m_newVertices.insert(m_newVertices.end(), vertices.begin(), vertices.end());
    for (size_t i = 0; i < vertices.size(); ++i)
        m_vertexSource.push_back((int64_t)(inputOffset + i));
    const auto &tris = *mesh->triangles();
    m_newTriangles.reserve(m_newTriangles.size() + tris.size());
    if (!halfEdges) {
        // the GPU builds the half-edge table later
        for (size_t i = 0; i < tris.size(); ++i)
            if (!used[i])
                m_newTriangles.push_back(
                    {tris[i][0] + oldVertexCount, tris[i][1] + oldVertexCount, tris[i][2] + oldVertexCount});
        return true;
    }
    halfEdges->reserve(3 * tris.size());
    constexpr size_t kAhead = 16;  // slots prefetched this many triangles ahead
    for (size_t i = 0; i < tris.size(); ++i) {
        if (i + kAhead < tris.size() && !used[i + kAhead]) {
            const auto &f = tris[i + kAhead];
            const size_t fa = f[0] + oldVertexCount, fb = f[1] + oldVertexCount, fc = f[2] + oldVertexCount;
            halfEdges->prefetch(halfEdgeKey(fa, fb));
            halfEdges->prefetch(halfEdgeKey(fb, fc));
            halfEdges->prefetch(halfEdgeKey(fc, fa));
        }
        if (used[i])
            continue;
        const size_t index = m_newTriangles.size();
        const size_t a = tris[i][0] + oldVertexCount, b = tris[i][1] + oldVertexCount, c = tris[i][2] + oldVertexCount;
        m_newTriangles.push_back({a, b, c});
        if (!halfEdges->insert(halfEdgeKey(a, b), index) || !halfEdges->insert(halfEdgeKey(b, c), index) ||
            !halfEdges->insert(halfEdgeKey(c, a), index)) {
            sbhost::diag("Found repeated halfedge");
            return false;
        }
    }
    return true;
}

bool SolidBoolean::reTriangulate(const ContextMap &contexts, const SolidMesh *mesh, size_t startOldVertex,
                                 EdgeMap &edges, HalfEdgeMap *halfEdges)
{
    // (reference: src/solidboolean.cpp:352-407)
    const auto &tris = *mesh->triangles();
    const auto &verts = *mesh->vertices();
    const std::vector<Vector3> *normals = contexts.empty() ? nullptr : mesh->triangleNormals();
    for (const auto &kv : contexts) {
        const auto &t = tris[kv.first];
        ReTriangulator rt({verts[t[0]], verts[t[1]], verts[t[2]]}, (*normals)[kv.first]);
        rt.setEdges(kv.second.points, &kv.second.neighborMap);
        std::vector<int> hints;
        if (m_weld != Weld::Cells) {
            // a cut point made on one of this triangle's own edges lies on it
            // (with hinted cells: any crossing welded into the point; the first
            // edge of the triangle that one lies on)
            hints.assign(kv.second.points.size(), -1);
            auto edgeOf = [&](int64_t o) {
                if (o < 0 || (size_t)o >= m_endpointKeys.size() || m_endpointKeys[(size_t)o].edge == ~0ull)
                    return -1;
                const uint64_t e = m_endpointKeys[(size_t)o].edge;
                for (int i = 0; i < 3; ++i) {
                    const uint64_t x = startOldVertex + t[i], y = startOldVertex + t[(i + 1) % 3];
                    if (e == (std::min(x, y) << 32 | std::max(x, y)))
                        return i;
                }
                return -1;
            };
            for (size_t k = 0; k < hints.size(); ++k) {
                if (m_weld == Weld::Crossings) {
                    hints[k] = edgeOf(kv.second.origin[k]);
                    continue;
                }
                for (int64_t o : kv.second.merged[k]) {
                    const int i = edgeOf(o);
                    if (i >= 0 && (hints[k] < 0 || i < hints[k]))
                        hints[k] = i;
                }
            }
            rt.setEdgeHints(&hints);
        }
        if (!rt.reTriangulate()) {
            sbhost::diag("Retriangle failed");
            return false;
        }
        std::vector<size_t> newIndices;
        newIndices.reserve(3 + kv.second.points.size());
        newIndices.push_back(startOldVertex + t[0]);
        newIndices.push_back(startOldVertex + t[1]);
        newIndices.push_back(startOldVertex + t[2]);
        for (size_t k = 0; k < kv.second.points.size(); ++k)
            newIndices.push_back(addNewPoint(kv.second.points[k], kv.second.origin[k]));
        for (const auto &nt : rt.triangles()) {
            const size_t index = m_newTriangles.size();
            const size_t a = newIndices[nt[0]], b = newIndices[nt[1]], c = newIndices[nt[2]];
            m_newTriangles.push_back({a, b, c});
            if (!halfEdges)
                continue;
            for (const uint64_t he : {halfEdgeKey(a, b), halfEdgeKey(b, c), halfEdgeKey(c, a)})
                if (!halfEdges->insert(he, index)) {
                    m_repeatedHalfEdge = true;
                    sbhost::diag("Found repeated halfedge");
                }
        }
        for (const auto &nb : kv.second.neighborMap) {
            const size_t from = newIndices[nb.first];
            for (size_t k : nb.second) {
                const size_t to = newIndices[k];
                edges[from].insert(to);
                edges[to].insert(from);
            }
        }
    }
    return true;
}

bool SolidBoolean::buildPolygonsFromEdges(const EdgeMap &edges, std::vector<std::vector<size_t>> &polygons)
{
    // (reference: src/solidboolean.cpp:124-165)
    std::unordered_set<size_t> visited;
    for (const auto &kv : edges) {
        const size_t start = kv.first;
        if (visited.count(start))
            continue;
        std::vector<size_t> loop;
        size_t cur = start;
        for (;;) {
            visited.insert(cur);
            loop.push_back(cur);
            auto it = edges.find(cur);
            if (it == edges.end())
                break;
            bool stepped = false;
            for (size_t nb : it->second)
                if (!visited.count(nb)) {
                    cur = nb;
                    stepped = true;
                    break;
                }
            if (!stepped)
                break;
        }
        if (loop.size() <= 2) {
            sbhost::diag("buildPolygonsFromEdges failed, too short");
            return false;
        }
        auto last = edges.find(loop.back());
        if (!last->second.count(start)) {
            sbhost::diag("buildPolygonsFromEdges failed, could not form a ring");
            return false;
        }
        polygons.push_back(loop);
    }
    return true;
}

// Half-edge tables of one mesh's new triangles (its untouched ones, then its
// re-triangulated ones) built on the GPU: per local triangle the local
// triangle across each edge, the owners of the loop half-edges (forward, then
// backward, per loop edge in loop order), and the first triangle at which a
// directed edge repeats.  False when the indices do not fit the device table.
bool SolidBoolean::gpuHalfEdges(const std::vector<std::vector<size_t>> &intersections, size_t remainingStart,
                                size_t remainingCount, size_t retriStart, size_t retriEnd, GpuTables &out) const
{
    const size_t n = remainingCount + (retriEnd - retriStart);
    std::vector<uint32_t> tris(3 * n);
    size_t l = 0;
    for (const auto &r : {std::make_pair(remainingStart, remainingStart + remainingCount),
                          std::make_pair(retriStart, retriEnd)})
        for (size_t t = r.first; t < r.second; ++t, ++l)
            for (size_t k = 0; k < 3; ++k) {
                if (m_newTriangles[t][k] >= 0xffffffffu)
                    return false;
                tris[3 * l + k] = (uint32_t)m_newTriangles[t][k];
            }
    std::vector<uint64_t> queries;
    for (const auto &loop : intersections)
        for (size_t i = 0; i < loop.size(); ++i) {
            const size_t j = (i + 1) % loop.size();
            if (loop[i] >= 0xffffffffu || loop[j] >= 0xffffffffu)
                return false;
            queries.push_back(halfEdgeKey(loop[i], loop[j]));
            queries.push_back(halfEdgeKey(loop[j], loop[i]));
        }
    out.opposite.assign(3 * n, 0);
    out.loopOwner.assign(queries.size(), 0);
    sbhost::GpuLock lock(sbhost::gpuMutex());
    sbhost::check(sbg_half_edges(sbhost::gpu(), tris.data(), n, queries.data(), queries.size(), out.opposite.data(),
                                 out.loopOwner.data(), &out.firstRepeat),
                  "sbg_half_edges");
    // local -> global triangle indices
    auto global = [&](uint32_t x) -> uint32_t {
        if (x == ~0u)
            return x;
        return (uint32_t)(x < remainingCount ? remainingStart + x : retriStart + (x - remainingCount));
    };
    for (uint32_t &x : out.opposite)
        x = global(x);
    for (uint32_t &x : out.loopOwner)
        x = global(x);
    out.remainingStart = remainingStart;
    out.remainingCount = remainingCount;
    out.retriStart = retriStart;
    return true;
}

void SolidBoolean::buildFaceGroups(const std::vector<std::vector<size_t>> &intersections,
                                   const HalfEdgeMap &halfEdges, const GpuTables *tables, size_t remainingStart,
                                   size_t remainingCount, size_t retriStart, size_t retriEnd,
                                   std::vector<std::vector<size_t>> &groups)
{
    // (reference: src/solidboolean.cpp:167-239) seeds on both sides of every
    // loop edge, flood fill across half-edges that are not loop edges.  The
    // reference also records every half-edge a filled triangle owns, but a
    // triangle is filled once, so such a record is only ever looked up again
    // when another triangle has the same directed edge.  Unless that happened
    // (m_repeatedHalfEdge), the fill reads a per-half-edge array instead of
    // the maps: the triangle across the edge, or "loop edge", or nothing.
    // the owner of loop half-edge number q (forward, backward per loop edge,
    // in loop order), from the GPU tables or the host map
    constexpr uint32_t kNone = ~0u, kLoop = ~1u;
    auto loopOwner = [&](size_t q, uint64_t key) -> uint32_t {
        if (tables)
            return tables->loopOwner[q];
        const size_t *t = halfEdges.find(key);
        return t ? (uint32_t)*t : kNone;
    };
    HalfEdgeMap groupOf;
    size_t groupIndex = 0, q = 0;
    std::vector<std::pair<size_t, size_t>> queue;  // FIFO: [head, end)
    size_t head = 0;
    for (const auto &loop : intersections) {
        for (size_t i = 0; i < loop.size(); ++i) {
            const size_t j = (i + 1) % loop.size();
            const uint64_t fwd = halfEdgeKey(loop[i], loop[j]);
            groupOf.insert(fwd, groupIndex);
            const uint32_t tf = loopOwner(q++, fwd);
            if (tf != kNone)
                queue.push_back({tf, groupIndex});
            const uint64_t rev = halfEdgeKey(loop[j], loop[i]);
            groupOf.insert(rev, groupIndex + 1);
            const uint32_t tr = loopOwner(q++, rev);
            if (tr != kNone)
                queue.push_back({tr, groupIndex + 1});
        }
        groupIndex += 2;
    }
    groups.resize(groupIndex);
    const bool recordAll = !tables && (m_repeatedHalfEdge || m_newTriangles.size() >= kLoop);
    std::vector<uint8_t> visited(m_newTriangles.size(), 0);
    // Every triangle the fill can reach lies in this mesh's two ranges (its
    // untouched triangles, then its retriangulated ones).  Without repeated
    // half-edges the triangle across each half-edge is looked up once, in
    // triangle order with the slots prefetched ahead, so the misses in the
    // large table overlap instead of following the fill's order one by one.
    if (!recordAll) {
        m_opposite.resize(3 * m_newTriangles.size());
        constexpr size_t kAhead = 16;
        size_t l = 0;
        for (const auto &r : {std::make_pair(remainingStart, remainingStart + remainingCount),
                              std::make_pair(retriStart, retriEnd)})
            for (size_t t = r.first; t < r.second; ++t) {
                if (tables) {
                    for (size_t i = 0; i < 3; ++i)
                        m_opposite[3 * t + i] = tables->opposite[3 * l + i];
                    ++l;
                    continue;
                }
                if (t + kAhead < r.second) {
                    const auto &f = m_newTriangles[t + kAhead];
                    halfEdges.prefetch(halfEdgeKey(f[1], f[0]));
                    halfEdges.prefetch(halfEdgeKey(f[2], f[1]));
                    halfEdges.prefetch(halfEdgeKey(f[0], f[2]));
                }
                const auto &tri = m_newTriangles[t];
                for (size_t i = 0; i < 3; ++i) {
                    const size_t *o = halfEdges.find(halfEdgeKey(tri[(i + 1) % 3], tri[i]));
                    m_opposite[3 * t + i] = o ? (uint32_t)*o : kNone;
                }
            }
        // A loop half-edge, when some triangle has it, belongs to exactly that
        // triangle, so the fill's test "is this half-edge a loop edge" is a
        // flag on the owner's edge.  Cleared edges point nowhere.
        q = 0;
        for (const auto &loop : intersections)
            for (size_t i = 0; i < loop.size(); ++i) {
                const size_t j = (i + 1) % loop.size();
                for (const uint64_t he : {halfEdgeKey(loop[i], loop[j]), halfEdgeKey(loop[j], loop[i])}) {
                    const uint32_t o = loopOwner(q++, he);
                    if (o == kNone)
                        continue;
                    const auto &tri = m_newTriangles[o];
                    for (size_t k = 0; k < 3; ++k)
                        if (halfEdgeKey(tri[k], tri[(k + 1) % 3]) == he)
                            m_opposite[3 * (size_t)o + k] = kLoop;
                }
            }
    }
    auto drain = [&]() {
        while (head < queue.size()) {
            const auto tg = queue[head++];
            if (visited[tg.first])
                continue;
            visited[tg.first] = 1;
            groups[tg.second].push_back(tg.first);
            if (!recordAll) {
                for (size_t i = 0; i < 3; ++i) {
                    const uint32_t o = m_opposite[3 * tg.first + i];
                    if (o < kLoop)
                        queue.push_back({o, tg.second});
                }
                continue;
            }
            const auto &t = m_newTriangles[tg.first];
            for (size_t i = 0; i < 3; ++i) {
                const size_t j = (i + 1) % 3;
                const uint64_t he = halfEdgeKey(t[i], t[j]);
                if (groupOf.find(he))
                    continue;
                groupOf.insert(he, tg.second);
                if (const size_t *o = halfEdges.find(halfEdgeKey(t[j], t[i])))
                    queue.push_back({*o, tg.second});
            }
        }
        queue.clear();
        head = 0;
    };
    drain();
    const size_t end = remainingStart + remainingCount;
    for (size_t t = remainingStart; t < end; ++t) {
        if (visited[t])
            continue;
        groups.push_back({});
        queue.push_back({t, groupIndex});
        ++groupIndex;
        drain();
    }
}

void SolidBoolean::decideGroupSide(const std::vector<std::vector<size_t>> &groups, const SolidMesh *mesh,
                                   std::vector<bool> &groupSides)
{
    // (reference: src/solidboolean.cpp:482-510) one representative per group:
    // the centroid of its first triangle, voted by three rays on the GPU
    groupSides.assign(groups.size(), false);
    std::vector<double> pts;
    std::vector<size_t> which;
    for (size_t i = 0; i < groups.size(); ++i) {
        if (groups[i].empty())
            continue;
        const auto &t = m_newTriangles[groups[i][0]];
        const Vector3 c = (m_newVertices[t[0]] + m_newVertices[t[1]] + m_newVertices[t[2]]) / 3.0;
        pts.push_back(c.x());
        pts.push_back(c.y());
        pts.push_back(c.z());
        which.push_back(i);
    }
    if (which.empty())
        return;
    std::vector<uint8_t> inside(which.size());
    {
        sbhost::GpuLock lock(sbhost::gpuMutex());
        sbhost::check(sbg_classify_points(mesh->gpuMesh(), pts.data(), which.size(), nullptr, inside.data()),
                      "sbg_classify_points");
    }
    for (size_t k = 0; k < which.size(); ++k)
        groupSides[which[k]] = inside[k] != 0;
}

bool SolidBoolean::combine()
{
    if (!m_firstMesh->gpuMesh() || !m_secondMesh->gpuMesh())
        throw std::logic_error("SolidBoolean::combine: both meshes must be prepared");
    if (!m_robustWeld) {
        m_weld = Weld::Cells;
        return combineOnce();
    }
    // robust: crossings; if they cannot close the loops, PositionKey cells
    // with edge hints; then plain PositionKey cells, each from the start
    for (Weld w : {Weld::Crossings, Weld::HintedCells, Weld::Cells}) {
        m_weld = w;
        m_newVertices.clear();
        m_vertexSource.clear();
        m_newTriangles.clear();
        m_newPositionMap.clear();
        m_newKeyMap.clear();
        m_endpointKeys.clear();
        m_repeatedHalfEdge = false;
        m_firstTriangleGroups.clear();
        m_secondTriangleGroups.clear();
        m_firstGroupSides.clear();
        m_secondGroupSides.clear();
        if (combineOnce())
            return true;
        sbhost::diag("robust weld: attempt failed, trying the next weld");
    }
    return false;
}

bool SolidBoolean::combineOnce()
{
    const size_t na = m_firstMesh->triangles()->size();
    const size_t nb = m_secondMesh->triangles()->size();

    benchBegin_searchPotentialIntersectedPairs = Clock::now();
    std::unique_lock<std::recursive_mutex> gpuLock(sbhost::gpuMutex());  // until the hits are on the host
    sbg_pairs pairs = nullptr;
    sbhost::check(sbg_find_pairs(m_firstMesh->gpuMesh(), m_secondMesh->gpuMesh(), &pairs), "sbg_find_pairs");
    benchEnd_searchPotentialIntersectedPairs = Clock::now();

    benchBegin_processPotentialIntersectedPairs = Clock::now();
    std::vector<uint32_t> ab;
    std::vector<double> seg;
    {
        int rc = sbg_tri_tri(m_firstMesh->gpuMesh(), m_secondMesh->gpuMesh(), pairs, SBG_PRECISION_FP32_FILTERED,
                             nullptr);
        size_t nh = 0;
        if (rc == SBG_OK)
            rc = sbg_pairs_hits(pairs, &nh, nullptr, nullptr);
        if (rc == SBG_OK) {
            ab.resize(2 * nh);
            seg.resize(6 * nh);
            rc = sbg_pairs_hits(pairs, &nh, ab.data(), seg.data());
        }
        sbg_pairs_destroy(pairs);
        sbhost::check(rc, "sbg_tri_tri");
    }
    gpuLock.unlock();
    if (m_referencePairOrder && ab.size() > 2) {
        const sbhost::ReferenceTree &first = *m_firstMesh->referenceTree();
        const sbhost::ReferenceTree &second = *m_secondMesh->referenceTree();
        const std::vector<size_t> order = sbhost::referencePairOrder(first, second, ab.data(), ab.size() / 2);
        std::vector<uint32_t> ab2(ab.size());
        std::vector<double> seg2(seg.size());
        for (size_t i = 0; i < order.size(); ++i) {
            ab2[2 * i] = ab[2 * order[i]];
            ab2[2 * i + 1] = ab[2 * order[i] + 1];
            std::memcpy(&seg2[6 * i], &seg[6 * order[i]], 6 * sizeof(double));
        }
        ab.swap(ab2);
        seg.swap(seg2);
    }
    m_firstIntersected.assign(na, 0);
    m_secondIntersected.assign(nb, 0);
    if (m_weld != Weld::Cells) {
        // every endpoint's crossing: an edge of one mesh with a triangle of the
        // other; endpoints that match none are keyed by their PositionKey cell
        const size_t nh = ab.size() / 2;
        m_endpointKeys.assign(2 * nh, WeldKey{~0ull, 0});
        std::map<PositionKey, uint64_t> cells;
        const uint32_t offB = (uint32_t)m_firstMesh->vertices()->size();
        for (size_t h = 0; h < nh; ++h)
            for (size_t e = 0; e < 2; ++e) {
                uint32_t d[5];
                WeldKey &k = m_endpointKeys[2 * h + e];
                if (endpointSource(ab.data(), seg.data(), h, e, d)) {
                    const uint64_t x = d[0], y = d[1];
                    const bool edgeOfA = x < offB;
                    k.edge = std::min(x, y) << 32 | std::max(x, y);
                    k.other = edgeOfA ? na + ab[2 * h + 1] : ab[2 * h];
                } else {
                    const double *p = &seg[6 * h + 3 * e];
                    k.other = cells.insert({PositionKey(Vector3(p[0], p[1], p[2])), cells.size()}).first->second;
                }
            }
    }
    ContextMap firstCtx, secondCtx;
    auto addPoint = [this](IntersectedContext &c, const Vector3 &p, int64_t origin) {
        if (m_weld == Weld::Crossings) {
            auto r = c.keyMap.insert({m_endpointKeys[(size_t)origin], c.points.size()});
            if (r.second) {
                c.points.push_back(p);
                c.origin.push_back(origin);
            }
            return r.first->second;
        }
        auto r = c.positionMap.insert({PositionKey(p), c.points.size()});
        if (r.second) {
            c.points.push_back(p);
            c.origin.push_back(origin);
            if (m_weld == Weld::HintedCells)
                c.merged.emplace_back();
        }
        if (m_weld == Weld::HintedCells)
            c.merged[r.first->second].push_back(origin);
        return r.first->second;
    };
    auto addSegment = [&](IntersectedContext &c, const Vector3 &s, const Vector3 &t, size_t hit) {
        const size_t i = 3 + addPoint(c, s, 2 * (int64_t)hit);
        const size_t j = 3 + addPoint(c, t, 2 * (int64_t)hit + 1);
        if (i != j) {
            c.neighborMap[i].insert(j);
            c.neighborMap[j].insert(i);
        }
    };
    const size_t nh = ab.size() / 2;
    for (size_t h = 0; h < nh; ++h) {
        const size_t a = ab[2 * h], b = ab[2 * h + 1];
        const Vector3 s(seg[6 * h], seg[6 * h + 1], seg[6 * h + 2]);
        const Vector3 t(seg[6 * h + 3], seg[6 * h + 4], seg[6 * h + 5]);
        m_firstIntersected[a] = 1;
        m_secondIntersected[b] = 1;
        addSegment(firstCtx[a], s, t, h);
        addSegment(secondCtx[b], s, t, h);
    }
    m_hitPairs = ab;
    m_hitSegments = seg;
    benchEnd_processPotentialIntersectedPairs = Clock::now();

    // the half-edge tables are built on the GPU unless a directed edge
    // repeats, in which case the host maps, with the reference's insertion
    // order semantics, are built instead and the host stages rerun
    for (int attempt = 0; attempt < 2; ++attempt) {
        const int r = finishCombine(firstCtx, secondCtx, attempt == 0);
        if (r >= 0)
            return r == 1;
        m_newVertices.clear();
        m_vertexSource.clear();
        m_newTriangles.clear();
        m_newPositionMap.clear();
        m_newKeyMap.clear();
        m_repeatedHalfEdge = false;
        m_firstTriangleGroups.clear();
        m_secondTriangleGroups.clear();
    }
    return false;
}

// The host stages after the pairs: 1 / 0 as combine() returns, or -1 when the
// GPU half-edge tables found a repeated directed edge (rerun with host maps).
int SolidBoolean::finishCombine(const ContextMap &firstCtx, const ContextMap &secondCtx, bool gpuTables)
{
    EdgeMap firstEdges, secondEdges;
    std::vector<std::vector<size_t>> firstIntersections;
    HalfEdgeMap firstHalfEdges, secondHalfEdges;
    HalfEdgeMap *firstMap = gpuTables ? nullptr : &firstHalfEdges;
    HalfEdgeMap *secondMap = gpuTables ? nullptr : &secondHalfEdges;

    benchBegin_addUnintersectedTriangles = Clock::now();
    m_firstRemainingStart = m_newTriangles.size();
    if (!addUnintersectedTriangles(m_firstMesh, m_firstIntersected, firstMap, 0))
        sbhost::diag("Add first mesh remaining triangles failed");
    const size_t firstRemainingCount = m_newTriangles.size() - m_firstRemainingStart;
    m_firstRemainingCount = firstRemainingCount;
    m_secondRemainingStart = m_newTriangles.size();
    if (!addUnintersectedTriangles(m_secondMesh, m_secondIntersected, secondMap,
                                   m_firstMesh->vertices()->size()))
        sbhost::diag("Add second mesh remaining triangles failed");
    const size_t secondRemainingCount = m_newTriangles.size() - m_secondRemainingStart;
    m_secondRemainingCount = secondRemainingCount;
    benchEnd_addUnintersectedTriangles = Clock::now();

    benchBegin_reTriangulate = Clock::now();
    const size_t firstRetriStart = m_newTriangles.size();
    if (!reTriangulate(firstCtx, m_firstMesh, 0, firstEdges, firstMap)) {
        sbhost::diag("Retriangulate first mesh failed");
        return 0;
    }
    const size_t secondRetriStart = m_newTriangles.size();
    if (!reTriangulate(secondCtx, m_secondMesh, m_firstMesh->vertices()->size(), secondEdges, secondMap)) {
        sbhost::diag("Retriangulate second mesh failed");
        return 0;
    }
    const size_t secondRetriEnd = m_newTriangles.size();
    benchEnd_reTriangulate = Clock::now();

    benchBegin_buildPolygonsFromEdges = Clock::now();
    if (!buildPolygonsFromEdges(firstEdges, firstIntersections)) {
        sbhost::diag("Build polygons from edges failed");
        return 0;
    }
    benchEnd_buildPolygonsFromEdges = Clock::now();

    // both meshes are grouped by the first mesh's loops, as the reference does
    // (src/solidboolean.cpp:451-462)
    benchBegin_buildFaceGroups = Clock::now();
    GpuTables firstTables, secondTables;
    if (gpuTables) {
        if (!gpuHalfEdges(firstIntersections, m_firstRemainingStart, firstRemainingCount, firstRetriStart,
                          secondRetriStart, firstTables) ||
            !gpuHalfEdges(firstIntersections, m_secondRemainingStart, secondRemainingCount, secondRetriStart,
                          secondRetriEnd, secondTables) ||
            firstTables.firstRepeat != ~0u || secondTables.firstRepeat != ~0u)
            return -1;
    }
    buildFaceGroups(firstIntersections, firstHalfEdges, gpuTables ? &firstTables : nullptr, m_firstRemainingStart,
                    firstRemainingCount, firstRetriStart, secondRetriStart, m_firstTriangleGroups);
    buildFaceGroups(firstIntersections, secondHalfEdges, gpuTables ? &secondTables : nullptr,
                    m_secondRemainingStart, secondRemainingCount, secondRetriStart, secondRetriEnd,
                    m_secondTriangleGroups);
    benchEnd_buildFaceGroups = Clock::now();

    benchBegin_decideGroupSide = Clock::now();
    decideGroupSide(m_firstTriangleGroups, m_secondMesh, m_firstGroupSides);
    decideGroupSide(m_secondTriangleGroups, m_firstMesh, m_secondGroupSides);
    benchEnd_decideGroupSide = Clock::now();
    return 1;
}

void SolidBoolean::fetchUnion(std::vector<std::vector<size_t>> &out)
{
    for (size_t i = 0; i < m_firstGroupSides.size(); ++i)
        if (!m_firstGroupSides[i])
            for (size_t t : m_firstTriangleGroups[i])
                out.push_back({m_newTriangles[t][0], m_newTriangles[t][1], m_newTriangles[t][2]});
    for (size_t i = 0; i < m_secondGroupSides.size(); ++i)
        if (!m_secondGroupSides[i])
            for (size_t t : m_secondTriangleGroups[i])
                out.push_back({m_newTriangles[t][0], m_newTriangles[t][1], m_newTriangles[t][2]});
}

void SolidBoolean::fetchDiff(std::vector<std::vector<size_t>> &out)
{
    for (size_t i = 0; i < m_firstGroupSides.size(); ++i)
        if (!m_firstGroupSides[i])
            for (size_t t : m_firstTriangleGroups[i])
                out.push_back({m_newTriangles[t][0], m_newTriangles[t][1], m_newTriangles[t][2]});
    for (size_t i = 0; i < m_secondGroupSides.size(); ++i)
        if (m_secondGroupSides[i])
            for (size_t t : m_secondTriangleGroups[i]) {
                const auto &tri = m_newTriangles[t];
                out.push_back({tri[2], tri[1], tri[0]});  // B's inside, reversed (:541-546)
            }
}

void SolidBoolean::fetchIntersect(std::vector<std::vector<size_t>> &out)
{
    for (size_t i = 0; i < m_firstGroupSides.size(); ++i)
        if (m_firstGroupSides[i])
            for (size_t t : m_firstTriangleGroups[i])
                out.push_back({m_newTriangles[t][0], m_newTriangles[t][1], m_newTriangles[t][2]});
    for (size_t i = 0; i < m_secondGroupSides.size(); ++i)
        if (m_secondGroupSides[i])
            for (size_t t : m_secondTriangleGroups[i])
                out.push_back({m_newTriangles[t][0], m_newTriangles[t][1], m_newTriangles[t][2]});
}

const std::vector<Vector3> &SolidBoolean::resultVertices() { return m_newVertices; }

namespace {

struct D3 {
    double x, y, z;
};
D3 dsub(const D3 &a, const D3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 dcross(const D3 &a, const D3 &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double ddot(const D3 &a, const D3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// One segment endpoint as CONSTRUCT_INTERSECTION makes it
// (thirdparty/GuigueDevillers03/tri_tri_intersect.c:297-301 and siblings):
// v1 = base - onplane, v2 = base - other, alpha = DOT(v1,N)/DOT(v2,N),
// out = base - alpha * v2.
D3 edge_plane(const D3 &base, const D3 &onplane, const D3 &other, const D3 &N)
{
    const D3 v1 = dsub(base, onplane), v2 = dsub(base, other);
    const double alpha = ddot(v1, N) / ddot(v2, N);
    return dsub(base, {alpha * v2.x, alpha * v2.y, alpha * v2.z});
}

bool same_bits(const D3 &a, const double *b)
{
    return std::memcmp(&a.x, b, 8) == 0 && std::memcmp(&a.y, b + 1, 8) == 0 && std::memcmp(&a.z, b + 2, 8) == 0;
}

}  // namespace

bool SolidBoolean::endpointSource(const uint32_t *ab, const double *seg, size_t h, size_t e, uint32_t d[5]) const
{
    const auto &ta = *m_firstMesh->triangles();
    const auto &tb = *m_secondMesh->triangles();
    const uint32_t offB = (uint32_t)m_firstMesh->vertices()->size();
    const size_t a = ab[2 * h], b = ab[2 * h + 1];
    const double *pt = &seg[6 * h + 3 * e];
    // the two triangles, ids into the input vertices, as the GD test saw them
    uint32_t id[2][3];
    D3 v[2][3];
    for (int k = 0; k < 3; ++k) {
        id[0][k] = (uint32_t)ta[a][k];
        id[1][k] = offB + (uint32_t)tb[b][k];
        const Vector3 &pa = (*m_firstMesh->vertices())[ta[a][k]];
        const Vector3 &pb = (*m_secondMesh->vertices())[tb[b][k]];
        v[0][k] = {pa.x(), pa.y(), pa.z()};
        v[1][k] = {pb.x(), pb.y(), pb.z()};
    }
    // N1 = (q1 - p1) x (r1 - p1), N2 = (p2 - r2) x (q2 - r2) (tri_tri_intersect.c:403-431)
    const D3 N[2] = {dcross(dsub(v[0][1], v[0][0]), dsub(v[0][2], v[0][0])),
                     dcross(dsub(v[1][0], v[1][2]), dsub(v[1][1], v[1][2]))};
    for (int owner = 0; owner < 2; ++owner)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                if (i == j)
                    continue;
                for (int k = 0; k < 3; ++k) {
                    if (!same_bits(edge_plane(v[owner][i], v[1 - owner][k], v[owner][j], N[1 - owner]), pt))
                        continue;
                    d[0] = id[owner][i];
                    d[1] = id[owner][j];
                    d[2] = id[1 - owner][0];
                    d[3] = id[1 - owner][1];
                    d[4] = id[1 - owner][2];
                    return true;
                }
            }
    return false;
}

bool SolidBoolean::vertexSources(std::vector<uint32_t> &src5) const
{
    const uint32_t none = 0xffffffffu;
    src5.assign(5 * m_vertexSource.size(), none);
    for (size_t r = 0; r < m_vertexSource.size(); ++r) {
        const int64_t o = m_vertexSource[r];
        if (o >= 0) {
            src5[5 * r] = (uint32_t)o;
            continue;
        }
        const size_t h = (size_t)((-o - 1) / 2), e = (size_t)((-o - 1) % 2);
        if (!endpointSource(m_hitPairs.data(), m_hitSegments.data(), h, e, &src5[5 * r]))
            return false;
    }
    return true;
}

std::vector<size_t> SolidBoolean::intersectedFaces(int which) const
{
    const auto &f = which == 0 ? m_firstIntersected : m_secondIntersected;
    std::vector<size_t> out;
    for (size_t i = 0; i < f.size(); ++i)
        if (f[i])
            out.push_back(i);
    return out;
}

std::vector<uint8_t> SolidBoolean::triangleSides(int which) const
{
    const auto &flags = which == 0 ? m_firstIntersected : m_secondIntersected;
    const auto &groups = which == 0 ? m_firstTriangleGroups : m_secondTriangleGroups;
    const auto &sides = which == 0 ? m_firstGroupSides : m_secondGroupSides;
    const size_t start = which == 0 ? m_firstRemainingStart : m_secondRemainingStart;
    const size_t count = which == 0 ? m_firstRemainingCount : m_secondRemainingCount;
    std::vector<size_t> kept;
    for (size_t i = 0; i < flags.size() && kept.size() < count; ++i)
        if (!flags[i])
            kept.push_back(i);
    std::vector<uint8_t> out(flags.size(), 2);
    for (size_t g = 0; g < groups.size() && g < sides.size(); ++g)
        for (size_t t : groups[g])
            if (t >= start && t < start + kept.size())
                out[kept[t - start]] = sides[g] ? 1 : 0;
    return out;
}
