// Re-triangulation of one cut triangle; see retriangulator.h.
#include "retriangulator.h"

#include <algorithm>
#include <array>
#include <deque>

#include "earclip.h"
#include "sbhost_log.h"

ReTriangulator::ReTriangulator(const std::vector<Vector3> &points, const Vector3 &normal) : m_projectNormal(normal)
{
    // plane frame: x along the first edge, origin at the first corner
    // (reference: src/retriangulator.cpp:29-37)
    m_projectAxis = (points[1] - points[0]).normalized();
    m_projectOrigin = points[0];
    Vector3::project(points, &m_points, m_projectNormal, m_projectAxis, m_projectOrigin);
}

void ReTriangulator::setEdges(const std::vector<Vector3> &points,
                              const std::unordered_map<size_t, std::unordered_set<size_t>> *neighborMapFrom3)
{
    Vector3::project(points, &m_points, m_projectNormal, m_projectAxis, m_projectOrigin);
    m_neighborMap = neighborMapFrom3;
}

// Split the cut graph into closed inner loops and open polylines: walks start
// from degree-1 points first, then from the rest, always stepping to the first
// unvisited neighbour (reference: src/retriangulator.cpp:47-92).
void ReTriangulator::lookupPolylines()
{
    const auto &nm = *m_neighborMap;
    std::vector<size_t> starts;
    starts.reserve(nm.size());
    for (const auto &kv : nm)
        if (kv.second.size() == 1)
            starts.push_back(kv.first);
    for (const auto &kv : nm)
        if (kv.second.size() > 1)
            starts.push_back(kv.first);
    std::unordered_set<size_t> visited;
    for (size_t start : starts) {
        if (visited.count(start))
            continue;
        std::vector<size_t> line;
        size_t cur = start;
        for (;;) {
            visited.insert(cur);
            line.push_back(cur);
            auto it = nm.find(cur);
            if (it == nm.end())
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
        if (line.size() >= 3) {
            auto last = nm.find(line.back());
            if (last->second.count(start)) {
                m_innerPolygons.push_back(line);
                continue;
            }
        }
        if (line.size() >= 2)
            m_polylines.push_back(line);
    }
}

int ReTriangulator::attachPointToTriangleEdge(size_t index) const
{
    if (m_edgeHints && index >= 3 && index - 3 < m_edgeHints->size() && (*m_edgeHints)[index - 3] >= 0)
        return (*m_edgeHints)[index - 3];
    const Vector2 &p = m_points[index];
    for (int i = 0; i < 3; ++i)
        if (p.isOnLine(m_points[i], m_points[(i + 1) % 3]))
            return i;
    return -1;
}

// Walk the triangle's boundary as a ring of corners and polyline ends (ordered
// by distance from each edge's first corner); every polyline end jumps across
// to its other end, which closes one polygon per start (reference:
// src/retriangulator.cpp:125-235).
bool ReTriangulator::buildPolygons()
{
    struct EdgePoint {
        size_t pointIndex = 0;
        size_t linkToPointIndex = 0;
        int polylineIndex = -1;
        bool reversed = false;
        double squaredDistance = 0.0;
        int linkTo = -1;
    };
    std::array<std::vector<EdgePoint>, 3> edgePoints;
    for (int pi = 0; pi < (int)m_polylines.size(); ++pi) {
        const auto &line = m_polylines[pi];
        const int fe = attachPointToTriangleEdge(line.front());
        const int be = attachPointToTriangleEdge(line.back());
        if (fe == -1 || be == -1) {
            sbhost::diag("Attach point to triangle edge failed");
            return false;
        }
        EdgePoint f;
        f.pointIndex = line.front();
        f.linkToPointIndex = line.back();
        f.polylineIndex = pi;
        f.reversed = false;
        f.squaredDistance = (m_points[line.front()] - m_points[fe]).lengthSquared();
        edgePoints[fe].push_back(f);
        EdgePoint b;
        b.pointIndex = line.back();
        b.linkToPointIndex = line.front();
        b.polylineIndex = pi;
        b.reversed = true;
        b.squaredDistance = (m_points[line.back()] - m_points[be]).lengthSquared();
        edgePoints[be].push_back(b);
    }
    for (auto &v : edgePoints)
        std::sort(v.begin(), v.end(),
                  [](const EdgePoint &x, const EdgePoint &y) { return x.squaredDistance < y.squaredDistance; });
    std::vector<EdgePoint> ring;
    for (size_t i = 0; i < 3; ++i) {
        EdgePoint c;
        c.pointIndex = i;
        ring.push_back(c);
        for (const auto &e : edgePoints[i])
            ring.push_back(e);
    }
    std::unordered_map<size_t, size_t> ringPos;
    for (size_t i = 0; i < ring.size(); ++i)
        if (ring[i].polylineIndex != -1)
            ringPos.insert({ring[i].pointIndex, i});
    size_t budget = 4 * ring.size();
    for (auto &e : ring) {
        if (e.polylineIndex == -1)
            continue;
        auto f = ringPos.find(e.linkToPointIndex);
        if (f != ringPos.end())
            e.linkTo = (int)f->second;
        budget += m_polylines[e.polylineIndex].size();
    }
    std::vector<bool> visited(ring.size(), false);
    std::deque<size_t> starts{0};
    while (!starts.empty()) {
        const size_t startIndex = starts.front();
        starts.pop_front();
        if (visited[startIndex])
            continue;
        std::vector<size_t> polygon;
        size_t at = startIndex;
        size_t steps = 0;
        do {
            if (++steps > budget) {
                sbhost::diag("polygon walk did not close");
                return false;
            }
            const EdgePoint &e = ring[at];
            visited[at] = true;
            if (e.polylineIndex == -1) {
                polygon.push_back(e.pointIndex);
                at = (at + 1) % ring.size();
            } else if (e.linkTo != -1) {
                const auto &line = m_polylines[e.polylineIndex];
                if (e.reversed)
                    polygon.insert(polygon.end(), line.rbegin(), line.rend());
                else
                    polygon.insert(polygon.end(), line.begin(), line.end());
                starts.push_back((at + 1) % ring.size());
                at = ((size_t)e.linkTo + 1) % ring.size();
            } else {
                sbhost::diag("linkTo failed");
                return false;
            }
        } while (at != startIndex);
        m_polygons.push_back(polygon);
    }
    return true;
}

// Nest inner loops: loop i is a child of loop j when its first point lies
// inside j; top-level loops become holes of the polygons containing them
// (reference: src/retriangulator.cpp:103-123).
void ReTriangulator::buildPolygonHierarchy()
{
    const size_t n = m_innerPolygons.size();
    for (size_t i = 0; i < n; ++i)
        for (size_t j = i + 1; j < n; ++j) {
            if (m_points[m_innerPolygons[i][0]].isInPolygon(m_points, m_innerPolygons[j])) {
                m_innerChildren[j].insert(i);
                m_innerParents[i].insert(j);
            } else if (m_points[m_innerPolygons[j][0]].isInPolygon(m_points, m_innerPolygons[i])) {
                m_innerChildren[i].insert(j);
                m_innerParents[j].insert(i);
            }
        }
    for (size_t i = 0; i < n; ++i) {
        if (m_innerParents.count(i))
            continue;
        for (size_t j = 0; j < m_polygons.size(); ++j)
            if (m_points[m_innerPolygons[i][0]].isInPolygon(m_points, m_polygons[j]))
                m_polygonHoles[j].push_back(i);
    }
}

void ReTriangulator::triangulate()
{
    auto clip = [&](const std::vector<size_t> &border, const std::vector<size_t> &holeLoops) {
        std::vector<sbhost::Ring> rings;
        std::vector<size_t> index;
        sbhost::Ring r;
        for (size_t p : border) {
            index.push_back(p);
            r.push_back({m_points[p].x(), m_points[p].y()});
        }
        rings.push_back(r);
        for (size_t h : holeLoops) {
            sbhost::Ring hr;
            for (size_t p : m_innerPolygons[h]) {
                index.push_back(p);
                hr.push_back({m_points[p].x(), m_points[p].y()});
            }
            rings.push_back(hr);
        }
        const std::vector<size_t> tri = sbhost::earclip(rings);
        for (size_t i = 0; i + 2 < tri.size(); i += 3)
            m_triangles.push_back({index[tri[i]], index[tri[i + 1]], index[tri[i + 2]]});
    };
    for (size_t pi = 0; pi < m_polygons.size(); ++pi) {
        auto h = m_polygonHoles.find(pi);
        clip(m_polygons[pi], h == m_polygonHoles.end() ? std::vector<size_t>() : h->second);
    }
    // inner loops are filled too, with their direct children (not grandchildren) as holes
    for (size_t pi = 0; pi < m_innerPolygons.size(); ++pi) {
        std::vector<size_t> holes;
        auto c = m_innerChildren.find(pi);
        if (c != m_innerChildren.end()) {
            std::unordered_set<size_t> children = c->second;
            for (size_t child : c->second) {
                auto g = m_innerChildren.find(child);
                if (g != m_innerChildren.end())
                    for (size_t grand : g->second)
                        children.erase(grand);
            }
            holes.assign(children.begin(), children.end());
        }
        clip(m_innerPolygons[pi], holes);
    }
}

bool ReTriangulator::reTriangulate()
{
    lookupPolylines();
    if (!buildPolygons()) {
        sbhost::diag("Build polygons failed");
        return false;
    }
    buildPolygonHierarchy();
    triangulate();
    return true;
}
