// Re-triangulation of one cut triangle (reference: src/retriangulator.h,
// src/retriangulator.cpp:29-337).  The triangle and the intersection points
// inside it are projected to the triangle's plane; the cut graph is split into
// open polylines (ending on the triangle's edges) and closed inner loops; the
// polylines cut the triangle into polygons, inner loops nest as holes, and
// every polygon is ear-clipped.  The cut graph and the loop nesting are held
// in the reference's container types (unordered_map / unordered_set keyed by
// point or loop index), filled in the reference's order: the walks visit them
// in their iteration order, which decides where polylines start, which way
// they run and the order of holes, and with that the output triangles.
#pragma once
#include <cstddef>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "vector2.h"
#include "vector3.h"

class ReTriangulator {
public:
    // points: the triangle's three corners; normal: its unit normal
    ReTriangulator(const std::vector<Vector3> &points, const Vector3 &normal);
    // points: the cut points (index 3 + k in neighborMapFrom3)
    void setEdges(const std::vector<Vector3> &points,
                  const std::unordered_map<size_t, std::unordered_set<size_t>> *neighborMapFrom3);
    // hints[k] (cut point 3 + k): the corner edge it lies on (0: 0-1, 1: 1-2,
    // 2: 2-0) or -1 to test it; kept by pointer until reTriangulate()
    void setEdgeHints(const std::vector<int> *hints) { m_edgeHints = hints; }
    bool reTriangulate();
    const std::vector<std::vector<size_t>> &polygons() const { return m_polygons; }
    const std::vector<std::vector<size_t>> &triangles() const { return m_triangles; }

private:
    Vector3 m_projectAxis, m_projectOrigin, m_projectNormal;
    std::vector<Vector2> m_points;
    const std::unordered_map<size_t, std::unordered_set<size_t>> *m_neighborMap = nullptr;
    const std::vector<int> *m_edgeHints = nullptr;
    std::vector<std::vector<size_t>> m_polylines;
    std::vector<std::vector<size_t>> m_innerPolygons;
    std::vector<std::vector<size_t>> m_polygons;
    std::unordered_map<size_t, std::unordered_set<size_t>> m_innerParents;
    std::unordered_map<size_t, std::unordered_set<size_t>> m_innerChildren;
    std::unordered_map<size_t, std::vector<size_t>> m_polygonHoles;
    std::vector<std::vector<size_t>> m_triangles;

    void lookupPolylines();
    int attachPointToTriangleEdge(size_t index) const;
    bool buildPolygons();
    void buildPolygonHierarchy();
    void triangulate();
};
