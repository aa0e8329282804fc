// SolidBoolean: union / difference / intersection of two closed meshes, with
// the reference's interface (src/solidboolean.h:33-108).
//
// combine() runs the data-parallel stages on the GPU through libsbgpu
// (candidate pairs, Guigue-Devillers segments, ray-parity group sides) and the
// topological stages on the host with the reference's rules: PositionKey
// welding of intersection points (first inserted wins), re-triangulation of
// cut triangles, intersection loops, half-edge flood fill into groups.  The
// containers whose iteration order the results depend on (cut-triangle
// contexts, their cut graphs, the intersection edges) are the reference's own
// types, unordered_map / unordered_set keyed by triangle or point index, filled
// by the same operations in the same order; intersecting pairs are processed
// in (a, b) order, or in the reference's search order (setReferencePairOrder),
// with which the numbering of new points, the direction of loops and the
// first triangle of every group are the reference's.
#pragma once
#include <chrono>
#include <cstdint>
#include <array>
#include <map>
#include <set>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "positionkey.h"
#include "solidmesh.h"
#include "vector3.h"

class SolidBoolean {
public:
    SolidBoolean(const SolidMesh *firstMesh, const SolidMesh *secondMesh);
    ~SolidBoolean();
    bool combine();
    void fetchUnion(std::vector<std::vector<size_t>> &resultTriangles);
    void fetchDiff(std::vector<std::vector<size_t>> &resultTriangles);
    void fetchIntersect(std::vector<std::vector<size_t>> &resultTriangles);
    const std::vector<Vector3> &resultVertices();

    std::chrono::time_point<std::chrono::high_resolution_clock> benchBegin_searchPotentialIntersectedPairs;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchEnd_searchPotentialIntersectedPairs;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchBegin_processPotentialIntersectedPairs;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchEnd_processPotentialIntersectedPairs;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchBegin_addUnintersectedTriangles;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchEnd_addUnintersectedTriangles;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchBegin_reTriangulate;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchEnd_reTriangulate;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchBegin_buildPolygonsFromEdges;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchEnd_buildPolygonsFromEdges;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchBegin_buildFaceGroups;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchEnd_buildFaceGroups;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchBegin_decideGroupSide;
    std::chrono::time_point<std::chrono::high_resolution_clock> benchEnd_decideGroupSide;

    // Inspection (no counterpart in the reference's public API): the
    // intersected triangles of mesh 0 / 1, and for every original triangle of
    // mesh 0 / 1 the side of the group it landed in (0 outside, 1 inside) or
    // 2 when it was intersected.
    std::vector<size_t> intersectedFaces(int which) const;
    // Where each result vertex comes from, 5 indices per vertex into the input
    // vertices (first mesh's, then second's): an input vertex v is
    // (v, ~0, ~0, ~0, ~0); a new vertex is (x, y, p0, p1, p2), the point where
    // the line through x and y (an edge of one triangle) meets the plane through
    // p0, p1, p2 (the other triangle).  False if some endpoint matches no edge.
    bool vertexSources(std::vector<uint32_t> &src5) const;
    std::vector<uint8_t> triangleSides(int which) const;
    // Before combine() (no counterpart in the reference; on by default): take
    // the intersecting pairs in the order the reference's own tree search
    // reports them (on) or in (a, b) order (off).  The weld keeps the first
    // point it meets in each PositionKey cell and the first triangle of a
    // group decides its side, so with this on the results are the
    // reference's bit for bit and in its order, also where points share a
    // cell or a group's triangles disagree.  It needs the reference's tree
    // of both meshes, which SolidMesh::prepare() builds on a host thread
    // beside the GPU work (about 25 ms per 82k-triangle mesh, 0.4 s per
    // 1.31M-triangle mesh on one core); off, combine() never waits for it.
    void setReferencePairOrder(bool on) { m_referencePairOrder = on; }
    // Before combine() (not the reference's behaviour, off by default): weld
    // intersection points by what they are, the crossing of one mesh's edge
    // with the other mesh's triangle, recovered bit for bit from the segment
    // endpoints, instead of by 1e-5 PositionKey cells; and tell the
    // re-triangulation which of its triangle's edges a cut point lies on
    // instead of testing it.  Distinct points closer than 1e-5 then stay
    // distinct, and points on an edge are never rejected, which is where the
    // reference's combine() fails on dense meshes ("Attach point to triangle
    // edge failed").  An endpoint that matches no edge-plane crossing (a
    // degenerate touch) falls back to its PositionKey cell.  When that fails,
    // combine() runs again welding by PositionKey cell as the reference does
    // but still hinting each cut point's edge (from every crossing welded into
    // it), and then with the plain PositionKey weld, so it succeeds wherever
    // the default does.
    void setRobustWeld(bool on) { m_robustWeld = on; }

    // open-addressed map from a half-edge key (first << 32 | second) to a triangle
    class HalfEdgeMap {
    public:
        void reserve(size_t n);
        bool insert(uint64_t key, size_t value);  // false when the key is present
        const size_t *find(uint64_t key) const;
        void prefetch(uint64_t key) const;  // the first slot `key` probes, into cache

    private:
        struct Slot {
            uint64_t key;
            size_t value;
        };
        std::vector<Slot> m_slots;  // key and value side by side: one cache line per probe
        size_t m_count = 0;
        size_t m_mask = 0;
        void grow();
    };

private:
    // robust weld: per hit endpoint (2 h + e) its key; (edge (lo << 32 | hi) of
    // input vertex ids, other triangle (first mesh's, then second's)), or
    // (~0, PositionKey cell number) when no crossing matched
    struct WeldKey {
        uint64_t edge, other;
        bool operator<(const WeldKey &o) const { return edge != o.edge ? edge < o.edge : other < o.other; }
    };
    struct IntersectedContext {
        std::vector<Vector3> points;
        std::vector<int64_t> origin;  // per point: 2 * hit + endpoint of its first insertion
        std::map<PositionKey, size_t> positionMap;
        std::map<WeldKey, size_t> keyMap;  // robust weld
        std::vector<std::vector<int64_t>> merged;  // hinted cells: every endpoint welded into each point
        std::unordered_map<size_t, std::unordered_set<size_t>> neighborMap;
    };
    using ContextMap = std::unordered_map<size_t, IntersectedContext>;
    using EdgeMap = std::unordered_map<size_t, std::unordered_set<size_t>>;

    const SolidMesh *m_firstMesh;
    const SolidMesh *m_secondMesh;
    std::vector<Vector3> m_newVertices;
    std::vector<std::array<size_t, 3>> m_newTriangles;
    bool m_repeatedHalfEdge = false;  // some directed edge belongs to two new triangles
    bool m_referencePairOrder = true;
    bool m_robustWeld = false;
    // the weld of the current attempt: PositionKey cells; crossings (keys);
    // or PositionKey cells with the edge hints of every endpoint in the cell
    enum class Weld { Cells, Crossings, HintedCells };
    Weld m_weld = Weld::Cells;
    std::vector<WeldKey> m_endpointKeys;
    std::map<WeldKey, size_t> m_newKeyMap;
    bool endpointSource(const uint32_t *ab, const double *seg, size_t h, size_t e, uint32_t d[5]) const;
    std::map<PositionKey, size_t> m_newPositionMap;
    std::vector<std::vector<size_t>> m_firstTriangleGroups, m_secondTriangleGroups;
    std::vector<bool> m_firstGroupSides, m_secondGroupSides;
    std::vector<uint8_t> m_firstIntersected, m_secondIntersected;
    size_t m_firstRemainingStart = 0, m_secondRemainingStart = 0;
    size_t m_firstRemainingCount = 0, m_secondRemainingCount = 0;
    // per result vertex: >= 0 an input vertex (first mesh's, then second's),
    // < 0 a hit endpoint -(2 * hit + endpoint) - 1
    std::vector<int64_t> m_vertexSource;
    std::vector<uint32_t> m_hitPairs;  // the intersecting pairs (a, b)
    std::vector<double> m_hitSegments;  // their segments (source, target)

    static uint64_t halfEdgeKey(size_t first, size_t second) { return ((uint64_t)first << 32) | (uint64_t)second; }
    size_t addNewPoint(const Vector3 &position, int64_t origin);
    bool addUnintersectedTriangles(const SolidMesh *mesh, const std::vector<uint8_t> &used, HalfEdgeMap *halfEdges,
                                   size_t inputOffset);
    bool reTriangulate(const ContextMap &contexts, const SolidMesh *mesh, size_t startOldVertex, EdgeMap &edges,
                       HalfEdgeMap *halfEdges);
    struct GpuTables {
        std::vector<uint32_t> opposite;   // per half-edge of the mesh's triangles, local order
        std::vector<uint32_t> loopOwner;  // per loop half-edge
        uint32_t firstRepeat = ~0u;
        size_t remainingStart = 0, remainingCount = 0, retriStart = 0;
    };
    bool gpuHalfEdges(const std::vector<std::vector<size_t>> &intersections, size_t remainingStart,
                      size_t remainingCount, size_t retriStart, size_t retriEnd, GpuTables &out) const;
    int finishCombine(const ContextMap &firstCtx, const ContextMap &secondCtx, bool gpuTables);
    bool combineOnce();
    bool buildPolygonsFromEdges(const EdgeMap &edges, std::vector<std::vector<size_t>> &polygons);
    void buildFaceGroups(const std::vector<std::vector<size_t>> &intersections, const HalfEdgeMap &halfEdges,
                         const GpuTables *tables, size_t remainingStart, size_t remainingCount, size_t retriStart,
                         size_t retriEnd, std::vector<std::vector<size_t>> &groups);
    std::vector<uint32_t> m_opposite;  // buildFaceGroups: per half-edge (3 t + i), the triangle across it
    void decideGroupSide(const std::vector<std::vector<size_t>> &groups, const SolidMesh *mesh,
                         std::vector<bool> &groupSides);
};
