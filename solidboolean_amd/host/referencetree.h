// The reference's own box tree and the order its search reports pairs in.
//
// The GPU search finds the same candidate set as the reference
// (src/solidboolean.cpp:94-101) but reports it in (a, b) order.  The
// reference reports it in the order of a simultaneous depth-first walk of two
// trees built top-down by mean splits (src/axisalignedboundingboxtree.cpp:
// 27-141, src/axisalignedboundingboxtree.h:54-95), and the weld keeps the
// first point it meets in each PositionKey cell, so where two intersection
// points share a cell the kept one depends on that order.  ReferenceTree
// rebuilds such a tree on the host and referencePairOrder sorts any list of
// (a, b) pairs into the walk's order.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "axisalignedboundingboxtree.h"
#include "vector3.h"

namespace sbhost {

class ReferenceTree {
public:
    ReferenceTree(const std::vector<Vector3> &vertices, const std::vector<std::vector<size_t>> &triangles);
    // the same tables read off a tree already built with those rules (root
    // and the number of boxes it indexes)
    ReferenceTree(const AxisAlignedBoudingBoxTree::Node *root, size_t nboxes);

    // the walk's turns (0 left, 1 right) from the root down to triangle t's leaf
    const std::string &pathTo(size_t t) const { return m_paths[m_leafOf[t]]; }
    // t's place among its leaf's triangles, in the leaf's order
    uint32_t placeInLeaf(size_t t) const { return m_place[t]; }

    struct Node {
        uint32_t left = 0, right = 0;  // children, both 0 for a leaf
        uint32_t count = 0;            // triangles below
        bool leaf() const { return left == 0 && right == 0; }
    };
    const std::vector<Node> &nodes() const { return m_nodes; }

private:
    std::vector<Node> m_nodes;          // [0] is the root
    std::vector<uint32_t> m_leafOf;     // per triangle: its leaf's node
    std::vector<uint32_t> m_place;      // per triangle
    std::vector<std::string> m_paths;   // per node (filled for leaves)
};

// The permutation that puts pairs ab[2 i], ab[2 i + 1] (a in the first tree's
// mesh, b in the second's) into the order in which the reference's search
// over (first, second) reports them.
std::vector<size_t> referencePairOrder(const ReferenceTree &first, const ReferenceTree &second, const uint32_t *ab,
                                       size_t n);

}  // namespace sbhost
