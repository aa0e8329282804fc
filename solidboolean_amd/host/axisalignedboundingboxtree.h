// AxisAlignedBoudingBoxTree with the reference's public interface
// (src/axisalignedboundingboxtree.h:26-106): a top-down tree over boxes, split
// at the mean of the members' centers along the longest axis of the node box,
// leaves of at most 20 boxes, and the simultaneous depth-first search of two
// trees that reports every pair of closed-overlapping boxes.
//
// Two kinds of tree share the type:
// * one built from boxes by the public constructor, as the reference builds
//   it (isPointInMesh builds a one-box tree per ray, src/solidboolean.cpp:55-63);
// * a mesh's tree (SolidMesh::axisAlignedBoundingBoxTree()), backed by the
//   mesh's LBVH in device memory.  Its host nodes are built on the first
//   root() call, with the reference's split rules over the reference's
//   triangle boxes, so a caller that walks nodes sees the reference's tree.
//
// test(first, second, secondBoxes, pairs) on the roots of two mesh trees (the
// call SolidBoolean::searchPotentialIntersectedPairs makes,
// src/solidboolean.cpp:94-101) finds the pairs on the GPU and appends them in
// the order the reference's walk reports them; any other call walks the host
// nodes as the reference does.  test(other, pairs) is the GPU search alone,
// in (a, b) order, with no host nodes.
#pragma once
#include <cstddef>
#include <utility>
#include <vector>

#include "axisalignedboundingbox.h"
#include "vector3.h"

class SolidMesh;

class AxisAlignedBoudingBoxTree {
public:
    struct Node {
        AxisAlignedBoudingBox boundingBox;
        Vector3 center;
        std::vector<size_t> boxIndices;
        Node *left = nullptr;
        Node *right = nullptr;
        bool isLeaf() const { return nullptr == left && nullptr == right; }
    };

    // the tree over boxes[boxIndices] with outterBox as the root's box
    // (src/axisalignedboundingboxtree.cpp:27-45); boxes must outlive the tree
    AxisAlignedBoudingBoxTree(const std::vector<AxisAlignedBoudingBox> *boxes, const std::vector<size_t> &boxIndices,
                              const AxisAlignedBoudingBox &outterBox);
    ~AxisAlignedBoudingBoxTree();
    AxisAlignedBoudingBoxTree(const AxisAlignedBoudingBoxTree &) = delete;
    AxisAlignedBoudingBoxTree &operator=(const AxisAlignedBoudingBoxTree &) = delete;

    const Node *root() const;
    const std::vector<AxisAlignedBoudingBox> *boxes() const;
    void splitNode(Node *node);
    void deleteNode(Node *node);
    // every (a under first, b under second) whose boxes closed-overlap, a from
    // this tree's boxes and b from secondBoxes, in the walk's order
    // (src/axisalignedboundingboxtree.h:54-88)
    void testNodes(const Node *first, const Node *second, const std::vector<AxisAlignedBoudingBox> *secondBoxes,
                   std::vector<std::pair<size_t, size_t>> *pairs) const;
    void test(const Node *first, const Node *second, const std::vector<AxisAlignedBoudingBox> *secondBoxes,
              std::vector<std::pair<size_t, size_t>> *pairs) const;

    // Mesh trees only (no counterpart in the reference): every (a in this
    // mesh, b in other's mesh) whose triangle boxes closed-overlap, found on
    // the GPU and appended in (a, b) order.
    void test(const AxisAlignedBoudingBoxTree *other, std::vector<std::pair<size_t, size_t>> *pairs) const;
    // the mesh of a mesh tree, else null
    const SolidMesh *mesh() const { return m_mesh; }

private:
    friend class SolidMesh;
    explicit AxisAlignedBoudingBoxTree(const SolidMesh *mesh);
    void build(const std::vector<size_t> &boxIndices, const AxisAlignedBoudingBox &outterBox) const;
    // the mesh tree whose root is `node`, or null
    static const AxisAlignedBoudingBoxTree *meshTreeOfRoot(const Node *node);

    const SolidMesh *m_mesh = nullptr;
    mutable const std::vector<AxisAlignedBoudingBox> *m_boxes = nullptr;
    mutable Node *m_root = nullptr;
    static constexpr size_t kLeafMaxNodeSize = 20;  // src/axisalignedboundingboxtree.cpp:25
};
