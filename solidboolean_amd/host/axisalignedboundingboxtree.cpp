// See axisalignedboundingboxtree.h.  The split rules follow
// src/axisalignedboundingboxtree.cpp:27-141 and the walk
// src/axisalignedboundingboxtree.h:54-95; both are written here without
// recursion (an explicit stack), which changes neither the tree nor the order
// of the reported pairs.
#include "axisalignedboundingboxtree.h"

#include <algorithm>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <unordered_map>

#include "referencetree.h"
#include "sbhost_log.h"
#include "solidmesh.h"

namespace {

std::mutex g_rootsMu;
std::unordered_map<const AxisAlignedBoudingBoxTree::Node *, const AxisAlignedBoudingBoxTree *> g_meshRoots;

}  // namespace

AxisAlignedBoudingBoxTree::AxisAlignedBoudingBoxTree(const std::vector<AxisAlignedBoudingBox> *boxes,
                                                     const std::vector<size_t> &boxIndices,
                                                     const AxisAlignedBoudingBox &outterBox)
{
    m_boxes = boxes;
    build(boxIndices, outterBox);
}

AxisAlignedBoudingBoxTree::AxisAlignedBoudingBoxTree(const SolidMesh *mesh) : m_mesh(mesh) {}

AxisAlignedBoudingBoxTree::~AxisAlignedBoudingBoxTree()
{
    if (m_mesh && m_root) {
        std::lock_guard<std::mutex> lock(g_rootsMu);
        g_meshRoots.erase(m_root);
    }
    deleteNode(m_root);
}

void AxisAlignedBoudingBoxTree::build(const std::vector<size_t> &boxIndices, const AxisAlignedBoudingBox &outterBox) const
{
    Node *root = new Node;
    root->boundingBox = outterBox;
    root->boxIndices = boxIndices;
    if (!boxIndices.empty()) {
        for (size_t i : boxIndices)
            root->center += (*m_boxes)[i].center();
        root->center /= (float)boxIndices.size();
    }
    m_root = root;
    const_cast<AxisAlignedBoudingBoxTree *>(this)->splitNode(root);
}

const AxisAlignedBoudingBoxTree::Node *AxisAlignedBoudingBoxTree::root() const
{
    if (!m_root && m_mesh) {
        // the mesh's triangle boxes and the box of all its triangles' corners
        // (src/solidmesh.cpp:57-75)
        m_boxes = m_mesh->triangleAxisAlignedBoundingBoxes();
        const auto &tris = *m_mesh->triangles();
        const auto &verts = *m_mesh->vertices();
        AxisAlignedBoudingBox all;
        std::vector<size_t> indices(tris.size());
        for (size_t t = 0; t < tris.size(); ++t) {
            indices[t] = t;
            for (size_t k = 0; k < 3; ++k)
                all.update(verts[tris[t][k]]);
        }
        all.updateCenter();
        build(indices, all);
        std::lock_guard<std::mutex> lock(g_rootsMu);
        g_meshRoots[m_root] = this;
    }
    return m_root;
}

const std::vector<AxisAlignedBoudingBox> *AxisAlignedBoudingBoxTree::boxes() const
{
    if (!m_boxes && m_mesh)
        m_boxes = m_mesh->triangleAxisAlignedBoundingBoxes();
    return m_boxes;
}

void AxisAlignedBoudingBoxTree::deleteNode(Node *node)
{
    std::vector<Node *> todo;
    if (node)
        todo.push_back(node);
    while (!todo.empty()) {
        Node *x = todo.back();
        todo.pop_back();
        if (x->left)
            todo.push_back(x->left);
        if (x->right)
            todo.push_back(x->right);
        delete x;
    }
    if (node == m_root)
        m_root = nullptr;
}

void AxisAlignedBoudingBoxTree::splitNode(Node *node)
{
    std::vector<Node *> todo{node};
    std::vector<size_t> below, rest;
    while (!todo.empty()) {
        Node *x = todo.back();
        todo.pop_back();
        const std::vector<size_t> &members = x->boxIndices;
        const size_t n = members.size();
        if (n <= kLeafMaxNodeSize)
            continue;
        // the longest side of the node's box, spans narrowed to float, the
        // first of equal ones; the split is the node's center on that axis
        const Vector3 &lo = x->boundingBox.lowerBound(), &hi = x->boundingBox.upperBound();
        size_t axis = 0;
        float longest = (float)(hi[0] - lo[0]);
        for (size_t k = 1; k < 3; ++k) {
            const float span = (float)(hi[k] - lo[k]);
            if (longest < span) {
                longest = span;
                axis = k;
            }
        }
        const double split = x->center[axis];
        below.clear();
        rest.clear();
        for (size_t i : members)
            ((*m_boxes)[i].center()[axis] < split ? below : rest).push_back(i);
        // the left child takes the members below the split in reverse scan
        // order, the right one the others in scan order; when one side is
        // empty the other is halved, the left child taking the rounded-down
        // half of a right side or the rounded-up half of a left side
        std::reverse(below.begin(), below.end());
        size_t cut = below.size();
        if (below.empty())
            cut = rest.size() / 2;
        else if (rest.empty())
            cut = below.size() - below.size() / 2;
        below.insert(below.end(), rest.begin(), rest.end());
        x->left = new Node;
        x->right = new Node;
        for (size_t j = 0; j < n; ++j) {
            Node *c = j < cut ? x->left : x->right;
            const AxisAlignedBoudingBox &b = (*m_boxes)[below[j]];
            c->boundingBox.update(b.lowerBound());
            c->boundingBox.update(b.upperBound());
            c->boxIndices.push_back(below[j]);
            c->center += b.center();
        }
        x->left->center /= (float)x->left->boxIndices.size();
        x->right->center /= (float)x->right->boxIndices.size();
        todo.push_back(x->right);
        todo.push_back(x->left);
    }
}

void AxisAlignedBoudingBoxTree::testNodes(const Node *first, const Node *second,
                                          const std::vector<AxisAlignedBoudingBox> *secondBoxes,
                                          std::vector<std::pair<size_t, size_t>> *pairs) const
{
    // node pairs in the order the recursive walk visits them: the pair pushed
    // last is taken first
    std::vector<std::pair<const Node *, const Node *>> todo{{first, second}};
    while (!todo.empty()) {
        const Node *a = todo.back().first, *b = todo.back().second;
        todo.pop_back();
        if (!a->boundingBox.intersectWith(b->boundingBox))
            continue;
        if (a->isLeaf() && b->isLeaf()) {
            for (size_t i : a->boxIndices)
                for (size_t j : b->boxIndices)
                    if ((*m_boxes)[i].intersectWith((*secondBoxes)[j]))
                        pairs->push_back({i, j});
            continue;
        }
        // descend the inner node, or of two inner nodes the one with more
        // boxes (the first on a tie)
        const bool descendSecond = a->isLeaf() || (!b->isLeaf() && a->boxIndices.size() < b->boxIndices.size());
        if (descendSecond) {
            todo.push_back({a, b->right});
            todo.push_back({a, b->left});
        } else {
            todo.push_back({a->right, b});
            todo.push_back({a->left, b});
        }
    }
}

const AxisAlignedBoudingBoxTree *AxisAlignedBoudingBoxTree::meshTreeOfRoot(const Node *node)
{
    std::lock_guard<std::mutex> lock(g_rootsMu);
    auto it = g_meshRoots.find(node);
    return it == g_meshRoots.end() ? nullptr : it->second;
}

void AxisAlignedBoudingBoxTree::test(const Node *first, const Node *second,
                                     const std::vector<AxisAlignedBoudingBox> *secondBoxes,
                                     std::vector<std::pair<size_t, size_t>> *pairs) const
{
    const AxisAlignedBoudingBoxTree *other = meshTreeOfRoot(second);
    if (m_mesh && first && first == m_root && other && secondBoxes == other->boxes()) {
        // two whole mesh trees: the GPU search, put in the walk's order
        std::vector<std::pair<size_t, size_t>> found;
        test(other, &found);
        std::vector<uint32_t> ab(2 * found.size());
        for (size_t i = 0; i < found.size(); ++i) {
            ab[2 * i] = (uint32_t)found[i].first;
            ab[2 * i + 1] = (uint32_t)found[i].second;
        }
        const sbhost::ReferenceTree ta(m_root, m_boxes->size()), tb(other->m_root, other->m_boxes->size());
        const std::vector<size_t> order = sbhost::referencePairOrder(ta, tb, ab.data(), found.size());
        pairs->reserve(pairs->size() + found.size());
        for (size_t i : order)
            pairs->push_back(found[i]);
        return;
    }
    testNodes(first, second, secondBoxes, pairs);
}

void AxisAlignedBoudingBoxTree::test(const AxisAlignedBoudingBoxTree *other,
                                     std::vector<std::pair<size_t, size_t>> *pairs) const
{
    if (!m_mesh || !other || !other->m_mesh)
        throw std::logic_error("AxisAlignedBoudingBoxTree::test: both trees must belong to prepared meshes");
    size_t n = 0;
    std::vector<uint32_t> ab;
    {
        sbhost::GpuLock lock(sbhost::gpuMutex());
        sbg_pairs p = nullptr;
        sbhost::check(sbg_find_pairs(m_mesh->gpuMesh(), other->m_mesh->gpuMesh(), &p), "sbg_find_pairs");
        sbg_pairs_count(p, &n);
        ab.resize(2 * n);
        int rc = sbg_pairs_copy(p, ab.data());
        sbg_pairs_destroy(p);
        sbhost::check(rc, "sbg_pairs_copy");
    }
    pairs->reserve(pairs->size() + n);
    for (size_t i = 0; i < n; ++i)
        pairs->push_back({ab[2 * i], ab[2 * i + 1]});
}
