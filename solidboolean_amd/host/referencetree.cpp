// See referencetree.h.  The split rules follow
// src/axisalignedboundingboxtree.cpp:27-141 over the boxes and centers of
// src/solidmesh.cpp:42-76 and src/axisalignedboundingbox.h:31-53; the walk
// follows src/axisalignedboundingboxtree.h:54-95.  Every value the shape
// depends on is computed with the same FP64 operations in the same order
// (sums of centers in member order, a division by the member count narrowed to
// float, spans narrowed to float).
#include "referencetree.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <deque>
#include <thread>
#include <utility>

namespace sbhost {

namespace {

constexpr size_t kLeafMax = 20;  // src/axisalignedboundingboxtree.cpp:25

struct Box {
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    void add(const double *p)
    {
        for (int k = 0; k < 3; ++k) {
            if (p[k] > hi[k])
                hi[k] = p[k];
            if (p[k] < lo[k])
                lo[k] = p[k];
        }
    }
};

struct Work {
    uint32_t node;
    size_t begin, end;
    Box box;
    Vector3 center;
};

constexpr size_t kParallelMin = 1u << 16;  // triangles from which subtrees are built on several threads

// One node's split (src/axisalignedboundingboxtree.cpp:71-141): a leaf records
// where its members sit and returns false; otherwise the two children are
// appended to `nodes` and described in child[0..1].
struct Splitter {
    const std::vector<Box> &box;
    const std::vector<Vector3> &center;
    std::vector<uint32_t> &members, &scratch;
    std::vector<uint32_t> &leafOf, &place;

    bool split(const Work &w, std::vector<ReferenceTree::Node> &nodes, Work child[2]) const
    {
        const size_t n = w.end - w.begin;
        if (n <= kLeafMax) {
            for (size_t i = w.begin; i < w.end; ++i) {
                leafOf[members[i]] = w.node;
                place[members[i]] = (uint32_t)(i - w.begin);
            }
            return false;
        }
        // the longest axis of the node's box (spans narrowed to float; the
        // first of equal ones), split at the node's center on it
        float span[3];
        for (int k = 0; k < 3; ++k)
            span[k] = (float)(w.box.hi[k] - w.box.lo[k]);
        int axis = 0;
        for (int k = 1; k < 3; ++k)
            if (span[axis] < span[k])
                axis = k;
        const double split = w.center.constData()[axis];
        // left members fill scratch from the top down (so in reverse scan
        // order once read upwards), right ones from the bottom up
        size_t nl = 0, nr = 0;
        uint32_t *lo = scratch.data() + w.begin;
        for (size_t i = w.begin; i < w.end; ++i) {
            const uint32_t t = members[i];
            if (center[t].constData()[axis] < split)
                lo[n - 1 - nl++] = t;
            else
                lo[nr++] = t;
        }
        // members = left (reverse scan order), then right (scan order)
        for (size_t j = 0; j < nl; ++j)
            members[w.begin + j] = lo[n - nl + j];
        for (size_t j = 0; j < nr; ++j)
            members[w.begin + nl + j] = lo[j];
        size_t cut = nl;
        if (nl == 0)
            cut = n / 2;
        else if (nr == 0)
            cut = n - n / 2;
        const size_t bounds[3] = {w.begin, w.begin + cut, w.end};
        for (int c = 0; c < 2; ++c) {
            Work &x = child[c];
            x.begin = bounds[c];
            x.end = bounds[c + 1];
            x.box = Box();
            x.center = Vector3(0.0, 0.0, 0.0);
            for (size_t j = x.begin; j < x.end; ++j) {
                const uint32_t t = members[j];
                x.box.add(box[t].lo);
                x.box.add(box[t].hi);
                x.center += center[t];
            }
            x.center /= (double)(float)(x.end - x.begin);
            x.node = (uint32_t)nodes.size();
            nodes.emplace_back();
            nodes.back().count = (uint32_t)(x.end - x.begin);
        }
        nodes[w.node].left = child[0].node;
        nodes[w.node].right = child[1].node;
        return true;
    }
};

// the whole subtree below w, depth first, into `nodes`
void buildSubtree(const Splitter &sp, const Work &w, std::vector<ReferenceTree::Node> &nodes)
{
    std::vector<Work> stack{w};
    while (!stack.empty()) {
        const Work x = stack.back();
        stack.pop_back();
        Work child[2];
        if (!sp.split(x, nodes, child))
            continue;
        stack.push_back(child[1]);
        stack.push_back(child[0]);
    }
}

}  // namespace

ReferenceTree::ReferenceTree(const std::vector<Vector3> &vertices, const std::vector<std::vector<size_t>> &triangles)
{
    const size_t nt = triangles.size();
    m_leafOf.assign(nt, 0);
    m_place.assign(nt, 0);
    // per triangle: box and center (the running sum from 0, divided by 3),
    // independent per triangle, so spread over threads for large meshes
    std::vector<Box> box(nt);
    std::vector<Vector3> center(nt);
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned threads = hw < 16 ? (hw ? hw : 1) : 16;
    auto prepare = [&](size_t t0, size_t t1) {
        for (size_t t = t0; t < t1; ++t) {
            Vector3 sum(0.0, 0.0, 0.0);
            for (size_t k = 0; k < 3; ++k) {
                const Vector3 &p = vertices[triangles[t][k]];
                box[t].add(p.constData());
                sum += p;
            }
            center[t] = sum /= (double)(float)3;
        }
    };
    if (nt < kParallelMin || threads < 2) {
        prepare(0, nt);
    } else {
        std::vector<std::thread> pool;
        for (unsigned i = 0; i < threads; ++i)
            pool.emplace_back(prepare, nt * i / threads, nt * (i + 1) / threads);
        for (auto &t : pool)
            t.join();
    }
    // the root's box (min / max, any order) and center (a sum in triangle order)
    Box all;
    Vector3 rootCenter(0.0, 0.0, 0.0);
    for (size_t t = 0; t < nt; ++t) {
        all.add(box[t].lo);
        all.add(box[t].hi);
        rootCenter += center[t];
    }
    if (nt)
        rootCenter /= (double)(float)nt;

    // nodes own ranges of `members`; a split rewrites its range as the left
    // side's members in reverse scan order, then the right side's in order
    std::vector<uint32_t> members(nt), scratch(nt);
    for (size_t t = 0; t < nt; ++t)
        members[t] = (uint32_t)t;
    m_nodes.emplace_back();
    m_nodes[0].count = (uint32_t)nt;
    const Splitter sp{box, center, members, scratch, m_leafOf, m_place};
    const Work root{0, 0, nt, all, rootCenter};
    if (nt < kParallelMin || threads < 2) {
        buildSubtree(sp, root, m_nodes);
    } else {
        // Every node's sums run over its own members in their order whoever
        // builds it, so the tree does not depend on how the work is shared:
        // split breadth-first on this thread until there are enough disjoint
        // subtrees, then build those on worker threads into local node lists,
        // and splice them in.
        std::deque<Work> top{root};
        std::vector<Work> pending;
        while (!top.empty() && top.size() + pending.size() < 8 * (size_t)threads) {
            const Work w = top.front();
            top.pop_front();
            Work child[2];
            if (!sp.split(w, m_nodes, child))
                continue;
            for (const Work &c : child) {
                if (c.end - c.begin > 2 * kLeafMax)
                    top.push_back(c);
                else
                    pending.push_back(c);
            }
        }
        pending.insert(pending.end(), top.begin(), top.end());
        std::vector<std::vector<Node>> local(pending.size());
        std::atomic<size_t> next{0};
        auto worker = [&]() {
            for (size_t k = next++; k < pending.size(); k = next++) {
                Work w = pending[k];
                w.node = 0;
                local[k].emplace_back();
                local[k][0].count = (uint32_t)(w.end - w.begin);
                buildSubtree(sp, w, local[k]);
            }
        };
        std::vector<std::thread> pool;
        for (unsigned i = 1; i < threads; ++i)
            pool.emplace_back(worker);
        worker();
        for (auto &t : pool)
            t.join();
        // local node 0 is the pending node itself, local j >= 1 goes to base + j - 1
        for (size_t k = 0; k < pending.size(); ++k) {
            const uint32_t self = pending[k].node;
            const uint32_t base = (uint32_t)m_nodes.size();
            auto map = [&](uint32_t j) { return j == 0 ? self : base + j - 1; };
            for (size_t j = 1; j < local[k].size(); ++j) {
                Node x = local[k][j];
                if (!x.leaf()) {
                    x.left = map(x.left);
                    x.right = map(x.right);
                }
                m_nodes.push_back(x);
            }
            if (!local[k][0].leaf()) {
                m_nodes[self].left = map(local[k][0].left);
                m_nodes[self].right = map(local[k][0].right);
            }
            for (size_t i = pending[k].begin; i < pending[k].end; ++i)
                m_leafOf[members[i]] = map(m_leafOf[members[i]]);
        }
    }
    // the turns down to every leaf
    m_paths.assign(m_nodes.size(), std::string());
    std::vector<uint32_t> todo{0};
    while (!todo.empty()) {
        const uint32_t v = todo.back();
        todo.pop_back();
        const Node &x = m_nodes[v];
        if (x.leaf())
            continue;
        m_paths[x.left] = m_paths[v] + '0';
        m_paths[x.right] = m_paths[v] + '1';
        todo.push_back(x.left);
        todo.push_back(x.right);
    }
    for (size_t v = 0; v < m_nodes.size(); ++v)
        if (!m_nodes[v].leaf())
            std::string().swap(m_paths[v]);
}

ReferenceTree::ReferenceTree(const AxisAlignedBoudingBoxTree::Node *root, size_t nboxes)
{
    m_leafOf.assign(nboxes, 0);
    m_place.assign(nboxes, 0);
    m_nodes.emplace_back();
    m_paths.emplace_back();
    struct Item {
        const AxisAlignedBoudingBoxTree::Node *node;
        uint32_t id;
    };
    std::vector<Item> todo;
    if (root)
        todo.push_back({root, 0});
    while (!todo.empty()) {
        const Item it = todo.back();
        todo.pop_back();
        const auto &members = it.node->boxIndices;
        m_nodes[it.id].count = (uint32_t)members.size();
        if (it.node->isLeaf()) {
            for (size_t k = 0; k < members.size(); ++k) {
                m_leafOf[members[k]] = it.id;
                m_place[members[k]] = (uint32_t)k;
            }
            continue;
        }
        const uint32_t l = (uint32_t)m_nodes.size(), r = l + 1;
        m_nodes.resize(m_nodes.size() + 2);
        m_paths.resize(m_paths.size() + 2);
        m_nodes[it.id].left = l;
        m_nodes[it.id].right = r;
        m_paths[l] = m_paths[it.id] + '0';
        m_paths[r] = m_paths[it.id] + '1';
        todo.push_back({it.node->right, r});
        todo.push_back({it.node->left, l});
    }
    for (size_t v = 0; v < m_nodes.size(); ++v)
        if (!m_nodes[v].leaf())
            std::string().swap(m_paths[v]);
}

std::vector<size_t> referencePairOrder(const ReferenceTree &first, const ReferenceTree &second, const uint32_t *ab,
                                       size_t n)
{
    // A pair's key is the walk's turns from the two roots to its two leaves
    // (which tree turns at each step is fixed by the node pair: the inner
    // node, or of two inner nodes the one with more triangles, the first on a
    // tie), then the two places in the leaves, which the walk's nested loop
    // takes first-major.  Keys of different leaf pairs differ before either
    // ends, so plain string order is the walk's order.
    const auto &A = first.nodes();
    const auto &B = second.nodes();
    std::vector<std::string> key(n);
    for (size_t i = 0; i < n; ++i) {
        const uint32_t a = ab[2 * i], b = ab[2 * i + 1];
        const std::string &pa = first.pathTo(a), &pb = second.pathTo(b);
        std::string &k = key[i];
        k.reserve(pa.size() + pb.size() + 2);
        uint32_t na = 0, nb = 0;
        size_t ia = 0, ib = 0;
        while (!(A[na].leaf() && B[nb].leaf())) {
            const bool turnB = A[na].leaf() || (!B[nb].leaf() && A[na].count < B[nb].count);
            if (turnB) {
                const char c = pb[ib++];
                nb = c == '0' ? B[nb].left : B[nb].right;
                k += c;
            } else {
                const char c = pa[ia++];
                na = c == '0' ? A[na].left : A[na].right;
                k += c;
            }
        }
        k += (char)('A' + first.placeInLeaf(a));
        k += (char)('A' + second.placeInLeaf(b));
    }
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i)
        order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return key[x] < key[y]; });
    return order;
}

}  // namespace sbhost
