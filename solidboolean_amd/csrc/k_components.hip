// Connected components of a mesh's triangles across opposite half-edges: the
// groups that buildFaceGroups (src/solidboolean.cpp:167-239) forms for
// triangles no intersection loop touches (its leftover pass, :230-238, starts a
// group at each unvisited triangle in index order and floods it with a
// breadth-first search across half-edges, :203-226).  Triangle t's edge i -> j
// leads to the triangle that owns the half-edge j -> i.  Only unblocked
// triangles enter the half-edge table, as only untouched triangles enter the
// reference's map (addUnintersectedTriangles, :250-285), and a directed edge
// that two of them share is the reference's "repeated halfedge" failure
// (:271-282): the table flags it and the call fails.  Without repeats the
// relation is symmetric, so the flood fill's groups are exactly the
// connected components.
//
// Half-edges go into an open-addressed hash table (64-bit key (from << 32) |
// to), then a concurrent union-find links every triangle to the owner of each
// reversed half-edge, always hooking the larger root under the smaller, so a
// component's root is its smallest triangle index, where the leftover pass
// starts that group.
#include "sbg_internal.h"

namespace sbg {

constexpr unsigned long long kEmptyKey = ~0ull;
constexpr int kComponentBlock = 256;

__device__ __forceinline__ uint32_t he_slot(unsigned long long key, uint32_t mask)
{
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 32) & mask;
}

// every half-edge of every unblocked triangle; owners start at ~0u.  A key
// met twice sets *repeated (the owner then is the smaller triangle, so the
// table stays deterministic, though the call reports failure).
__global__ __launch_bounds__(kComponentBlock) void k_he_insert(const uint32_t *__restrict__ T, uint32_t nt,
                                                               const uint8_t *__restrict__ blocked,
                                                               unsigned long long *__restrict__ keys,
                                                               uint32_t *__restrict__ owners, uint32_t mask,
                                                               uint32_t *__restrict__ repeated)
{
    const uint32_t e = blockIdx.x * kComponentBlock + threadIdx.x;
    if (e >= 3u * nt)
        return;
    const uint32_t t = e / 3, k = e - 3 * t;
    if (blocked && blocked[t])
        return;
    const uint32_t from = T[3 * (size_t)t + k], to = T[3 * (size_t)t + (k + 1) % 3];
    const unsigned long long key = (unsigned long long)from << 32 | to;
    uint32_t s = he_slot(key, mask);
    // the table holds at least twice as many slots as half-edges, so a free
    // slot is always found
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const unsigned long long cur = atomicCAS(&keys[s], kEmptyKey, key);
        if (cur == kEmptyKey || cur == key) {
            if (cur == key)
                atomicOr(repeated, 1u);
            atomicMin(&owners[s], t);
            return;
        }
        s = (s + 1) & mask;
    }
}

// owner of half-edge `key`, or ~0u
__device__ __forceinline__ uint32_t he_owner(const unsigned long long *__restrict__ keys,
                                             const uint32_t *__restrict__ owners, uint32_t mask,
                                             unsigned long long key)
{
    uint32_t s = he_slot(key, mask);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const unsigned long long cur = keys[s];
        if (cur == key)
            return owners[s];
        if (cur == kEmptyKey)
            return ~0u;
        s = (s + 1) & mask;
    }
    return ~0u;
}

__device__ __forceinline__ uint32_t cc_find(uint32_t *parent, uint32_t x)
{
    uint32_t p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (p != x) {
        const uint32_t gp = __hip_atomic_load(&parent[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (gp != p)  // path halving; a lost race only leaves a longer path
            __hip_atomic_store(&parent[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
        p = gp;
    }
    return x;
}

__global__ __launch_bounds__(kComponentBlock) void k_cc_init(uint32_t nt, uint32_t *__restrict__ parent)
{
    const uint32_t t = blockIdx.x * kComponentBlock + threadIdx.x;
    if (t < nt)
        parent[t] = t;
}

// one lane per half-edge of an unblocked triangle: union with the owner of the
// reversed half-edge when that one is unblocked too
__global__ __launch_bounds__(kComponentBlock) void k_cc_union(const uint32_t *__restrict__ T, uint32_t nt,
                                                              const uint8_t *__restrict__ blocked,
                                                              const unsigned long long *__restrict__ keys,
                                                              const uint32_t *__restrict__ owners, uint32_t mask,
                                                              uint32_t *parent)
{
    const uint32_t e = blockIdx.x * kComponentBlock + threadIdx.x;
    if (e >= 3u * nt)
        return;
    const uint32_t t = e / 3, k = e - 3 * t;
    if (blocked && blocked[t])
        return;
    const uint32_t from = T[3 * (size_t)t + k], to = T[3 * (size_t)t + (k + 1) % 3];
    // the table holds unblocked triangles only
    const uint32_t o = he_owner(keys, owners, mask, (unsigned long long)to << 32 | from);
    if (o == ~0u || o == t)
        return;
    uint32_t a = cc_find(parent, t), b = cc_find(parent, o);
    while (a != b) {
        if (a > b) {
            const uint32_t x = a;
            a = b;
            b = x;
        }
        // hook root b under the smaller root a, unless b stopped being a root
        const uint32_t old = atomicCAS(&parent[b], b, a);
        if (old == b)
            break;
        b = cc_find(parent, old);
        a = cc_find(parent, a);
    }
}

__global__ __launch_bounds__(kComponentBlock) void k_cc_labels(const uint8_t *__restrict__ blocked, uint32_t nt,
                                                               uint32_t *parent, uint32_t *__restrict__ labels)
{
    const uint32_t t = blockIdx.x * kComponentBlock + threadIdx.x;
    if (t >= nt)
        return;
    labels[t] = (blocked && blocked[t]) ? ~0u : cc_find(parent, t);
}

// Half-edge table queries for the host's face grouping.  repeat: triangle t
// repeats a directed edge when an earlier triangle owns one of its half-edges
// or two of its own are equal (where inserting its half-edges in order would
// first fail); *firstRepeat = the smallest such t.
__global__ __launch_bounds__(kComponentBlock) void k_he_opposite(const uint32_t *__restrict__ T, uint32_t nt,
                                                                 const unsigned long long *__restrict__ keys,
                                                                 const uint32_t *__restrict__ owners, uint32_t mask,
                                                                 uint32_t *__restrict__ opposite,
                                                                 uint32_t *__restrict__ firstRepeat)
{
    const uint32_t e = blockIdx.x * kComponentBlock + threadIdx.x;
    if (e >= 3u * nt)
        return;
    const uint32_t t = e / 3, k = e - 3 * t;
    const uint32_t from = T[3 * (size_t)t + k], to = T[3 * (size_t)t + (k + 1) % 3];
    opposite[e] = he_owner(keys, owners, mask, (unsigned long long)to << 32 | from);
    bool repeat = he_owner(keys, owners, mask, (unsigned long long)from << 32 | to) < t;
    for (uint32_t j = 0; j < k; ++j)
        repeat |= T[3 * (size_t)t + j] == from && T[3 * (size_t)t + (j + 1) % 3] == to;
    if (repeat)
        atomicMin(firstRepeat, t);
}

__global__ __launch_bounds__(kComponentBlock) void k_he_query(const unsigned long long *__restrict__ queries,
                                                              uint32_t nq,
                                                              const unsigned long long *__restrict__ keys,
                                                              const uint32_t *__restrict__ owners, uint32_t mask,
                                                              uint32_t *__restrict__ out)
{
    const uint32_t i = blockIdx.x * kComponentBlock + threadIdx.x;
    if (i < nq)
        out[i] = he_owner(keys, owners, mask, queries[i]);
}

size_t component_table_slots(uint32_t nt)
{
    size_t cap = 1024;
    while (cap < 6 * (size_t)nt)
        cap <<= 1;
    return cap;
}

// keys: component_table_slots(nt) entries; owners: that many + 1 (the last
// word is the repeated-half-edge flag); parent, labels: nt
void launch_components(const uint32_t *T, uint32_t nt, const uint8_t *blocked, unsigned long long *keys,
                       uint32_t *owners, uint32_t *parent, uint32_t *labels, hipStream_t s)
{
    if (nt == 0)
        return;
    const size_t slots = component_table_slots(nt);
    const uint32_t mask = (uint32_t)(slots - 1);
    launch_fill_u32(reinterpret_cast<uint32_t *>(keys), ~0u, 2 * slots, s);
    launch_fill_u32(owners, ~0u, slots, s);
    launch_fill_u32(owners + slots, 0u, 1, s);
    const uint32_t he = 3u * nt;
    hipLaunchKernelGGL(k_he_insert, dim3((he + kComponentBlock - 1) / kComponentBlock), dim3(kComponentBlock), 0, s,
                       T, nt, blocked, keys, owners, mask, owners + slots);
    hipLaunchKernelGGL(k_cc_init, dim3((nt + kComponentBlock - 1) / kComponentBlock), dim3(kComponentBlock), 0, s,
                       nt, parent);
    hipLaunchKernelGGL(k_cc_union, dim3((he + kComponentBlock - 1) / kComponentBlock), dim3(kComponentBlock), 0, s,
                       T, nt, blocked, keys, owners, mask, parent);
    hipLaunchKernelGGL(k_cc_labels, dim3((nt + kComponentBlock - 1) / kComponentBlock), dim3(kComponentBlock), 0, s,
                       blocked, nt, parent, labels);
}

}  // namespace sbg

namespace sbg {

// keys: component_table_slots(nt) entries; owners: that many + 1 (the last
// word takes the repeat flag of the insert and is unused here); opposite: 3 nt;
// firstRepeat: 1 word; queries / queryOwner: nq
void launch_half_edges(const uint32_t *T, uint32_t nt, const unsigned long long *queries, uint32_t nq,
                       unsigned long long *keys, uint32_t *owners, uint32_t *opposite, uint32_t *queryOwner,
                       uint32_t *firstRepeat, hipStream_t s)
{
    const size_t slots = component_table_slots(nt);
    const uint32_t mask = (uint32_t)(slots - 1);
    launch_fill_u32(firstRepeat, ~0u, 1, s);
    if (nt) {
        launch_fill_u32(reinterpret_cast<uint32_t *>(keys), ~0u, 2 * slots, s);
        launch_fill_u32(owners, ~0u, slots, s);
        launch_fill_u32(owners + slots, 0u, 1, s);
        const uint32_t he = 3u * nt;
        hipLaunchKernelGGL(k_he_insert, dim3((he + kComponentBlock - 1) / kComponentBlock), dim3(kComponentBlock), 0,
                           s, T, nt, nullptr, keys, owners, mask, owners + slots);
        hipLaunchKernelGGL(k_he_opposite, dim3((he + kComponentBlock - 1) / kComponentBlock), dim3(kComponentBlock),
                           0, s, T, nt, keys, owners, mask, opposite, firstRepeat);
    } else {
        launch_fill_u32(reinterpret_cast<uint32_t *>(keys), ~0u, 2 * slots, s);
    }
    if (nq)
        hipLaunchKernelGGL(k_he_query, dim3((nq + kComponentBlock - 1) / kComponentBlock), dim3(kComponentBlock), 0, s,
                           queries, nq, keys, owners, mask, queryOwner);
}

}  // namespace sbg
