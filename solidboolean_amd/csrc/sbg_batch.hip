// C-ABI of the batched booleans (include/sbgpu.h, sbg_batch_*): many
// independent mesh pairs run through one sequence of launches over one
// combined mesh (kernels in k_batch.hip plus the single-pair kernels).
//
// The sequence allocates nothing and never waits on the host, so once the
// buffers are sized it is captured into one hipGraph and each later run is a
// single graph launch.  Counts the host needs (candidates, queue overflows,
// grid entries) are read back after the run; a run that overflowed a buffer
// grows it, drops the graph and runs again, so no result is ever truncated.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sbg_host.h"

namespace {

constexpr int kBatchMaxLevels = 32;
constexpr uint32_t kBatchSpillInit = 2048, kBatchSpillMax = 1u << 15;
// readback words: [0, 33) frontier sizes, 33 candidates, 34 search flags,
// 35 depth-first claims, [36, 40) ray counters (36 ray flags), 40 grid entries
constexpr int kCandW = kBatchMaxLevels + 1, kFlagW = kBatchMaxLevels + 2, kHeadW = kBatchMaxLevels + 3;
constexpr int kRayFlagW = kBatchMaxLevels + 4, kGridW = kBatchMaxLevels + 8;
constexpr int kWords = kBatchMaxLevels + 10;
// the batch's ray-grid build: one workgroup per tile of 512 triangles (the
// batch's grids are millions of triangles, and the branch beside them is the
// batch's own search)
constexpr uint32_t kBatchGridBlocks = 0;

int bits_for(uint64_t n)
{
    int b = 0;
    while (b < 63 && (1ull << b) < n)
        ++b;
    return b;
}

}  // namespace

struct sbg_batch_s {
    sbg_context ctx = nullptr;
    size_t nmeshes = 0, npairs = 0, nv = 0, nt = 0;
    std::vector<uint64_t> vBegin, tBegin;
    int localBits = 0, mortonBits = 30, meshBits = 0, triBits = 1;
    sbg_batch_tuning tune = {16, 1024, 4, 1};

    DevBuf<double> V;
    DevBuf<uint32_t> T, meshOf, target, triBegin;
    DevBuf<double> tri64;
    DevBuf<float> tri32, box32, blockBounds, rootBox, meshBox, refitScratch;
    DevBuf<int> meshBounds, waveBounds;
    DevBuf<uint64_t> keys, keysAlt, codes;
    DevBuf<uint32_t> ids, leafParent, work, meshRoot, hist;
    DevBuf<BNode> nodes;
    DevBuf<uint2> front[2], spill;
    DevBuf<uint32_t> words;  // kWords device counters
    DevBuf<uint64_t> pairs, pairsAlt;
    DevBuf<uint8_t> status;
    DevBuf<double> seg;
    DevBuf<uint32_t> gdWork;
    DevBuf<uint8_t> flags, bits, votes;
    DevBuf<double> pts;
    DevBuf<uint32_t> rayList;
    DevBuf<uint2> gridDesc;
    DevBuf<uint32_t> gridCells, gridTemp;
    DevBuf<uint2> gridCode;
    DevBuf<uint32_t> gridTri;

    size_t frontCap = 0, candCap = 0, gridCap = 0;
    uint32_t spillPerWave = kBatchSpillInit;
    uint32_t ncells = 0;
    uint32_t *host = nullptr;  // pinned copy of the words
    bool keysInAlt = false, pairsInAlt = false;

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graphPrecision = -1;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;  // the ray grid's branch on the second stream
    double lastMs = 0.0;
    bool ran = false;
    uint32_t nCand = 0;
};

namespace {

void drop_graph(sbg_batch b)
{
    if (b->exec)
        (void)hipGraphExecDestroy(b->exec);
    if (b->graph)
        (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graphPrecision = -1;
}

// per-mesh grid sizes under the current density; returns the total cell count
uint32_t grid_layout(sbg_batch b, std::vector<uint2> &desc)
{
    desc.resize(b->nmeshes);
    uint64_t base = 0;
    for (size_t m = 0; m < b->nmeshes; ++m) {
        const double nt = (double)(b->tBegin[m + 1] - b->tBegin[m]);
        const uint32_t G = grid_dim((uint32_t)nt, b->tune.grid_density);
        desc[m] = make_uint2(G, (uint32_t)base);
        base += 3ull * G * G;
    }
    return (uint32_t)base;
}

// device buffers for the current capacities and tuning (allocation happens
// only here, never inside a run)
int size_buffers(sbg_batch b)
{
    const size_t nt = b->nt;
    std::vector<uint2> desc;
    const uint64_t ncells = grid_layout(b, desc);
    if (ncells >= 0xffffffffull)
        return fail(SBG_ERR_INVALID, "ray grids need %llu cells", (unsigned long long)ncells);
    b->ncells = (uint32_t)ncells;
    TRY(b->gridDesc.ensure(b->nmeshes));
    HIP_TRY(hipMemcpy(b->gridDesc.p, desc.data(), sizeof(uint2) * b->nmeshes, hipMemcpyHostToDevice));
    if (!b->frontCap)
        b->frontCap = std::max<size_t>((size_t)1 << 16, 2 * nt);
    if (!b->candCap)
        b->candCap = std::max<size_t>((size_t)1 << 14, nt / 4);
    if (!b->gridCap)
        b->gridCap = 8 * nt + 1024;
    if (b->frontCap > 0xffffffffu || b->candCap > 0x7fffffffu || b->gridCap > 0xffffffffu)
        return fail(SBG_ERR_OVERFLOW, "batch queues exceed 32-bit indexing");
    TRY(b->front[0].ensure(b->frontCap));
    TRY(b->front[1].ensure(b->frontCap));
    TRY(b->spill.ensure((size_t)b->tune.dfs_waves * b->spillPerWave));
    TRY(b->pairs.ensure(b->candCap));
    TRY(b->pairsAlt.ensure(b->candCap));
    TRY(b->status.ensure(b->candCap));
    TRY(b->seg.ensure(6 * b->candCap));
    TRY(b->gdWork.ensure(tritri_work_u32((uint32_t)b->candCap)));
    TRY(b->hist.ensure(std::max(sort_hist_entries(nt), sort_hist_entries(b->candCap))));
    TRY(b->gridCells.ensure((size_t)b->ncells + 1));
    TRY(b->gridTemp.ensure(scan_large_temp((size_t)b->ncells + 1)));
    TRY(b->gridCode.ensure(b->gridCap));
    TRY(b->gridTri.ensure(b->gridCap));
    return SBG_OK;
}

// The whole run as stream-ordered launches: no allocation, no host wait, no
// memset (buffers are cleared by fill kernels), so it can be captured.
int enqueue(sbg_batch b, int precision)
{
    hipStream_t s = b->ctx->stream;
    const uint32_t nt = (uint32_t)b->nt;
    uint32_t *w = b->words.p;
    const int nb = records_blocks(nt);
    launch_fill_u32(w, 0u, kWords, s);
    launch_records(b->V.p, b->T.p, nt, b->tri64.p, b->tri32.p, b->box32.p, b->blockBounds.p, nb, s);
    // combined LBVH: every mesh a subtree
    launch_batch_bounds_init(b->meshBounds.p, (uint32_t)b->nmeshes, s);
    launch_batch_bounds(b->box32.p, b->meshOf.p, b->triBegin.p, (uint32_t)b->nmeshes, nt, b->meshBounds.p,
                        b->waveBounds.p, s);
    launch_batch_keys(b->box32.p, b->meshOf.p, b->triBegin.p, b->meshBounds.p, nt, b->localBits, b->mortonBits,
                      b->keys.p, s);
    const int keyRange[2] = {b->localBits, b->localBits + b->mortonBits + b->meshBits};
    b->keysInAlt = radix_sort_u64(b->keys.p, b->keysAlt.p, nt, keyRange, 1, b->hist.p, s) != 0;
    const uint64_t *sortedKeys = b->keysInAlt ? b->keysAlt.p : b->keys.p;
    launch_batch_codes(sortedKeys, nt, b->localBits, b->mortonBits, b->triBegin.p, b->codes.p, b->ids.p, s);
    launch_hierarchy64(b->codes.p, b->ids.p, nt, b->nodes.p, b->leafParent.p, b->work.p, s);
    launch_refit(b->ids.p, b->box32.p, nt, b->nodes.p, b->leafParent.p, b->work.p, b->refitScratch.p, b->rootBox.p,
                 s);
    launch_fill_u32(b->meshRoot.p, ~0u, b->nmeshes, s);
    launch_fill_u32(reinterpret_cast<uint32_t *>(b->meshBox.p), 0u, 6 * b->nmeshes, s);  // empty meshes: zero box
    launch_mesh_roots(b->nodes.p, b->ids.p, b->triBegin.p, (uint32_t)b->nmeshes, nt, b->box32.p, b->rootBox.p,
                      b->meshRoot.p, b->meshBox.p, s);
    TRY(check_launch("batch tree"));
    // The ray grids depend on the meshes alone: they are built on a branch
    // (the context's second stream, default priority) beside the search and
    // the GD, and joined before the rays.  Captured, the branch becomes
    // parallel nodes of the graph.
    GridSet gs = {};
    gs.meshOf = b->meshOf.p;
    gs.target = b->target.p;
    gs.desc = b->gridDesc.p;
    gs.boxes = b->meshBox.p;
    gs.cells = b->gridCells.p;
    gs.gcode = b->gridCode.p;
    gs.gtri = b->gridTri.p;
    gs.cap = (uint32_t)b->gridCap;
    gs.ncells = b->ncells;
    gs.order = b->ids.p;
    const hipStream_t aux = b->ctx->aux;
    GridSet gsRays = gs;  // the ray kernels' view: FP32 certificates in the filtered precision
    gsRays.tri32 = precision == SBG_PRECISION_FP32_FILTERED ? b->tri32.p : nullptr;
    // the context's lane cap (default 64): 48 left more of the batch's rays
    // to the one-wave kernel (2.265 vs 2.253 ms); a captured graph keeps the
    // value it was captured with, which changes timing only
    gsRays.fastCands = (uint32_t)b->ctx->tune.ray_cell_cap;
    HIP_TRY(hipEventRecord(b->fork, s));
    HIP_TRY(hipStreamWaitEvent(aux, b->fork, 0));
    launch_grid_build(b->box32.p, nt, gs, b->gridCells.p, b->gridTemp.p, b->gridCode.p, b->gridTri.p, kBatchGridBlocks, aux);
    // the ray queries (centroids in tree order, target mesh as tag) do not
    // depend on the intersections: the rays skip intersected triangles by flag
    launch_centroids_ordered(b->tri64.p, nullptr, b->ids.p, b->target.p, nt, b->pts.p, aux);
    HIP_TRY(hipEventRecord(b->join, aux));
    TRY(check_launch("batch grid"));
    // candidate search: every pair's root pair, breadth-first levels, then the
    // persistent depth-first kernel over what is left
    const TraverseArgs t = {b->nodes.p, b->box32.p, b->tri64.p, b->nodes.p, b->box32.p, b->tri64.p};
    const uint32_t fcap = (uint32_t)b->frontCap, ccap = (uint32_t)b->candCap;
    uint32_t *candCount = w + kCandW, *flagsW = w + kFlagW;
    launch_batch_seed(t, b->meshRoot.p, b->meshBox.p, (uint32_t)b->npairs, b->front[0].p, w, fcap, b->pairs.p,
                      candCount, ccap, flagsW, s);
    int level = 0;
    for (; level < b->tune.bfs_levels; ++level)
        launch_expand(t, b->front[level & 1].p, w + level, b->front[(level + 1) & 1].p, w + level + 1, fcap, b->pairs.p,
                      candCount, ccap, flagsW, 1024, s);
    launch_traverse_dfs(t, b->front[level & 1].p, w + level, fcap, w + kHeadW, b->spill.p, b->spillPerWave,
                        b->tune.dfs_waves, b->pairs.p, candCount, ccap, flagsW, s);
    TRY(check_launch("batch search"));
    // canonical (a, b) order of the candidates (their count read on the device)
    const int pairRanges[4] = {0, b->triBits, 32, 32 + b->triBits};
    b->pairsInAlt =
        radix_sort_u64(b->pairs.p, b->pairsAlt.p, b->candCap, pairRanges, 2, b->hist.p, s, candCount) != 0;
    const uint64_t *sorted = b->pairsInAlt ? b->pairsAlt.p : b->pairs.p;
    launch_tritri(b->tri64.p, b->tri64.p, b->tri32.p, b->tri32.p, sorted, ccap, candCount, precision, b->status.p,
                  b->seg.p, b->gdWork.p, nullptr, s);
    TRY(check_launch("batch tri_tri"));
    // classification of every non-intersected triangle against its partner
    launch_fill_u32((uint32_t *)b->flags.p, 0u, (nt + 3) / 4, s);
    launch_mark_both(sorted, b->status.p, candCount, ccap, b->flags.p, s);
    HIP_TRY(hipStreamWaitEvent(s, b->join, 0));
    launch_rays_ordered(gsRays, b->tri64.p, b->pts.p, b->ids.p, b->flags.p, nt, false, b->bits.p, b->rayList.p,
                        w + kRayFlagW, s);
    launch_vote(b->bits.p, b->flags.p, nt, b->votes.p, s);
    launch_copy_u32(b->gridCells.p + b->ncells, w + kGridW, 1, s);
    TRY(check_launch("batch classify"));
    return SBG_OK;
}

}  // namespace

extern "C" {

int sbg_batch_create(sbg_context ctx, const double *xyz, const uint32_t *tris, const uint64_t *vert_begin,
                     const uint64_t *tri_begin, size_t nmeshes, sbg_batch *out)
{
    if (!out)
        return fail(SBG_ERR_INVALID, "null output");
    *out = nullptr;
    TRY(use(ctx));
    if (!vert_begin || !tri_begin || nmeshes == 0 || (nmeshes & 1))
        return fail(SBG_ERR_INVALID, "need an even, non-zero number of meshes with offset arrays");
    const uint64_t nv = vert_begin[nmeshes], nt = tri_begin[nmeshes];
    if ((nv && !xyz) || (nt && !tris))
        return fail(SBG_ERR_INVALID, "null vertex or triangle array");
    if (vert_begin[0] != 0 || tri_begin[0] != 0)
        return fail(SBG_ERR_INVALID, "offset arrays must start at 0");
    if (nt == 0 || nt >= 0x7fffffffull || nv >= 0xffffffffull)
        return fail(SBG_ERR_INVALID, "batch needs 1 to 2^31-1 triangles in total");
    uint64_t maxTris = 1;
    std::vector<uint32_t> T(3 * nt), meshOf(nt), target(nt), triBegin(nmeshes + 1);
    for (size_t m = 0; m < nmeshes; ++m) {
        if (vert_begin[m + 1] < vert_begin[m] || tri_begin[m + 1] < tri_begin[m])
            return fail(SBG_ERR_INVALID, "offsets of mesh %zu decrease", m);
        const uint64_t v0 = vert_begin[m], nvm = vert_begin[m + 1] - v0;
        maxTris = std::max<uint64_t>(maxTris, tri_begin[m + 1] - tri_begin[m]);
        triBegin[m] = (uint32_t)tri_begin[m];
        for (uint64_t t = tri_begin[m]; t < tri_begin[m + 1]; ++t) {
            for (int k = 0; k < 3; ++k) {
                const uint32_t v = tris[3 * t + k];
                if (v >= nvm)
                    return fail(SBG_ERR_INVALID, "triangle %llu of mesh %zu indexes vertex %u of %llu",
                                (unsigned long long)(t - tri_begin[m]), m, v, (unsigned long long)nvm);
                T[3 * t + k] = (uint32_t)(v0 + v);
            }
            meshOf[t] = (uint32_t)m;
            target[t] = (uint32_t)(m ^ 1u);
        }
    }
    triBegin[nmeshes] = (uint32_t)nt;
    sbg_batch b = new sbg_batch_s;
    b->ctx = ctx;
    ++ctx->refs;
    b->nmeshes = nmeshes;
    b->npairs = nmeshes / 2;
    b->nv = nv;
    b->nt = nt;
    b->vBegin.assign(vert_begin, vert_begin + nmeshes + 1);
    b->tBegin.assign(tri_begin, tri_begin + nmeshes + 1);
    b->localBits = bits_for(maxTris);
    b->meshBits = std::max(1, bits_for(nmeshes));
    b->mortonBits = std::min(30, 64 - b->localBits - b->meshBits);
    // the tree sort runs 8-bit passes over mesh id + Morton code: drop the
    // Morton bits that would cost a pass of their own, down to 8 per axis
    // (the single-mesh keys' precision)
    const int rounded = 8 * ((b->meshBits + b->mortonBits) / 8) - b->meshBits;
    if (rounded >= 24 && rounded < b->mortonBits)
        b->mortonBits = rounded;
    b->triBits = std::max(1, bits_for(nt));
    int rc = SBG_OK;
    auto bail = [&](int r) {
        sbg_batch_destroy(b);
        return r;
    };
    if (b->mortonBits < 6)
        return bail(fail(SBG_ERR_INVALID, "too many meshes / triangles per mesh for 64-bit tree keys"));
    if ((rc = b->V.ensure(3 * nv)) || (rc = b->T.ensure(3 * nt)) || (rc = b->meshOf.ensure(nt)) ||
        (rc = b->target.ensure(nt)) || (rc = b->triBegin.ensure(nmeshes + 1)) ||
        (rc = b->tri64.ensure((size_t)kTri64Stride * nt)) || (rc = b->tri32.ensure((size_t)kTri32Stride * nt)) ||
        (rc = b->box32.ensure((size_t)kBox32Stride * nt)) ||
        (rc = b->blockBounds.ensure(6 * (size_t)records_blocks((uint32_t)nt))) || (rc = b->rootBox.ensure(6)) ||
        (rc = b->meshBox.ensure(6 * nmeshes)) || (rc = b->meshBounds.ensure(6 * nmeshes)) ||
        (rc = b->waveBounds.ensure(6 * ((nt + 63) / 64))) ||
        (rc = b->keys.ensure(nt)) || (rc = b->keysAlt.ensure(nt)) || (rc = b->codes.ensure(nt)) ||
        (rc = b->ids.ensure(nt)) || (rc = b->leafParent.ensure(nt)) || (rc = b->work.ensure(2 * nt + 1)) ||
        (rc = b->refitScratch.ensure(refit_scratch_floats((uint32_t)nt))) ||
        (rc = b->meshRoot.ensure(nmeshes)) || (rc = b->nodes.ensure(nt > 1 ? nt - 1 : 1)) ||
        (rc = b->words.ensure(kWords)) || (rc = b->flags.ensure(nt + 4)) || (rc = b->bits.ensure(3 * nt + 4)) ||
        (rc = b->votes.ensure(nt)) || (rc = b->pts.ensure(4 * nt)) || (rc = b->rayList.ensure(ray_list_u32(nt))))
        return bail(rc);
    if (hipHostMalloc((void **)&b->host, sizeof(uint32_t) * kWords, hipHostMallocDefault) != hipSuccess) {
        b->host = nullptr;
        return bail(fail(SBG_ERR_OOM, "hipHostMalloc failed"));
    }
    if (hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess)
        return bail(fail(SBG_ERR_HIP, "hipEventCreate failed"));
    if (hipMemcpy(b->V.p, xyz, sizeof(double) * 3 * nv, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->T.p, T.data(), sizeof(uint32_t) * 3 * nt, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->meshOf.p, meshOf.data(), sizeof(uint32_t) * nt, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->target.p, target.data(), sizeof(uint32_t) * nt, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->triBegin.p, triBegin.data(), sizeof(uint32_t) * (nmeshes + 1), hipMemcpyHostToDevice) !=
            hipSuccess)
        return bail(fail(SBG_ERR_HIP, "batch upload failed"));
    if ((rc = size_buffers(b)) != SBG_OK)
        return bail(rc);
    *out = b;
    return SBG_OK;
}

int sbg_batch_set_vertices(sbg_batch b, const double *xyz)
{
    if (!b || (b->nv && !xyz))
        return fail(SBG_ERR_INVALID, "null argument");
    TRY(use(b->ctx));
    HIP_TRY(hipMemcpyAsync(b->V.p, xyz, sizeof(double) * 3 * b->nv, hipMemcpyHostToDevice, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    b->ran = false;
    return SBG_OK;
}

int sbg_batch_set_tuning(sbg_batch b, const sbg_batch_tuning *t)
{
    if (!b || !t)
        return fail(SBG_ERR_INVALID, "null argument");
    if (t->bfs_levels < 0 || t->bfs_levels > kBatchMaxLevels || t->dfs_waves < 64 || t->dfs_waves > 16384 ||
        t->grid_density < 1 || t->grid_density > 64)
        return fail(SBG_ERR_INVALID, "tuning out of range");
    TRY(use(b->ctx));
    const bool regrid = t->grid_density != b->tune.grid_density;
    b->tune = *t;
    drop_graph(b);
    if (regrid)
        b->gridCap = 0;
    return size_buffers(b);
}

int sbg_batch_get_tuning(sbg_batch b, sbg_batch_tuning *t)
{
    if (!b || !t)
        return fail(SBG_ERR_INVALID, "null argument");
    *t = b->tune;
    return SBG_OK;
}

// one run (graph replay when allowed and sized), then the readback
static int run_once(sbg_batch b, int precision, bool allowGraph)
{
    hipStream_t s = b->ctx->stream;
    HIP_TRY(hipEventRecord(b->ev0, s));
    if (allowGraph && b->tune.use_graph) {
        if (!b->exec || b->graphPrecision != precision) {
            drop_graph(b);
            HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue(b, precision);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (rc != SBG_OK) {
                if (g)
                    (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess)
                return fail(SBG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
            b->graph = g;
            HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graphPrecision = precision;
        }
        HIP_TRY(hipGraphLaunch(b->exec, s));
    } else {
        TRY(enqueue(b, precision));
    }
    HIP_TRY(hipEventRecord(b->ev1, s));
    HIP_TRY(hipMemcpyAsync(b->host, b->words.p, sizeof(uint32_t) * kWords, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->lastMs = ms;
    return SBG_OK;
}

int sbg_batch_run(sbg_batch b, int precision)
{
    if (!b)
        return fail(SBG_ERR_INVALID, "null batch");
    if (precision != SBG_PRECISION_FP64 && precision != SBG_PRECISION_FP32_FILTERED)
        return fail(SBG_ERR_INVALID, "unknown precision %d", precision);
    TRY(use(b->ctx));
    b->ran = false;
    // A graph is captured only once a plain run has fitted its buffers.  Every
    // run, the capturing one and every replay included, goes through the same
    // flag check: how the depth-first search spills depends on how waves claim
    // seeds, so a replay can overflow where the plain run did not.
    bool fitted = b->exec != nullptr;
    for (int attempt = 0; attempt < 16; ++attempt) {
        const bool viaGraph = fitted && b->tune.use_graph;
        TRY(run_once(b, precision, viaGraph));
        const uint32_t *h = b->host;
        uint32_t maxFront = 0;
        for (int l = 0; l <= kBatchMaxLevels; ++l)
            maxFront = std::max(maxFront, h[l]);
        const uint32_t nCand = h[kCandW], sflags = h[kFlagW], rflags = h[kRayFlagW], gridTotal = h[kGridW];
        bool grow = false;
        if (sflags & 1u) {
            if (maxFront > b->frontCap)
                b->frontCap = std::max<size_t>(2 * b->frontCap, (size_t)maxFront + maxFront / 2);
            if (nCand > b->candCap)
                b->candCap = std::max<size_t>(2 * b->candCap, (size_t)nCand + nCand / 4);
            grow = true;
        }
        if (sflags & 12u) {
            if (b->spillPerWave >= kBatchSpillMax)
                return fail(SBG_ERR_OVERFLOW, "depth-first search exceeded its largest spill regions");
            b->spillPerWave = std::min(kBatchSpillMax, std::max(kBatchSpillInit, 4 * b->spillPerWave));
            grow = true;
        }
        if (rflags & 2u) {
            b->gridCap = (size_t)gridTotal + gridTotal / 16 + 1024;
            grow = true;
        }
        if (!grow && (rflags & 1u))
            return fail(SBG_ERR_OVERFLOW, "a ray has more distinct hits than a wave gathers");
        if (grow) {
            drop_graph(b);
            fitted = false;
            TRY(size_buffers(b));
            continue;
        }
        b->nCand = nCand;
        if (b->tune.use_graph && !viaGraph) {
            // the pair sort and the GD grid cover the whole capacity: keep it
            // near the real count before a graph fixes it (buffers stay)
            const size_t want = std::max<size_t>((size_t)1 << 14, (size_t)nCand + nCand / 4);
            if (b->candCap > 2 * want)
                b->candCap = want;
            fitted = true;  // capture now and check that run like any other
            continue;
        }
        b->ran = true;
        return SBG_OK;
    }
    return fail(SBG_ERR_OVERFLOW, "batch buffers did not fit after repeated growth");
}

int sbg_batch_set_spill(sbg_batch b, uint32_t per_wave)
{
    if (!b)
        return fail(SBG_ERR_INVALID, "null batch");
    if (per_wave > kBatchSpillMax)
        return fail(SBG_ERR_INVALID, "spill region of %u pairs exceeds %u", per_wave, kBatchSpillMax);
    TRY(use(b->ctx));
    b->spillPerWave = per_wave;
    drop_graph(b);
    b->ran = false;
    return size_buffers(b);
}

int sbg_batch_bench_tri_tri(sbg_batch b, int precision, int iters, double *avg_ms, sbg_tri_tri_stats *stats)
{
    if (!b || !avg_ms || iters <= 0)
        return fail(SBG_ERR_INVALID, "bad argument");
    if (precision != SBG_PRECISION_FP64 && precision != SBG_PRECISION_FP32_FILTERED)
        return fail(SBG_ERR_INVALID, "unknown precision %d", precision);
    if (!b->ran)
        return fail(SBG_ERR_STATE, "no completed run");
    TRY(use(b->ctx));
    hipStream_t s = b->ctx->stream;
    const uint64_t *sorted = b->pairsInAlt ? b->pairsAlt.p : b->pairs.p;
    const uint32_t n = b->nCand;
    // the launches rewrite status and segments with the same values
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s));
    for (int i = 0; i < iters; ++i)
        launch_tritri(b->tri64.p, b->tri64.p, b->tri32.p, b->tri32.p, sorted, n, nullptr, precision, b->status.p,
                      b->seg.p, b->gdWork.p, nullptr, s);
    HIP_TRY(hipEventRecord(e1, s));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / iters;
    if (stats) {
        DevBuf<GdStats> d;
        TRY(d.ensure(1));
        launch_fill_u32(reinterpret_cast<uint32_t *>(d.p), 0u, sizeof(GdStats) / 4, s);
        launch_tritri(b->tri64.p, b->tri64.p, b->tri32.p, b->tri32.p, sorted, n, nullptr, precision, b->status.p,
                      b->seg.p, b->gdWork.p, d.p, s);
        GdStats h;
        HIP_TRY(hipMemcpyAsync(&h, d.p, sizeof(GdStats), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        d.release();
        stats->tested = n;
        stats->exit_plane1 = h.exits[0];
        stats->exit_plane2 = h.exits[1];
        stats->exit_interval = h.exits[2];
        stats->hits = h.exits[3];
        stats->coplanar = h.exits[4];
        stats->fallback_pairs = h.fallback;
        stats->filtered_pairs = h.filtered;
    }
    return check_launch("batch bench tri_tri");
}

int sbg_batch_last_ms(sbg_batch b, double *ms)
{
    if (!b || !ms)
        return fail(SBG_ERR_INVALID, "null argument");
    if (!b->ran)
        return fail(SBG_ERR_STATE, "no completed run");
    *ms = b->lastMs;
    return SBG_OK;
}

int sbg_batch_counts(sbg_batch b, size_t *candidates, size_t *hits)
{
    if (!b)
        return fail(SBG_ERR_INVALID, "null batch");
    if (!b->ran)
        return fail(SBG_ERR_STATE, "no completed run");
    TRY(use(b->ctx));
    if (candidates)
        *candidates = b->nCand;
    if (hits) {
        std::vector<uint8_t> st(b->nCand);
        if (b->nCand)
            HIP_TRY(hipMemcpy(st.data(), b->status.p, b->nCand, hipMemcpyDeviceToHost));
        size_t h = 0;
        for (uint8_t v : st)
            h += v == 1;
        *hits = h;
    }
    return SBG_OK;
}

int sbg_batch_candidates(sbg_batch b, uint32_t *ab, uint8_t *status, double *seg)
{
    if (!b)
        return fail(SBG_ERR_INVALID, "null batch");
    if (!b->ran)
        return fail(SBG_ERR_STATE, "no completed run");
    TRY(use(b->ctx));
    const size_t n = b->nCand;
    if (!n)
        return SBG_OK;
    if (ab) {
        std::vector<uint64_t> p(n);
        HIP_TRY(hipMemcpy(p.data(), b->pairsInAlt ? b->pairsAlt.p : b->pairs.p, sizeof(uint64_t) * n,
                          hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            ab[2 * i] = (uint32_t)(p[i] >> 32);
            ab[2 * i + 1] = (uint32_t)p[i];
        }
    }
    if (status)
        HIP_TRY(hipMemcpy(status, b->status.p, n, hipMemcpyDeviceToHost));
    if (seg) {
        std::vector<uint8_t> st(n);
        HIP_TRY(hipMemcpy(st.data(), b->status.p, n, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(seg, b->seg.p, sizeof(double) * 6 * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i)
            if (st[i] != 1)
                std::memset(seg + 6 * i, 0, sizeof(double) * 6);
    }
    return SBG_OK;
}

int sbg_batch_sides(sbg_batch b, uint8_t *side)
{
    if (!b || !side)
        return fail(SBG_ERR_INVALID, "null argument");
    if (!b->ran)
        return fail(SBG_ERR_STATE, "no completed run");
    TRY(use(b->ctx));
    HIP_TRY(hipMemcpy(side, b->votes.p, b->nt, hipMemcpyDeviceToHost));
    return SBG_OK;
}

int sbg_batch_autotune(sbg_batch b, int precision, int reps, sbg_batch_tuning *best, double *best_ms)
{
    if (!b || reps <= 0)
        return fail(SBG_ERR_INVALID, "bad argument");
    TRY(use(b->ctx));
    sbg_batch_tuning cur = b->tune;
    const int useGraph = cur.use_graph;
    cur.use_graph = 0;  // settings are compared on plain runs
    auto time_of = [&](const sbg_batch_tuning &t, double &ms) -> int {
        TRY(sbg_batch_set_tuning(b, &t));
        TRY(sbg_batch_run(b, precision));  // sizes buffers for this setting
        double sum = 0.0;
        for (int r = 0; r < reps; ++r) {
            TRY(sbg_batch_run(b, precision));
            sum += b->lastMs;
        }
        ms = sum / reps;
        return SBG_OK;
    };
    double bestMs = 0.0;
    TRY(time_of(cur, bestMs));
    const int levels[] = {0, 4, 8, 12, 16};
    const int waves[] = {512, 1024, 2048, 4096, 8192};
    const int density[] = {2, 4, 8, 16};
    for (int pass = 0; pass < 2; ++pass) {
        for (int v : levels) {
            sbg_batch_tuning t = cur;
            t.bfs_levels = v;
            double ms = 0.0;
            if (v != cur.bfs_levels && time_of(t, ms) == SBG_OK && ms < bestMs) {
                bestMs = ms;
                cur = t;
            }
        }
        for (int v : waves) {
            sbg_batch_tuning t = cur;
            t.dfs_waves = v;
            double ms = 0.0;
            if (v != cur.dfs_waves && time_of(t, ms) == SBG_OK && ms < bestMs) {
                bestMs = ms;
                cur = t;
            }
        }
        for (int v : density) {
            sbg_batch_tuning t = cur;
            t.grid_density = v;
            double ms = 0.0;
            if (v != cur.grid_density && time_of(t, ms) == SBG_OK && ms < bestMs) {
                bestMs = ms;
                cur = t;
            }
        }
    }
    cur.use_graph = useGraph;
    TRY(sbg_batch_set_tuning(b, &cur));
    if (best)
        *best = cur;
    if (best_ms)
        *best_ms = bestMs;
    return SBG_OK;
}

int sbg_batch_destroy(sbg_batch b)
{
    if (!b)
        return SBG_OK;
    sbg_context c = b->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    drop_graph(b);
    if (b->ev0)
        (void)hipEventDestroy(b->ev0);
    if (b->ev1)
        (void)hipEventDestroy(b->ev1);
    if (b->fork)
        (void)hipEventDestroy(b->fork);
    if (b->join)
        (void)hipEventDestroy(b->join);
    if (b->host)
        (void)hipHostFree(b->host);
    DevBuf<double> *d64[] = {&b->V, &b->tri64, &b->seg, &b->pts};
    for (auto *p : d64)
        p->release();
    DevBuf<float> *f32[] = {&b->tri32, &b->box32, &b->blockBounds, &b->rootBox, &b->meshBox, &b->refitScratch};
    for (auto *p : f32)
        p->release();
    DevBuf<uint32_t> *u32[] = {&b->T,    &b->meshOf,     &b->target,   &b->triBegin, &b->ids,
                               &b->leafParent, &b->work, &b->meshRoot, &b->hist,     &b->words,
                               &b->gdWork, &b->rayList, &b->gridCells, &b->gridTemp};
    for (auto *p : u32)
        p->release();
    DevBuf<uint64_t> *u64[] = {&b->keys, &b->keysAlt, &b->codes, &b->pairs, &b->pairsAlt};
    for (auto *p : u64)
        p->release();
    DevBuf<uint8_t> *u8[] = {&b->status, &b->flags, &b->bits, &b->votes};
    for (auto *p : u8)
        p->release();
    b->meshBounds.release();
    b->waveBounds.release();
    b->nodes.release();
    b->front[0].release();
    b->front[1].release();
    b->spill.release();
    b->gridDesc.release();
    b->gridCode.release();
    b->gridTri.release();
    delete b;
    context_release(c);
    return SBG_OK;
}

}  // extern "C"
