// C-ABI implementation (include/sbgpu.h): contexts, meshes, pair lists and the
// host-side sequencing of the kernels in k_*.hip.
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "sbg_host.h"
#include "sbg_rayfilter.h"

namespace sbg {
// resolves this device's copy of the allocation-free scratch array (the
// device must be current); defined beside the array at the end of this file
int noalloc_attach(int device);
}  // namespace sbg

using namespace sbg;

namespace sbg {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int use(sbg_context ctx)
{
    if (!ctx)
        return fail(SBG_ERR_INVALID, "null context");
    if (ctx->destroyed)
        return fail(SBG_ERR_STATE, "context already destroyed");
    HIP_TRY(hipSetDevice(ctx->device));
    return SBG_OK;
}


void context_release(sbg_context ctx)
{
    if (--ctx->refs == 0)
        context_free(ctx);
}

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(SBG_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return SBG_OK;
}

int ceil_log2(size_t n)
{
    int b = 0;
    while (((size_t)1 << b) < n)
        ++b;
    return b > 0 ? b : 1;
}

}  // namespace sbg

struct sbg_mesh_s {
    sbg_context ctx = nullptr;
    size_t nv = 0, nt = 0;
    // sbg_mesh_prepare runs on this stream (meshes alternate between the
    // context's two, so the two meshes of a boolean prepare concurrently);
    // `used` marks the last work queued on the mesh, which everything that
    // reads or rewrites the mesh waits for
    hipStream_t stream = nullptr;
    hipEvent_t used = nullptr;
    hipEvent_t prepDone = nullptr;  // end of the last prepare (the grid stream's start)
    bool pendingOnAux = false;  // `used` was last recorded on the second stream and not yet waited for
    DevBuf<uint32_t> hist;  // the Morton sort's histograms (per mesh: prepares overlap)
    DevBuf<double> V;
    DevBuf<uint32_t> T;
    DevBuf<double> tri64;
    DevBuf<float> tri32, box32, blockBounds, bounds, rootBox, refitScratch;
    DevBuf<uint32_t> keys, vals, keysAlt, valsAlt, leafParent, flags;
    DevBuf<BNode> nodes;
    uint32_t root = 0;
    const uint32_t *sortedIds = nullptr;  // triangles in Morton order (vals or valsAlt)
    bool prepared = false;
    // per-axis ray grid, built on first use as a classification target
    DevBuf<uint32_t> gridCells, gridTemp;
    DevBuf<uint2> gridCode;
    DevBuf<uint32_t> gridTri;
    size_t gridCap = 0;
    uint32_t gridG = 0;
    int gridDensity = 0;
    bool gridReady = false;
    // the stream of the last grid build (null: none since the last prepare);
    // gridDone follows that build, and work on any other stream that reads the
    // grid or rewrites its inputs waits for it
    hipStream_t gridStream = nullptr;
    hipEvent_t gridDone = nullptr;
};

struct sbg_pairs_s {
    sbg_context ctx = nullptr;
    size_t n = 0;
    DevBuf<uint64_t> pairs, alt;
    DevBuf<uint8_t> status;
    DevBuf<double> seg, hitSeg;
    DevBuf<uint64_t> hitPairs;
    DevBuf<uint32_t> blockCounts, nHits, work;
    bool tested = false;
    // largest first / second index in the list (host-built lists are checked
    // against the meshes before any kernel indexes them)
    uint64_t maxA = 0, maxB = 0;
};

// Work queued on the context stream for up to two meshes: waits for what was
// queued on them elsewhere (a prepare on the other stream), and on leaving
// marks the context stream's work as their latest use, which a later prepare
// waits for.
struct MeshUse {
    sbg_mesh m[2];
    explicit MeshUse(sbg_mesh a, sbg_mesh b = nullptr) : m{a, b}
    {
        for (sbg_mesh x : m)
            if (x && x->pendingOnAux) {
                (void)hipStreamWaitEvent(x->ctx->stream, x->used, 0);
                x->pendingOnAux = false;
            }
    }
    ~MeshUse()
    {
        // only a mesh that prepares on the second stream needs the mark: one
        // that prepares on the context stream is ordered after this work anyway
        for (sbg_mesh x : m)
            if (x && x->stream != x->ctx->stream)
                (void)hipEventRecord(x->used, x->ctx->stream);
    }
};

// waits for the second stream when leaving (work queued there must not
// outlive the call, whatever path it returns by)
struct AuxJoin {
    sbg_context c;
    ~AuxJoin() { (void)hipStreamSynchronize(c->aux); }
};

extern "C" {

const char *sbg_last_error(void) { return g_err.c_str(); }

int sbg_device_count(int *count)
{
    if (!count)
        return fail(SBG_ERR_INVALID, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return fail(SBG_ERR_NO_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    *count = n;
    return SBG_OK;
}

int sbg_context_create(int device, sbg_context *out)
{
    if (!out)
        return fail(SBG_ERR_INVALID, "null output");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(SBG_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n)
        return fail(SBG_ERR_INVALID, "device %d out of range (%d devices)", device, n);
    HIP_TRY(hipSetDevice(device));
    TRY(noalloc_attach(device));
    sbg_context c = new sbg_context_s;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(SBG_ERR_HIP, "hipStreamCreateWithFlags failed");
    }
    if (hipHostMalloc((void **)&c->pinned, sizeof(uint32_t) * kPinnedWords, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(SBG_ERR_OOM, "hipHostMalloc failed");
    }
    if (hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->grids, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->auxReady, hipEventDisableTiming) != hipSuccess) {
        if (c->grids)
            (void)hipStreamDestroy(c->grids);
        if (c->aux)
            (void)hipStreamDestroy(c->aux);
        (void)hipHostFree(c->pinned);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(SBG_ERR_HIP, "second or third stream creation failed");
    }
    *out = c;
    return SBG_OK;
}

}  // extern "C"

namespace sbg {
constexpr size_t kSpareMeshes = 2;

static void mesh_free_buffers(sbg_mesh m)
{
    m->hist.release();
    m->V.release();
    m->T.release();
    m->tri64.release();
    m->tri32.release();
    m->box32.release();
    m->blockBounds.release();
    m->bounds.release();
    m->rootBox.release();
    m->refitScratch.release();
    m->keys.release();
    m->vals.release();
    m->keysAlt.release();
    m->valsAlt.release();
    m->leafParent.release();
    m->flags.release();
    m->nodes.release();
    m->gridCells.release();
    m->gridTemp.release();
    m->gridCode.release();
    m->gridTri.release();
}

static void mesh_free_events(sbg_mesh m)
{
    for (hipEvent_t *e : {&m->used, &m->prepDone, &m->gridDone})
        if (*e) {
            (void)hipEventDestroy(*e);
            *e = nullptr;
        }
}

// a new mesh, on the buffers of a destroyed one when there is one (a caller
// that makes and drops meshes per boolean then allocates nothing)
static sbg_mesh mesh_new(sbg_context c)
{
    sbg_mesh m = nullptr;
    if (!c->spareMeshes.empty()) {
        // work queued on the old mesh may still read these buffers, and the
        // new mesh's uploads do not go through the streams
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->aux);
        (void)hipStreamSynchronize(c->grids);
        m = c->spareMeshes.back();
        c->spareMeshes.pop_back();
    } else {
        m = new sbg_mesh_s;
    }
    if ((!m->used && hipEventCreateWithFlags(&m->used, hipEventDisableTiming) != hipSuccess) ||
        (!m->prepDone && hipEventCreateWithFlags(&m->prepDone, hipEventDisableTiming) != hipSuccess) ||
        (!m->gridDone && hipEventCreateWithFlags(&m->gridDone, hipEventDisableTiming) != hipSuccess)) {
        mesh_free_events(m);
        delete m;
        return nullptr;
    }
    m->stream = (c->meshCount++ & 1) ? c->aux : c->stream;
    (void)hipEventRecord(m->used, c->stream);
    m->pendingOnAux = false;
    m->ctx = c;
    m->nv = m->nt = 0;
    m->root = 0;
    m->sortedIds = nullptr;
    m->prepared = false;
    m->gridCap = 0;
    m->gridG = 0;
    m->gridDensity = 0;
    m->gridReady = false;
    m->gridStream = nullptr;
    ++c->refs;
    return m;
}

static void pairs_free_buffers(sbg_pairs p)
{
    p->pairs.release();
    p->alt.release();
    p->status.release();
    p->seg.release();
    p->hitSeg.release();
    p->hitPairs.release();
    p->blockCounts.release();
    p->nHits.release();
    p->work.release();
}

// a new pair list, on the buffers of the last destroyed one when there is one
// (a step that makes and drops one list per call then allocates nothing)
static sbg_pairs pairs_new(sbg_context c, bool hostUpload)
{
    sbg_pairs p = c->sparePairs;
    c->sparePairs = nullptr;
    // a list filled by kernels on the context stream follows every earlier
    // user of these buffers there (the second stream is joined before any
    // call returns); a host upload bypasses the streams and must wait
    if (!p) {
        p = new sbg_pairs_s;
    } else if (hostUpload) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->aux);
    }
    p->ctx = c;
    p->n = 0;
    p->tested = false;
    p->maxA = p->maxB = 0;
    ++c->refs;
    return p;
}

void context_free(sbg_context ctx)
{
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    ctx->front[0].release();
    ctx->front[1].release();
    ctx->spill.release();
    ctx->counters.release();
    ctx->hist.release();
    ctx->stats.release();
    ctx->pts.release();
    ctx->flags.release();
    ctx->bits.release();
    ctx->votes.release();
    ctx->ovf.release();
    ctx->rayList.release();
    ctx->pts2.release();
    ctx->flags2.release();
    ctx->bits2.release();
    ctx->votes2.release();
    ctx->ovf2.release();
    ctx->rayList2.release();
    ctx->votesHost.release();
    ctx->votesHost2.release();
    ctx->workCounts.release();
    ctx->ccKeys.release();
    ctx->ccOwners.release();
    ctx->ccParent.release();
    ctx->ccLabels.release();
    ctx->ccBlocked.release();
    ctx->rawA.release();
    ctx->rawB.release();
    ctx->rawSeg.release();
    ctx->rawStatus.release();
    ctx->adjXyz.release();
    ctx->adjAll.release();
    ctx->adjGradD.release();
    ctx->adjGall.release();
    ctx->adjVol.release();
    ctx->adjSrc.release();
    ctx->adjFaces.release();
    if (ctx->sparePairs) {
        pairs_free_buffers(ctx->sparePairs);
        delete ctx->sparePairs;
        ctx->sparePairs = nullptr;
    }
    (void)hipStreamSynchronize(ctx->aux);
    (void)hipStreamSynchronize(ctx->grids);
    for (sbg_mesh m : ctx->spareMeshes) {
        mesh_free_buffers(m);
        mesh_free_events(m);
        delete m;
    }
    ctx->spareMeshes.clear();
    for (hipStream_t u : ctx->upload)
        if (u) {
            (void)hipStreamSynchronize(u);
            (void)hipStreamDestroy(u);
        }
    for (auto &slots : ctx->stageDone)
        for (hipEvent_t e : slots)
            if (e)
                (void)hipEventDestroy(e);
    ctx->staging.release();
    (void)hipHostFree(ctx->pinned);
    (void)hipEventDestroy(ctx->auxReady);
    (void)hipStreamDestroy(ctx->grids);
    (void)hipStreamDestroy(ctx->aux);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}
}  // namespace sbg

namespace sbg {

// The upload streams, the staging slots' events and the page-locked staging,
// made on the context's first mesh upload.
static int upload_staging(sbg_context c)
{
    for (hipStream_t &u : c->upload)
        if (!u && hipStreamCreateWithFlags(&u, hipStreamNonBlocking) != hipSuccess) {
            u = nullptr;
            return fail(SBG_ERR_HIP, "upload stream creation failed");
        }
    for (auto &slots : c->stageDone)
        for (hipEvent_t &e : slots)
            if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                e = nullptr;
                return fail(SBG_ERR_HIP, "upload event creation failed");
            }
    return c->staging.ensure(2 * kUploadThreads * kStageSlot);
}

// One upload thread's share, src[first, last) into dst[first, last): the
// thread converts a slot's worth of elements into one of its two staging
// slots while the DMA of the other runs on its upload stream, and keeps the
// largest element of an integer source (the triangles' vertex indices).
template <typename S, typename D>
static void upload_share(sbg_context c, int t, D *dst, const S *src, size_t first, size_t last, int *status,
                         uint64_t *top)
{
    *status = SBG_OK;
    *top = 0;
    if (hipSetDevice(c->device) != hipSuccess) {
        *status = SBG_ERR_HIP;
        return;
    }
    const size_t per = kStageSlot / sizeof(D);
    S hi = 0;
    int k = 0;
    for (size_t i = first; i < last; i += per, k ^= 1) {
        const size_t n = std::min(per, last - i);
        D *slot = (D *)(c->staging.p + (size_t)(2 * t + k) * kStageSlot);
        if (hipEventSynchronize(c->stageDone[t][k]) != hipSuccess) {  // its last DMA is done
            *status = SBG_ERR_HIP;
            return;
        }
        if constexpr (std::is_integral<S>::value) {
            for (size_t j = 0; j < n; ++j) {
                const S v = src[i + j];
                slot[j] = (D)v;
                hi = v > hi ? v : hi;
            }
        } else {
            std::memcpy(slot, src + i, n * sizeof(D));
        }
        const hipStream_t u = c->upload[t % kUploadStreams];
        if (hipMemcpyAsync(dst + i, slot, n * sizeof(D), hipMemcpyHostToDevice, u) != hipSuccess ||
            hipEventRecord(c->stageDone[t][k], u) != hipSuccess) {
            *status = SBG_ERR_HIP;
            return;
        }
    }
    *top = (uint64_t)hi;
}

// The mesh's two host-to-device copies, vertices and triangles (narrowed from
// 64-bit indices when S is), through the page-locked staging: up to
// kUploadThreads host threads (one per kUploadShare bytes), each taking one
// contiguous share of both arrays, the calling thread the first.  Returns
// when every copy has landed, with the largest vertex index in *top.
template <typename S>
static int upload_mesh(sbg_context c, sbg_mesh m, const double *xyz, const S *tris, uint64_t *top)
{
    *top = 0;
    const size_t nx = 3 * m->nv, ni = 3 * m->nt;
    if (nx + ni == 0)
        return SBG_OK;
    TRY(upload_staging(c));
    const size_t bytes = sizeof(double) * nx + sizeof(uint32_t) * ni;
    const int shares = (int)std::max<size_t>(1, std::min<size_t>(kUploadThreads, bytes / kUploadShare));
    int status[kUploadThreads][2];
    uint64_t tops[kUploadThreads][2];
    auto share = [&](int t) {
        const size_t sx = (nx + shares - 1) / shares, si = (ni + shares - 1) / shares;
        const size_t x0 = std::min(nx, t * sx), i0 = std::min(ni, t * si);
        upload_share<double, double>(c, t, m->V.p, xyz, x0, std::min(nx, x0 + sx), &status[t][0], &tops[t][0]);
        upload_share<S, uint32_t>(c, t, m->T.p, tris, i0, std::min(ni, i0 + si), &status[t][1], &tops[t][1]);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < shares; ++t) {
        try {
            pool.emplace_back(share, t);
        } catch (...) {  // no thread to be had: this share on the calling thread
            share(t);
        }
    }
    share(0);
    for (std::thread &th : pool)
        th.join();
    hipError_t e = hipSuccess;
    for (hipStream_t u : c->upload) {
        const hipError_t eu = hipStreamSynchronize(u);
        e = e != hipSuccess ? e : eu;
    }
    for (int t = 0; t < shares; ++t) {
        if (status[t][0] != SBG_OK || status[t][1] != SBG_OK || e != hipSuccess)
            return fail(SBG_ERR_HIP, "mesh upload failed: %s", hipGetErrorString(e != hipSuccess ? e : hipGetLastError()));
        *top = std::max(*top, tops[t][1]);
    }
    return SBG_OK;
}

// sbg_mesh_create and sbg_mesh_create64: the indices are checked against
// nverts (and against 32 bits, which the device stores) after the upload,
// whose threads find their largest value on the way.
template <typename S>
static int mesh_create(sbg_context ctx, const double *xyz, size_t nverts, const S *tris, size_t ntris, sbg_mesh *out)
{
    if (!out)
        return fail(SBG_ERR_INVALID, "null output");
    *out = nullptr;
    TRY(use(ctx));
    if ((nverts && !xyz) || (ntris && !tris))
        return fail(SBG_ERR_INVALID, "null vertex or triangle array");
    if (ntris >= (size_t)kLeafBit)
        return fail(SBG_ERR_INVALID, "too many triangles (%zu)", ntris);
    sbg_mesh m = mesh_new(ctx);
    if (!m)
        return fail(SBG_ERR_HIP, "event creation failed");
    m->nv = nverts;
    m->nt = ntris;
    int r = m->V.ensure(3 * nverts);
    if (r == SBG_OK)
        r = m->T.ensure(3 * ntris);
    uint64_t top = 0;
    if (r == SBG_OK)
        r = upload_mesh(ctx, m, xyz, tris, &top);
    if (r == SBG_OK && ntris && (top >= nverts || top > 0xffffffffull)) {
        // the first offending index for the message
        r = fail(SBG_ERR_INVALID, "vertex index %llu beyond 32 bits", (unsigned long long)top);
        for (size_t i = 0; i < 3 * ntris; ++i)
            if ((uint64_t)tris[i] >= nverts) {
                r = fail(SBG_ERR_INVALID, "triangle %zu references vertex %llu of %zu", i / 3,
                         (unsigned long long)tris[i], nverts);
                break;
            }
    }
    if (r != SBG_OK) {
        sbg_mesh_destroy(m);
        return r;
    }
    *out = m;
    return SBG_OK;
}

}  // namespace sbg

extern "C" {

int sbg_context_destroy(sbg_context ctx)
{
    if (!ctx || ctx->destroyed)
        return SBG_OK;
    ctx->destroyed = true;
    context_release(ctx);
    return SBG_OK;
}

int sbg_context_synchronize(sbg_context ctx)
{
    TRY(use(ctx));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SBG_OK;
}

int sbg_context_stream(sbg_context ctx, void **hip_stream)
{
    if (!ctx || !hip_stream)
        return fail(SBG_ERR_INVALID, "null argument");
    *hip_stream = (void *)ctx->stream;
    return SBG_OK;
}

int sbg_context_set_timing(sbg_context ctx, int enable)
{
    TRY(use(ctx));
    ctx->timing = enable != 0;
    return SBG_OK;
}

int sbg_context_stage_ms(sbg_context ctx, double *ms, int reset)
{
    TRY(use(ctx));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (auto &e : ctx->events) {
        float t = 0.0f;
        HIP_TRY(hipEventElapsedTime(&t, e.a, e.b));
        ctx->acc[e.stage] += t;
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    ctx->events.clear();
    if (ms)
        for (int i = 0; i < SBG_STAGE_COUNT; ++i)
            ms[i] = ctx->acc[i];
    if (reset)
        for (int i = 0; i < SBG_STAGE_COUNT; ++i)
            ctx->acc[i] = 0.0;
    return SBG_OK;
}

// ---------------------------------------------------------------------------
// meshes
// ---------------------------------------------------------------------------
int sbg_mesh_create(sbg_context ctx, const double *xyz, size_t nverts, const uint32_t *tris, size_t ntris,
                    sbg_mesh *out)
{
    return mesh_create(ctx, xyz, nverts, tris, ntris, out);
}

int sbg_mesh_create64(sbg_context ctx, const double *xyz, size_t nverts, const uint64_t *tris, size_t ntris,
                      sbg_mesh *out)
{
    return mesh_create(ctx, xyz, nverts, tris, ntris, out);
}

int sbg_host_alloc(size_t bytes, void **out)
{
    if (!out)
        return fail(SBG_ERR_INVALID, "null output");
    *out = nullptr;
    if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        (void)hipGetLastError();
        return fail(SBG_ERR_OOM, "hipHostMalloc of %zu bytes failed", bytes);
    }
    return SBG_OK;
}

int sbg_host_free(void *p)
{
    if (p && hipHostFree(p) != hipSuccess)
        return fail(SBG_ERR_INVALID, "not a block from sbg_host_alloc");
    return SBG_OK;
}

}  // extern "C"

static int ensure_grid(sbg_mesh m, hipStream_t stream);

// Queues the ray grid of every pending mesh except `keep` on the grid stream,
// each after its last prepare.  A grid already built since (a classification
// asked for it) is not built again.
static int queue_pending_grids(sbg_context c, sbg_mesh keep)
{
    std::vector<sbg_mesh> rest;
    for (sbg_mesh m : c->gridsPending) {
        if (m == keep) {
            rest.push_back(m);
            continue;
        }
        if (!m->prepared || (m->gridReady && m->gridDensity == c->tune.grid_density))
            continue;
        HIP_TRY(hipStreamWaitEvent(c->grids, m->prepDone, 0));
        TRY(ensure_grid(m, c->grids));
    }
    c->gridsPending.swap(rest);
    return SBG_OK;
}

static void drop_pending_grid(sbg_mesh m)
{
    auto &v = m->ctx->gridsPending;
    v.erase(std::remove(v.begin(), v.end(), m), v.end());
}


extern "C" {

int sbg_mesh_destroy(sbg_mesh m)
{
    if (!m)
        return SBG_OK;
    sbg_context c = m->ctx;
    (void)hipSetDevice(c->device);
    drop_pending_grid(m);
    if (c->spareMeshes.size() < kSpareMeshes && !c->destroyed) {
        c->spareMeshes.push_back(m);  // keeps its buffers for the next mesh
    } else {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->aux);
        (void)hipStreamSynchronize(c->grids);
        mesh_free_buffers(m);
        mesh_free_events(m);
        delete m;
    }
    context_release(c);
    return SBG_OK;
}

int sbg_mesh_counts(sbg_mesh m, size_t *nverts, size_t *ntris)
{
    if (!m)
        return fail(SBG_ERR_INVALID, "null mesh");
    if (nverts)
        *nverts = m->nv;
    if (ntris)
        *ntris = m->nt;
    return SBG_OK;
}

int sbg_mesh_prepare(sbg_mesh m)
{
    if (!m)
        return fail(SBG_ERR_INVALID, "null mesh");
    sbg_context c = m->ctx;
    TRY(use(c));
    const size_t nt = m->nt;
    const int nb = records_blocks((uint32_t)nt);
    TRY(m->tri64.ensure((size_t)kTri64Stride * nt));
    TRY(m->tri32.ensure((size_t)kTri32Stride * nt));
    TRY(m->box32.ensure((size_t)kBox32Stride * nt));
    TRY(m->blockBounds.ensure(6 * (size_t)(nb > 0 ? nb : 1)));
    TRY(m->bounds.ensure(6));
    TRY(m->rootBox.ensure(6));
    TRY(m->keys.ensure(nt));
    TRY(m->vals.ensure(nt));
    TRY(m->keysAlt.ensure(nt));
    TRY(m->valsAlt.ensure(nt));
    TRY(m->leafParent.ensure(nt));
    TRY(m->nodes.ensure(nt > 1 ? nt - 1 : 1));
    TRY(m->flags.ensure(2 * nt + 1));
    TRY(m->refitScratch.ensure(refit_scratch_floats((uint32_t)nt)));
    TRY(m->hist.ensure(sort_hist_entries(nt)));
    const hipStream_t s = m->stream;
    const bool onAux = s != c->stream;
    if (onAux)
        HIP_TRY(hipStreamWaitEvent(s, m->used, 0));  // earlier context-stream work on this mesh
    if (m->gridStream && m->gridStream != s)
        HIP_TRY(hipStreamWaitEvent(s, m->gridDone, 0));  // the last grid build still reads the records
    m->gridStream = nullptr;
    m->prepared = false;
    m->gridReady = false;
    drop_pending_grid(m);
    if (nt == 0) {
        m->prepared = true;
        return SBG_OK;
    }
    {
        Stage st(c, SBG_STAGE_RECORDS, s);
        launch_records(m->V.p, m->T.p, (uint32_t)nt, m->tri64.p, m->tri32.p, m->box32.p,
                       m->blockBounds.p, nb, s);
        launch_bounds_reduce(m->blockBounds.p, nb, m->bounds.p, s);
        launch_morton(m->box32.p, m->bounds.p, (uint32_t)nt, m->keys.p, m->vals.p, m->flags.p,
                      nt > 1 ? hierarchy_zeroed_u32((uint32_t)nt) : 0u, s);
        TRY(check_launch("records"));
    }
    uint32_t *sk = m->keys.p, *sv = m->vals.p;
    {
        Stage st(c, SBG_STAGE_SORT, s);
        if (radix_sort_u32(m->keys.p, m->vals.p, m->keysAlt.p, m->valsAlt.p, nt, 0, kMortonBits, m->hist.p, s)) {
            sk = m->keysAlt.p;
            sv = m->valsAlt.p;
        }
        TRY(check_launch("morton sort"));
    }
    m->sortedIds = sv;
    {
        Stage st(c, SBG_STAGE_BVH, s);
        if (nt == 1) {
            m->root = kLeafBit;
            HIP_TRY(hipMemcpyAsync(m->rootBox.p, m->box32.p, 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
            HIP_TRY(hipMemcpyAsync(m->rootBox.p + 3, m->box32.p + 4, 3 * sizeof(float), hipMemcpyDeviceToDevice,
                                   s));
        } else {
            m->root = 0;
            launch_hierarchy(sk, sv, (uint32_t)nt, m->nodes.p, m->leafParent.p, m->flags.p, s);
            launch_refit(sv, m->box32.p, (uint32_t)nt, m->nodes.p, m->leafParent.p, m->flags.p, m->refitScratch.p,
                         m->rootBox.p, s);
        }
        TRY(check_launch("bvh"));
    }
    if (onAux) {
        HIP_TRY(hipEventRecord(m->used, s));
        m->pendingOnAux = true;
    }
    m->prepared = true;
    if (c->tune.early_grids == 2) {
        // the ray grid goes on the third stream after this prepare, so the
        // pair search waits for the prepare only and the grid builds beside
        // it; it is queued after the next prepare of another mesh or at the
        // next search, so that the host queues the two prepares first
        HIP_TRY(hipEventRecord(m->prepDone, s));
        TRY(queue_pending_grids(c, m));
        c->gridsPending.push_back(m);
    }
    return SBG_OK;
}

int sbg_mesh_normals(sbg_mesh m, double *out)
{
    if (!m || !out)
        return fail(SBG_ERR_INVALID, "null argument");
    if (!m->prepared)
        return fail(SBG_ERR_STATE, "mesh not prepared");
    TRY(use(m->ctx));
    MeshUse mu(m);
    if (m->nt) {
        sbg_context c = m->ctx;
        TRY(c->pts.ensure(3 * m->nt));
        launch_normals64(m->tri64.p, (uint32_t)m->nt, c->pts.p, c->stream);
        TRY(check_launch("normals64"));
        HIP_TRY(hipMemcpyAsync(out, c->pts.p, sizeof(double) * 3 * m->nt, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return SBG_OK;
}

int sbg_mesh_boxes(sbg_mesh m, double *out)
{
    if (!m || !out)
        return fail(SBG_ERR_INVALID, "null argument");
    if (!m->prepared)
        return fail(SBG_ERR_STATE, "mesh not prepared");
    sbg_context c = m->ctx;
    TRY(use(c));
    MeshUse mu(m);
    if (!m->nt)
        return SBG_OK;
    TRY(c->pts.ensure(6 * m->nt));
    launch_boxes64(m->tri64.p, (uint32_t)m->nt, c->pts.p, c->stream);
    TRY(check_launch("boxes64"));
    HIP_TRY(hipMemcpyAsync(out, c->pts.p, sizeof(double) * 6 * m->nt, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SBG_OK;
}

int sbg_components(sbg_mesh m, const uint8_t *blocked, uint32_t *labels)
{
    if (!m || !labels)
        return fail(SBG_ERR_INVALID, "null argument");
    sbg_context c = m->ctx;
    TRY(use(c));
    MeshUse mu(m);
    const size_t nt = m->nt;
    if (!nt)
        return SBG_OK;
    const size_t slots = component_table_slots((uint32_t)nt);
    TRY(c->ccKeys.ensure(slots));
    TRY(c->ccOwners.ensure(slots + 1));
    TRY(c->ccParent.ensure(nt));
    TRY(c->ccLabels.ensure(nt));
    if (blocked) {
        TRY(c->ccBlocked.ensure(nt));
        HIP_TRY(hipMemcpyAsync(c->ccBlocked.p, blocked, nt, hipMemcpyHostToDevice, c->stream));
    }
    launch_components(m->T.p, (uint32_t)nt, blocked ? c->ccBlocked.p : nullptr, c->ccKeys.p, c->ccOwners.p,
                      c->ccParent.p, c->ccLabels.p, c->stream);
    TRY(check_launch("components"));
    uint32_t *repeated = c->pinned + kComponentsPinned;
    HIP_TRY(hipMemcpyAsync(repeated, c->ccOwners.p + slots, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(labels, c->ccLabels.p, sizeof(uint32_t) * nt, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (*repeated)
        return fail(SBG_ERR_INVALID, "repeated half-edge: two unblocked triangles share a directed edge");
    return SBG_OK;
}

int sbg_half_edges(sbg_context c, const uint32_t *tris, size_t ntris, const uint64_t *queries, size_t nqueries,
                   uint32_t *opposite, uint32_t *query_owner, uint32_t *first_repeat)
{
    if ((ntris && (!tris || !opposite)) || (nqueries && (!queries || !query_owner)) || !first_repeat)
        return fail(SBG_ERR_INVALID, "null argument");
    if (ntris >= 0x50000000u || nqueries >= 0xffffffffu)
        return fail(SBG_ERR_INVALID, "too many triangles or queries");
    TRY(use(c));
    for (size_t i = 0; i < 3 * ntris; ++i)
        if (tris[i] == 0xffffffffu)
            return fail(SBG_ERR_INVALID, "vertex index 0xffffffff");
    const size_t slots = component_table_slots((uint32_t)ntris);
    TRY(c->ccKeys.ensure(slots));
    TRY(c->ccOwners.ensure(slots + 1));
    TRY(c->heTris.ensure(3 * ntris));
    TRY(c->heOpposite.ensure(3 * ntris));
    TRY(c->heQueries.ensure(nqueries));
    TRY(c->heQueryOwner.ensure(nqueries));
    TRY(c->heRepeat.ensure(1));
    if (ntris)
        HIP_TRY(hipMemcpyAsync(c->heTris.p, tris, sizeof(uint32_t) * 3 * ntris, hipMemcpyHostToDevice, c->stream));
    if (nqueries)
        HIP_TRY(hipMemcpyAsync(c->heQueries.p, queries, sizeof(uint64_t) * nqueries, hipMemcpyHostToDevice,
                               c->stream));
    launch_half_edges(c->heTris.p, (uint32_t)ntris, c->heQueries.p, (uint32_t)nqueries, c->ccKeys.p, c->ccOwners.p,
                      c->heOpposite.p, c->heQueryOwner.p, c->heRepeat.p, c->stream);
    TRY(check_launch("half edges"));
    if (ntris)
        HIP_TRY(hipMemcpyAsync(opposite, c->heOpposite.p, sizeof(uint32_t) * 3 * ntris, hipMemcpyDeviceToHost,
                               c->stream));
    if (nqueries)
        HIP_TRY(hipMemcpyAsync(query_owner, c->heQueryOwner.p, sizeof(uint32_t) * nqueries, hipMemcpyDeviceToHost,
                               c->stream));
    HIP_TRY(hipMemcpyAsync(first_repeat, c->heRepeat.p, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SBG_OK;
}

int sbg_mesh_bvh(sbg_mesh m, void *nodes_out, uint32_t *root_ref, float *root_box6)
{
    if (!m)
        return fail(SBG_ERR_INVALID, "null mesh");
    if (!m->prepared)
        return fail(SBG_ERR_STATE, "mesh not prepared");
    sbg_context c = m->ctx;
    TRY(use(c));
    MeshUse mu(m);
    if (root_ref)
        *root_ref = m->root;
    if (nodes_out && m->nt > 1)
        HIP_TRY(hipMemcpyAsync(nodes_out, m->nodes.p, sizeof(BNode) * (m->nt - 1), hipMemcpyDeviceToHost, c->stream));
    if (root_box6 && m->nt > 0)
        HIP_TRY(hipMemcpyAsync(root_box6, m->rootBox.p, sizeof(float) * 6, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SBG_OK;
}

// ---------------------------------------------------------------------------
// candidate pairs
// ---------------------------------------------------------------------------
int sbg_pairs_destroy(sbg_pairs p)
{
    if (!p)
        return SBG_OK;
    sbg_context c = p->ctx;
    (void)hipSetDevice(c->device);
    if (!c->sparePairs && !c->destroyed) {
        c->sparePairs = p;  // keeps its buffers for the next list
    } else {
        (void)hipStreamSynchronize(c->stream);
        pairs_free_buffers(p);
        delete p;
    }
    context_release(c);
    return SBG_OK;
}

static int queue_grids(sbg_mesh a, sbg_mesh b);

int sbg_find_pairs(sbg_mesh a, sbg_mesh b, sbg_pairs *out)
{
    if (!out || !a || !b)
        return fail(SBG_ERR_INVALID, "null argument");
    *out = nullptr;
    if (a->ctx != b->ctx)
        return fail(SBG_ERR_INVALID, "meshes belong to different contexts");
    if (!a->prepared || !b->prepared)
        return fail(SBG_ERR_STATE, "mesh not prepared");
    sbg_context c = a->ctx;
    TRY(use(c));
    TRY(queue_pending_grids(c, nullptr));
    MeshUse mu(a, b);
    const bool earlyGrids = c->tune.early_grids && a->nt && b->nt;
    if (earlyGrids) {
        // the second stream after everything queued so far (the prepares)
        HIP_TRY(hipEventRecord(c->auxReady, c->stream));
        HIP_TRY(hipStreamWaitEvent(c->aux, c->auxReady, 0));
    }
    sbg_pairs p = pairs_new(c, false);
    if (a->nt == 0 || b->nt == 0) {
        *out = p;
        return SBG_OK;
    }
    const size_t na = a->nt, nb = b->nt;
    if (!c->frontCap)
        c->frontCap = std::max<size_t>((size_t)1 << 18, 2 * (na + nb));
    if (!c->candCap)
        c->candCap = std::max<size_t>((size_t)1 << 16, (na + nb) / 4);
    TraverseArgs t = {a->nodes.p, a->box32.p, a->tri64.p, b->nodes.p, b->box32.p, b->tri64.p};
    uint32_t nCand = 0;
    int rc = SBG_OK;
    bool found = false;
    size_t preSorted = 0;  // slots of the sort queued before the count was read
    bool preInAlt = false;
    for (int attempt = 0; attempt < 16; ++attempt) {
        if ((rc = c->front[0].ensure(c->frontCap)) != SBG_OK || (rc = c->front[1].ensure(c->frontCap)) != SBG_OK ||
            (rc = c->counters.ensure(kMaxLevels + 8)) != SBG_OK || (rc = p->pairs.ensure(c->candCap)) != SBG_OK) {
            sbg_pairs_destroy(p);
            return rc;
        }
        uint32_t *cnt = c->counters.p;                 // [0, kMaxLevels]: frontier sizes
        uint32_t *candCount = cnt + kMaxLevels + 1;    // candidates
        uint32_t *overflow = cnt + kMaxLevels + 2;     // overflow flags
        uint32_t *head = cnt + kMaxLevels + 3;         // depth-first seed claims
        uint32_t *topLevel = cnt + kMaxLevels + 4;     // first level past the one-workgroup top
        uint32_t *runFlag = cnt + kMaxLevels + 5;      // a first triangle's run too long for the quick sort
        const uint32_t fcap = (uint32_t)std::min<size_t>(c->frontCap, 0xffffffffu);
        const uint32_t ccap = (uint32_t)std::min<size_t>(c->candCap, 0xffffffffu);
        const bool dfs = c->depthFirst && c->tune.depth_first;
        if (dfs && (rc = c->spill.ensure((size_t)c->tune.dfs_waves * c->spillPerWave)) != SBG_OK) {
            sbg_pairs_destroy(p);
            return rc;
        }
        int level = 0;
        bool done = false, ovf = false;
        {
            Stage st(c, SBG_STAGE_TRAVERSE);
            launch_fill_u32(cnt, 0u, kMaxLevels + 8, c->stream);
            launch_seed(t, a->root, b->root, a->rootBox.p, b->rootBox.p, c->front[0].p, cnt, p->pairs.p, candCount,
                        c->stream);
            if (dfs) {
                // up to bfs_levels breadth-first levels with no host round
                // trip (the small top levels in one workgroup), then one
                // persistent launch over the frontier, then one sync
                launch_search_levels(t, c->front[0].p, c->front[1].p, cnt, fcap, kTopMax, c->tune.bfs_levels,
                                     kExtraLevels, topLevel, head, c->spill.p, c->spillPerWave, c->tune.dfs_waves,
                                     p->pairs.p, candCount, ccap, overflow, kExpandGrid, c->stream);
                rc = check_launch("depth-first search");
                done = true;
            }
            while (!dfs && !done && level < kMaxLevels) {
                for (int k = 0; k < kLevelsPerSync && level < kMaxLevels; ++k, ++level)
                    launch_expand(t, c->front[level & 1].p, cnt + level, c->front[(level + 1) & 1].p, cnt + level + 1,
                                  fcap, p->pairs.p, candCount, ccap, overflow, kExpandGrid, c->stream);
                if ((rc = check_launch("expand")) != SBG_OK)
                    break;
                hipError_t e1 = hipMemcpyAsync(c->pinned, cnt + level, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
                hipError_t e2 = hipMemcpyAsync(c->pinned + 1, overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
                hipError_t e3 = hipStreamSynchronize(c->stream);
                if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
                    rc = fail(SBG_ERR_HIP, "traversal sync failed: %s", hipGetErrorString(e3));
                    break;
                }
                if (c->pinned[1]) {
                    ovf = true;
                    break;
                }
                if (c->pinned[0] == 0)
                    done = true;
            }
        }
        // queue the pair sort before the count comes back (launched over
        // sortHint slots, sorting as many as the count on the device says),
        // so the host's launches overlap the search
        preSorted = 0;
        preInAlt = false;
        if (rc == SBG_OK && dfs && c->sortHint > 1 && c->sortHint <= ccap) {
            const size_t ns = c->sortHint;
            // alt as large as pairs: the two swap, and neither may shrink
            if ((rc = p->alt.ensure(ccap)) == SBG_OK && (rc = c->hist.ensure(sort_hist_entries(ns))) == SBG_OK) {
                Stage st(c, SBG_STAGE_PAIR_SORT);
                // by the first triangle (radix passes over its bits only),
                // then each first triangle's short run by the second; a run
                // too long for that raises a flag read with the count, and
                // the list is then sorted the long way
                const int ranges[2] = {32, 32 + ceil_log2(na)};
                preInAlt = radix_sort_u64(p->pairs.p, p->alt.p, ns, ranges, 1, c->hist.p, c->stream, candCount) != 0;
                launch_sort_runs(preInAlt ? p->alt.p : p->pairs.p, ns, candCount, kSortRunMax, runFlag, c->stream);
                preSorted = ns;
                rc = check_launch("pair sort");
            }
        }
        // the ray grids go beside the depth-first search and the pair sort,
        // which leave most of the chip idle
        if (rc == SBG_OK && attempt == 0 && earlyGrids)
            rc = queue_grids(a, b);
        if (rc != SBG_OK) {
            sbg_pairs_destroy(p);
            return rc;
        }
        if (!done && !ovf) {
            sbg_pairs_destroy(p);
            return fail(SBG_ERR_OVERFLOW, "tree-vs-tree search exceeded %d levels", kMaxLevels);
        }
        // all counters, to size a retry or read the candidate count
        hipError_t e = hipMemcpyAsync(c->pinned, cnt, sizeof(uint32_t) * (kMaxLevels + 6), hipMemcpyDeviceToHost,
                                      c->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            sbg_pairs_destroy(p);
            return fail(SBG_ERR_HIP, "counter readback failed: %s", hipGetErrorString(e));
        }
        uint32_t maxFront = 0;
        for (int l = 0; l <= kMaxLevels; ++l)
            maxFront = std::max(maxFront, c->pinned[l]);
        nCand = c->pinned[kMaxLevels + 1];
        const uint32_t flags = c->pinned[kMaxLevels + 2];
        if (preSorted && c->pinned[kMaxLevels + 5]) {
            // a long run: the list is sorted again by both triangles below,
            // from the buffer that holds all of it (the quick sort's output
            // when it covered the whole count; otherwise `pairs`, where the
            // slots past the quick sort's are untouched and the ones before
            // hold a permutation of what it read)
            if (preInAlt && preSorted >= nCand)
                std::swap(p->pairs, p->alt);
            preSorted = 0;
            preInAlt = false;
        }
        if (!ovf && !flags) {
            found = true;
            break;
        }
        if (flags & 12u) {  // depth-first spill or iteration cap: larger regions, then breadth-first only
            if (c->spillPerWave < kSpillMax)
                c->spillPerWave *= 4;
            else
                c->depthFirst = false;
            if (!(flags & 1u))
                continue;
        }
        if (maxFront > fcap)
            c->frontCap = std::max<size_t>(2 * c->frontCap, (size_t)maxFront + maxFront / 2);
        if (nCand > ccap)
            c->candCap = std::max<size_t>(2 * c->candCap, (size_t)nCand + nCand / 4);
        if (maxFront <= fcap && nCand <= ccap) {  // overflow flagged at a deeper, unread level
            c->frontCap *= 2;
            c->candCap *= 2;
        }
    }
    if (!found) {
        sbg_pairs_destroy(p);
        return fail(SBG_ERR_OVERFLOW, "candidate search did not fit after repeated growth");
    }
    p->n = nCand;
    p->maxA = na - 1;
    p->maxB = nb - 1;
    c->sortHint = nCand > 1 ? nCand + nCand / 4 + 1024 : 0;
    if (nCand > 1 && preSorted >= nCand) {
        if (preInAlt)
            std::swap(p->pairs, p->alt);
    } else if (nCand > 1) {
        Stage st(c, SBG_STAGE_PAIR_SORT);
        int rc2 = p->alt.ensure(std::max<size_t>(nCand, p->pairs.cap));
        if (rc2 == SBG_OK)
            rc2 = c->hist.ensure(sort_hist_entries(nCand));
        if (rc2 != SBG_OK) {
            sbg_pairs_destroy(p);
            return rc2;
        }
        const int ranges[4] = {0, ceil_log2(nb), 32, 32 + ceil_log2(na)};
        if (radix_sort_u64(p->pairs.p, p->alt.p, nCand, ranges, 2, c->hist.p, c->stream))
            std::swap(p->pairs, p->alt);
        if ((rc2 = check_launch("pair sort")) != SBG_OK) {
            sbg_pairs_destroy(p);
            return rc2;
        }
    }
    *out = p;
    return SBG_OK;
}

int sbg_pairs_create(sbg_context ctx, const uint32_t *ab, size_t n, sbg_pairs *out)
{
    if (!out)
        return fail(SBG_ERR_INVALID, "null output");
    *out = nullptr;
    TRY(use(ctx));
    if (n && !ab)
        return fail(SBG_ERR_INVALID, "null pair array");
    sbg_pairs p = pairs_new(ctx, true);
    p->n = n;
    int rc = p->pairs.ensure(n);
    if (rc != SBG_OK) {
        sbg_pairs_destroy(p);
        return rc;
    }
    if (n) {
        std::vector<uint64_t> h(n);
        for (size_t i = 0; i < n; ++i) {
            h[i] = ((uint64_t)ab[2 * i] << 32) | ab[2 * i + 1];
            p->maxA = std::max<uint64_t>(p->maxA, ab[2 * i]);
            p->maxB = std::max<uint64_t>(p->maxB, ab[2 * i + 1]);
        }
        if (hipMemcpy(p->pairs.p, h.data(), sizeof(uint64_t) * n, hipMemcpyHostToDevice) != hipSuccess) {
            sbg_pairs_destroy(p);
            return fail(SBG_ERR_HIP, "pair upload failed");
        }
    }
    *out = p;
    return SBG_OK;
}

int sbg_pairs_count(sbg_pairs p, size_t *n)
{
    if (!p || !n)
        return fail(SBG_ERR_INVALID, "null argument");
    *n = p->n;
    return SBG_OK;
}

int sbg_pairs_copy(sbg_pairs p, uint32_t *ab)
{
    if (!p || (!ab && p->n))
        return fail(SBG_ERR_INVALID, "null argument");
    TRY(use(p->ctx));
    if (!p->n)
        return SBG_OK;
    std::vector<uint64_t> h(p->n);
    HIP_TRY(hipMemcpyAsync(h.data(), p->pairs.p, sizeof(uint64_t) * p->n, hipMemcpyDeviceToHost, p->ctx->stream));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    for (size_t i = 0; i < p->n; ++i) {
        ab[2 * i] = (uint32_t)(h[i] >> 32);
        ab[2 * i + 1] = (uint32_t)h[i];
    }
    return SBG_OK;
}

static int check_pairs_in_range(sbg_mesh a, sbg_mesh b, sbg_pairs p)
{
    if (p->n && (p->maxA >= a->nt || p->maxB >= b->nt))
        return fail(SBG_ERR_INVALID, "pair indices (max %llu, %llu) out of range for meshes of %zu and %zu triangles",
                    (unsigned long long)p->maxA, (unsigned long long)p->maxB, a->nt, b->nt);
    return SBG_OK;
}

static int run_tri_tri(sbg_mesh a, sbg_mesh b, sbg_pairs p, int precision, GdStats *dstats)
{
    sbg_context c = p->ctx;
    const uint32_t n = (uint32_t)p->n;
    TRY(p->status.ensure(p->n));
    TRY(p->seg.ensure(6 * p->n));
    TRY(p->hitPairs.ensure(p->n));
    TRY(p->hitSeg.ensure(6 * p->n));
    TRY(p->blockCounts.ensure((size_t)compact_blocks(n)));
    TRY(p->nHits.ensure(1));
    TRY(p->work.ensure(tritri_work_u32((uint32_t)p->n)));
    {
        Stage st(c, SBG_STAGE_TRI_TRI);
        launch_tritri(a->tri64.p, b->tri64.p, a->tri32.p, b->tri32.p, p->pairs.p, n, nullptr, precision, p->status.p,
                      p->seg.p, p->work.p, dstats, c->stream);
        TRY(check_launch("tri_tri"));
    }
    {
        Stage st(c, SBG_STAGE_COMPACT);
        if (n == 0)  // otherwise the compaction's last block writes the count
            launch_fill_u32(p->nHits.p, 0u, 1, c->stream);
        launch_compact_hits(p->pairs.p, p->status.p, p->seg.p, n, p->blockCounts.p, p->hitPairs.p, p->hitSeg.p,
                            p->nHits.p, c->stream);
        TRY(check_launch("compact"));
    }
    return SBG_OK;
}

int sbg_tri_tri(sbg_mesh a, sbg_mesh b, sbg_pairs p, int precision, sbg_tri_tri_stats *stats)
{
    if (!a || !b || !p)
        return fail(SBG_ERR_INVALID, "null argument");
    if (a->ctx != b->ctx || a->ctx != p->ctx)
        return fail(SBG_ERR_INVALID, "handles belong to different contexts");
    if (!a->prepared || !b->prepared)
        return fail(SBG_ERR_STATE, "mesh not prepared");
    if (precision != SBG_PRECISION_FP64 && precision != SBG_PRECISION_FP32_FILTERED)
        return fail(SBG_ERR_INVALID, "unknown precision %d", precision);
    sbg_context c = p->ctx;
    TRY(use(c));
    MeshUse mu(a, b);
    TRY(check_pairs_in_range(a, b, p));
    GdStats *dstats = nullptr;
    if (stats) {
        TRY(c->stats.ensure(1));
        dstats = c->stats.p;
        launch_fill_u32((uint32_t *)dstats, 0u, sizeof(GdStats) / 4, c->stream);
    }
    TRY(run_tri_tri(a, b, p, precision, dstats));
    p->tested = true;
    if (stats) {
        GdStats h;
        HIP_TRY(hipMemcpyAsync(&h, dstats, sizeof(GdStats), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        stats->tested = p->n;
        stats->exit_plane1 = h.exits[0];
        stats->exit_plane2 = h.exits[1];
        stats->exit_interval = h.exits[2];
        stats->hits = h.exits[3];
        stats->coplanar = h.exits[4];
        stats->fallback_pairs = h.fallback;
        stats->filtered_pairs = h.filtered;
    }
    return SBG_OK;
}

int sbg_pairs_results(sbg_pairs p, uint8_t *status, double *seg)
{
    if (!p)
        return fail(SBG_ERR_INVALID, "null pairs");
    if (!p->tested)
        return fail(SBG_ERR_STATE, "sbg_tri_tri has not run on this pair list");
    TRY(use(p->ctx));
    if (!p->n)
        return SBG_OK;
    if (status)
        HIP_TRY(hipMemcpyAsync(status, p->status.p, p->n, hipMemcpyDeviceToHost, p->ctx->stream));
    if (seg)
        HIP_TRY(hipMemcpyAsync(seg, p->seg.p, sizeof(double) * 6 * p->n, hipMemcpyDeviceToHost, p->ctx->stream));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    return SBG_OK;
}

int sbg_pairs_hits(sbg_pairs p, size_t *nhits, uint32_t *ab, double *seg)
{
    if (!p || !nhits)
        return fail(SBG_ERR_INVALID, "null argument");
    if (!p->tested)
        return fail(SBG_ERR_STATE, "sbg_tri_tri has not run on this pair list");
    TRY(use(p->ctx));
    *nhits = 0;
    if (!p->n)
        return SBG_OK;
    uint32_t h = 0;
    HIP_TRY(hipMemcpyAsync(&h, p->nHits.p, sizeof(uint32_t), hipMemcpyDeviceToHost, p->ctx->stream));
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    *nhits = h;
    if (!h || (!ab && !seg))
        return SBG_OK;
    if (ab) {
        std::vector<uint64_t> v(h);
        HIP_TRY(hipMemcpyAsync(v.data(), p->hitPairs.p, sizeof(uint64_t) * h, hipMemcpyDeviceToHost, p->ctx->stream));
        HIP_TRY(hipStreamSynchronize(p->ctx->stream));
        for (uint32_t i = 0; i < h; ++i) {
            ab[2 * i] = (uint32_t)(v[i] >> 32);
            ab[2 * i + 1] = (uint32_t)v[i];
        }
    }
    if (seg) {
        HIP_TRY(hipMemcpyAsync(seg, p->hitSeg.p, sizeof(double) * 6 * h, hipMemcpyDeviceToHost, p->ctx->stream));
        HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    }
    return SBG_OK;
}

int sbg_tri_tri_batch(sbg_context ctx, const double *a9, const double *b9, size_t n, uint8_t *status, double *seg)
{
    TRY(use(ctx));
    if (n && (!a9 || !b9 || !status || !seg))
        return fail(SBG_ERR_INVALID, "null argument");
    if (n > 0xffffffffu)
        return fail(SBG_ERR_INVALID, "too many pairs");
    if (!n)
        return SBG_OK;
    TRY(ctx->rawA.ensure(9 * n));
    TRY(ctx->rawB.ensure(9 * n));
    TRY(ctx->rawSeg.ensure(6 * n));
    TRY(ctx->rawStatus.ensure(n));
    HIP_TRY(hipMemcpyAsync(ctx->rawA.p, a9, sizeof(double) * 9 * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->rawB.p, b9, sizeof(double) * 9 * n, hipMemcpyHostToDevice, ctx->stream));
    launch_tritri_raw(ctx->rawA.p, ctx->rawB.p, (uint32_t)n, ctx->rawStatus.p, ctx->rawSeg.p, ctx->stream);
    TRY(check_launch("tri_tri_raw"));
    HIP_TRY(hipMemcpyAsync(status, ctx->rawStatus.p, n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(seg, ctx->rawSeg.p, sizeof(double) * 6 * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SBG_OK;
}

// ---------------------------------------------------------------------------
// classification
// ---------------------------------------------------------------------------
// m's ray grid as the ray kernels read it; with `filtered`, rays decide
// their candidates from m's FP32 records under the certificate
static GridSet grid_set(sbg_mesh m, bool filtered = false)
{
    GridSet gs = {};
    gs.tri32 = filtered ? m->tri32.p : nullptr;
    gs.boxes = m->rootBox.p;
    gs.cells = m->gridCells.p;
    gs.gcode = m->gridCode.p;
    gs.gtri = m->gridTri.p;
    gs.G0 = m->gridG;
    gs.cap = (uint32_t)m->gridCap;
    gs.ncells = (uint32_t)(grid_cells(m->gridG) - 1);
    gs.order = m->sortedIds;
    gs.fastCands = (uint32_t)m->ctx->tune.ray_cell_cap;
    return gs;
}

// Builds m's ray grid on `stream` unless it is current; a reader on another
// stream than the build's waits for the build.
static int ensure_grid(sbg_mesh m, hipStream_t stream)
{
    sbg_context c = m->ctx;
    const hipStream_t s = stream ? stream : c->stream;
    if (m->gridReady && m->gridDensity == c->tune.grid_density) {
        if (m->gridStream && m->gridStream != s)
            HIP_TRY(hipStreamWaitEvent(s, m->gridDone, 0));
        return SBG_OK;
    }
    // a rebuild must not overwrite a grid that another stream may still read
    if (m->gridStream && m->gridStream != s)
        HIP_TRY(hipStreamWaitEvent(s, m->gridDone, 0));
    const uint32_t G = grid_dim((uint32_t)m->nt, c->tune.grid_density);
    m->gridDensity = c->tune.grid_density;
    // entries per triangle grow with the density: about 9 at density 2 and
    // 12 at 4 over the three axes on closed meshes; a grid that turns out too
    // small is rebuilt at its exact size
    if (!m->gridCap)
        m->gridCap = (6 + 3 * (size_t)c->tune.grid_density) * m->nt + 1024;
    if (m->gridCap > 0xffffffffu)
        return fail(SBG_ERR_OVERFLOW, "ray grid needs %zu entries", m->gridCap);
    TRY(m->gridCells.ensure(grid_cells(G)));
    TRY(m->gridTemp.ensure(scan_large_temp(grid_cells(G))));
    TRY(m->gridCode.ensure(m->gridCap));
    TRY(m->gridTri.ensure(m->gridCap));
    m->gridG = G;
    {
        Stage st(c, SBG_STAGE_GRID, s);
        launch_grid_build(m->box32.p, (uint32_t)m->nt, grid_set(m), m->gridCells.p, m->gridTemp.p, m->gridCode.p,
                          m->gridTri.p, kGridBesideSearch, s);
        TRY(check_launch("ray grid"));
    }
    m->gridStream = s;
    HIP_TRY(hipEventRecord(m->gridDone, s));
    m->gridReady = true;
    return SBG_OK;
}

// The ray grids of both meshes of a boolean, queued on the second stream
// during the pair search: they depend on the meshes only, and the search
// leaves most of the chip idle.  The second stream must already be ordered
// after the meshes' prepares (sbg_find_pairs forks it before the search).
static int queue_grids(sbg_mesh a, sbg_mesh b)
{
    sbg_context c = a->ctx;
    // b's first: the context stream's side classifies a's triangles in b's grid
    for (sbg_mesh m : {b, a})
        if (m->nt && !(m->gridReady && m->gridDensity == c->tune.grid_density))
            TRY(ensure_grid(m, c->aux));
    return SBG_OK;
}

// One classification side's scratch buffers and the words of c->pinned its
// readbacks land in.
struct RayBufs {
    DevBuf<double> &pts;
    DevBuf<uint8_t> &flags, &bits, &votes;
    DevBuf<uint32_t> &ovf, &rayList;
    uint32_t *host;  // 4 words: overflow flags, grid entries, -, -
    hipStream_t s;   // the stream this side's work goes on
    PinnedBuf &votesHost;
};

static RayBufs ray_bufs(sbg_context c, int set)
{
    if (set)
        return {c->pts2,   c->flags2, c->bits2, c->votes2, c->ovf2, c->rayList2, c->pinned + kRayPinned + 4,
                c->aux,    c->votesHost2};
    return {c->pts,    c->flags,  c->bits,  c->votes,  c->ovf,  c->rayList,  c->pinned + kRayPinned,
            c->stream, c->votesHost};
}

// Page-locked host memory (hipHostMalloc, sbg_host_alloc, or registered by
// the caller): a device-to-host copy into it is plain DMA.
static bool page_locked(const void *p)
{
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory is not an error here
        return false;
    }
    return at.type == hipMemoryTypeHost;
}

// Queues the rays of n queries into target's grid, the vote into rb.votes,
// and the readbacks (overflow flags and the grid's entry count into rb.host;
// with hostVotes the votes, straight into hostVotes when it is page-locked,
// else into rb.votesHost for the caller to copy out); nothing waits.  *direct
// says which.  ordered: dpts is the visiting-order (x, y, z, tag) layout over
// dorder; otherwise xyz triples by index with optional dorder / dskip.
static int enqueue_rays(sbg_mesh target, RayBufs rb, const double *dpts, const uint32_t *dorder,
                        const uint8_t *dskip, size_t n, bool ordered, bool allAxes, uint8_t *hostVotes,
                        bool *direct)
{
    sbg_context c = target->ctx;
    const bool filtered = c->tune.ray_precision == SBG_PRECISION_FP32_FILTERED;
    const bool votes = hostVotes != nullptr;
    *direct = votes && page_locked(hostVotes);
    TRY(rb.ovf.ensure(4));
    if (votes && !*direct)
        TRY(rb.votesHost.ensure(n));
    TRY(rb.rayList.ensure(ray_list_u32(n)));
    rb.host[0] = rb.host[1] = 0;
    if (target->nt == 0) {
        Stage st(c, SBG_STAGE_CLASSIFY, rb.s);
        launch_fill_u32((uint32_t *)rb.bits.p, 0u, (3 * n + 3) / 4, rb.s);
        launch_vote(rb.bits.p, dskip, (uint32_t)n, rb.votes.p, rb.s);
        TRY(check_launch("rays"));
    } else {
        TRY(ensure_grid(target, rb.s));
        Stage st(c, SBG_STAGE_CLASSIFY, rb.s);
        launch_fill_u32(rb.ovf.p, 0u, 4, rb.s);
        if (ordered)
            launch_rays_ordered(grid_set(target, filtered), target->tri64.p, dpts, dorder, nullptr, (uint32_t)n, allAxes, rb.bits.p,
                                rb.rayList.p, rb.ovf.p, rb.s);
        else
            launch_rays(grid_set(target, filtered), target->tri64.p, dpts, dorder, dskip, (uint32_t)n, allAxes, rb.bits.p,
                        rb.rayList.p, rb.ovf.p, rb.s);
        launch_vote(rb.bits.p, dskip, (uint32_t)n, rb.votes.p, rb.s);
        TRY(check_launch("rays"));
        HIP_TRY(hipMemcpyAsync(rb.host, rb.ovf.p, sizeof(uint32_t), hipMemcpyDeviceToHost, rb.s));
        HIP_TRY(hipMemcpyAsync(rb.host + 1, target->gridCells.p + (grid_cells(target->gridG) - 1),
                               sizeof(uint32_t), hipMemcpyDeviceToHost, rb.s));
    }
    if (votes)
        HIP_TRY(hipMemcpyAsync(*direct ? hostVotes : rb.votesHost.p, rb.votes.p, n, hipMemcpyDeviceToHost, rb.s));
    return SBG_OK;
}

// After the stream has drained: SBG_OK, or `retry` when the grid turned out
// too small (it is marked for a rebuild at its exact size), or an error.
static int check_rays(sbg_mesh target, RayBufs rb, bool &retry)
{
    retry = false;
    if (rb.host[0] & 2u) {
        target->gridCap = (size_t)rb.host[1] + rb.host[1] / 16 + 1024;
        target->gridReady = false;
        retry = true;
        return SBG_OK;
    }
    if (rb.host[0])
        return fail(SBG_ERR_OVERFLOW, "a ray has more distinct hits than its key list holds");
    return SBG_OK;
}

// enqueue, wait, check; a grid that turned out too small is rebuilt at its
// exact size and the rays run again
static int run_rays(sbg_mesh target, RayBufs rb, const double *dpts, const uint32_t *dorder, const uint8_t *dskip,
                    size_t n, bool ordered, bool allAxes, uint8_t *hostVotes)
{
    for (int attempt = 0; attempt < 2; ++attempt) {
        bool direct = false;
        TRY(enqueue_rays(target, rb, dpts, dorder, dskip, n, ordered, allAxes, hostVotes, &direct));
        HIP_TRY(hipStreamSynchronize(rb.s));
        bool retry = false;
        TRY(check_rays(target, rb, retry));
        if (!retry) {
            if (hostVotes && !direct)
                std::memcpy(hostVotes, rb.votesHost.p, n);
            return SBG_OK;
        }
    }
    return fail(SBG_ERR_OVERFLOW, "ray grid did not fit after a resize");
}

int sbg_classify_points(sbg_mesh target, const double *pts, size_t n, uint8_t *axis_inside, uint8_t *inside)
{
    if (!target || (n && !pts))
        return fail(SBG_ERR_INVALID, "null argument");
    if (!target->prepared)
        return fail(SBG_ERR_STATE, "mesh not prepared");
    if (n > 0x50000000u)
        return fail(SBG_ERR_INVALID, "too many points");
    sbg_context c = target->ctx;
    TRY(use(c));
    MeshUse mu(target);
    if (!n)
        return SBG_OK;
    TRY(c->pts.ensure(3 * n));
    TRY(c->bits.ensure(3 * n + 4));
    TRY(c->votes.ensure(n));
    HIP_TRY(hipMemcpyAsync(c->pts.p, pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
    TRY(run_rays(target, ray_bufs(c, 0), c->pts.p, nullptr, nullptr, n, false, axis_inside != nullptr, inside));
    std::vector<uint8_t> bits;
    if (axis_inside) {
        bits.resize(3 * n);
        HIP_TRY(hipMemcpyAsync(bits.data(), c->bits.p, 3 * n, hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (axis_inside)
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k)
                axis_inside[3 * i + k] = bits[(size_t)k * n + i];
    return SBG_OK;
}

static int check_classify_args(sbg_mesh self, sbg_mesh other, sbg_pairs p, int side)
{
    if (self->ctx != other->ctx || self->ctx != p->ctx)
        return fail(SBG_ERR_INVALID, "handles belong to different contexts");
    if (!self->prepared || !other->prepared)
        return fail(SBG_ERR_STATE, "mesh not prepared");
    if (!p->tested && p->n)
        return fail(SBG_ERR_STATE, "sbg_tri_tri has not run on this pair list");
    if (side != 0 && side != 1)
        return fail(SBG_ERR_INVALID, "side must be 0 or 1");
    if (p->n && (side == 0 ? p->maxA : p->maxB) >= self->nt)
        return fail(SBG_ERR_INVALID, "pair list indexes triangle %llu of a %zu-triangle mesh (wrong side?)",
                    (unsigned long long)(side == 0 ? p->maxA : p->maxB), self->nt);
    return SBG_OK;
}

// the queries of sbg_classify_triangles(self, ..., side): intersected flags,
// then the centroids of the others in visiting order, into rb
static int enqueue_queries(sbg_mesh self, sbg_pairs p, int side, RayBufs rb)
{
    sbg_context c = self->ctx;
    const size_t nt = self->nt;
    TRY(rb.flags.ensure(nt + 4));
    TRY(rb.pts.ensure(4 * nt));
    TRY(rb.bits.ensure(3 * nt + 4));
    TRY(rb.votes.ensure(nt));
    Stage st(c, SBG_STAGE_CLASSIFY, rb.s);
    launch_fill_u32((uint32_t *)rb.flags.p, 0u, (nt + 3) / 4, rb.s);
    launch_mark_intersected(p->pairs.p, p->status.p, (uint32_t)p->n, side, rb.flags.p, rb.s);
    launch_centroids_ordered(self->tri64.p, rb.flags.p, self->sortedIds, nullptr, (uint32_t)nt, rb.pts.p, rb.s);
    return check_launch("classify");
}

int sbg_classify_triangles(sbg_mesh self, sbg_mesh other, sbg_pairs p, int side, uint8_t *out)
{
    if (!self || !other || !p || !out)
        return fail(SBG_ERR_INVALID, "null argument");
    TRY(check_classify_args(self, other, p, side));
    sbg_context c = self->ctx;
    TRY(use(c));
    MeshUse mu(self, other);
    if (!self->nt)
        return SBG_OK;
    const RayBufs rb = ray_bufs(c, 0);
    TRY(enqueue_queries(self, p, side, rb));
    return run_rays(other, rb, rb.pts.p, self->sortedIds, rb.flags.p, self->nt, true, false, out);
}

int sbg_classify_pair(sbg_mesh a, sbg_mesh b, sbg_pairs p, uint8_t *out_a, uint8_t *out_b)
{
    if (!a || !b || !p || !out_a || !out_b)
        return fail(SBG_ERR_INVALID, "null argument");
    TRY(check_classify_args(a, b, p, 0));
    TRY(check_classify_args(b, a, p, 1));
    sbg_context c = a->ctx;
    TRY(use(c));
    MeshUse mu(a, b);
    AuxJoin join{c};
    // side 0 on the context stream, side 1 on the second stream and scratch
    // set, ordered after everything queued so far; one wait for both
    const RayBufs ra = ray_bufs(c, 0), rb = ray_bufs(c, 1);
    HIP_TRY(hipEventRecord(c->auxReady, c->stream));
    HIP_TRY(hipStreamWaitEvent(c->aux, c->auxReady, 0));
    // both sides' queries first, then both sides' rays: the second stream
    // starts its work after a few launches instead of after the first side's
    // whole sequence
    if (a->nt)
        TRY(enqueue_queries(a, p, 0, ra));
    if (b->nt)
        TRY(enqueue_queries(b, p, 1, rb));
    bool directA = false, directB = false;
    if (a->nt)
        TRY(enqueue_rays(b, ra, ra.pts.p, a->sortedIds, ra.flags.p, a->nt, true, false, out_a, &directA));
    if (b->nt)
        TRY(enqueue_rays(a, rb, rb.pts.p, b->sortedIds, rb.flags.p, b->nt, true, false, out_b, &directB));
    // each side's votes land in page-locked output arrays directly, or are
    // copied out of the staging as soon as that side is done (the first
    // side's copy overlapping the second side's GPU work); a side whose
    // target grid turned out too small runs again on its own afterwards
    bool retryA = false, retryB = false;
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (a->nt) {
        TRY(check_rays(b, ra, retryA));
        if (!retryA && !directA)
            std::memcpy(out_a, ra.votesHost.p, a->nt);
    }
    HIP_TRY(hipStreamSynchronize(c->aux));
    if (b->nt) {
        TRY(check_rays(a, rb, retryB));
        if (!retryB && !directB)
            std::memcpy(out_b, rb.votesHost.p, b->nt);
    }
    if (retryA)
        TRY(run_rays(b, ra, ra.pts.p, a->sortedIds, ra.flags.p, a->nt, true, false, out_a));
    if (retryB)
        TRY(run_rays(a, rb, rb.pts.p, b->sortedIds, rb.flags.p, b->nt, true, false, out_b));
    return SBG_OK;
}

int sbg_context_set_tuning(sbg_context ctx, const sbg_tuning *t)
{
    TRY(use(ctx));
    if (!t)
        return fail(SBG_ERR_INVALID, "null tuning");
    if (t->bfs_levels < 0 || t->bfs_levels > 64 || t->dfs_waves < 64 || t->dfs_waves > 16384 ||
        t->grid_density < 1 || t->grid_density > 64 || t->early_grids < 0 || t->early_grids > 2 ||
        t->ray_cell_cap < 16 || t->ray_cell_cap > 4096 ||
        (t->ray_precision != SBG_PRECISION_FP64 && t->ray_precision != SBG_PRECISION_FP32_FILTERED))
        return fail(SBG_ERR_INVALID, "tuning out of range");
    ctx->tune = *t;
    ctx->tune.depth_first = t->depth_first ? 1 : 0;
    return SBG_OK;
}

int sbg_context_get_tuning(sbg_context ctx, sbg_tuning *t)
{
    TRY(use(ctx));
    if (!t)
        return fail(SBG_ERR_INVALID, "null tuning");
    *t = ctx->tune;
    return SBG_OK;
}

// wall time of prepare both -> find_pairs -> tri_tri -> classify both
// meshes, `reps` times
static int time_pair_chain(sbg_mesh a, sbg_mesh b, int reps, double &ms)
{
    std::vector<uint8_t> ca(a->nt + 1), cb(b->nt + 1);
    auto once = [&]() -> int {
        TRY(sbg_mesh_prepare(a));
        TRY(sbg_mesh_prepare(b));
        sbg_pairs p = nullptr;
        TRY(sbg_find_pairs(a, b, &p));
        int rc = sbg_tri_tri(a, b, p, SBG_PRECISION_FP32_FILTERED, nullptr);
        if (rc == SBG_OK)
            rc = sbg_classify_pair(a, b, p, ca.data(), cb.data());
        sbg_pairs_destroy(p);
        return rc;
    };
    TRY(once());  // builds grids and sizes queues for this setting
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r)
        TRY(once());
    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
    return SBG_OK;
}

int sbg_autotune_pair(sbg_mesh a, sbg_mesh b, int reps, sbg_tuning *best, double *best_ms)
{
    if (!a || !b || reps <= 0)
        return fail(SBG_ERR_INVALID, "bad argument");
    if (a->ctx != b->ctx)
        return fail(SBG_ERR_INVALID, "meshes belong to different contexts");
    if (!a->prepared || !b->prepared)
        return fail(SBG_ERR_STATE, "mesh not prepared");
    sbg_context c = a->ctx;
    TRY(use(c));
    sbg_tuning cur = c->tune;
    double bestMs = 0.0;
    TRY(time_pair_chain(a, b, reps, bestMs));
    const int levels[] = {8, 12, 16, 20, 24};
    const int waves[] = {256, 512, 1024, 2048, 4096};
    const int density[] = {2, 4, 8, 16};
    const int cellCap[] = {48, 64, 96, 128};
    auto trial = [&](const sbg_tuning &t) -> int {
        c->tune = t;
        double ms = 0.0;
        TRY(time_pair_chain(a, b, reps, ms));
        if (ms < bestMs) {
            bestMs = ms;
            cur = t;
        }
        return SBG_OK;
    };
    for (int pass = 0; pass < 2; ++pass) {
        for (int v : levels)
            if (v != cur.bfs_levels) {
                sbg_tuning t = cur;
                t.bfs_levels = v;
                TRY(trial(t));
            }
        for (int v : waves)
            if (v != cur.dfs_waves) {
                sbg_tuning t = cur;
                t.dfs_waves = v;
                TRY(trial(t));
            }
        for (int v : density)
            if (v != cur.grid_density) {
                sbg_tuning t = cur;
                t.grid_density = v;
                TRY(trial(t));
            }
        for (int v : cellCap)
            if (v != cur.ray_cell_cap) {
                sbg_tuning t = cur;
                t.ray_cell_cap = v;
                TRY(trial(t));
            }
        for (int v = 0; v <= 2; ++v)
            if (v != cur.early_grids) {
                sbg_tuning t = cur;
                t.early_grids = v;
                TRY(trial(t));
            }
    }
    c->tune = cur;
    if (best)
        *best = cur;
    if (best_ms)
        *best_ms = bestMs;
    return SBG_OK;
}

int sbg_adjoint(sbg_context ctx, const double *xyz_in, size_t nin, const uint32_t *src, size_t nd,
                const uint32_t *faces, size_t nf, double vol_weight, const double *grad_derived, double *derived_xyz,
                double *volume, double *grad_in)
{
    TRY(use(ctx));
    if ((nin && !xyz_in) || (nd && !src) || (nf && !faces))
        return fail(SBG_ERR_INVALID, "null input array");
    if (nin + nd >= 0x7fffffffu || nf >= 0x7fffffffu)
        return fail(SBG_ERR_INVALID, "too many vertices or faces");
    for (size_t i = 0; i < 5 * nd; ++i)
        if (src[i] >= nin)
            return fail(SBG_ERR_INVALID, "derived vertex %zu names input %u of %zu", i / 5, src[i], nin);
    for (size_t i = 0; i < 3 * nf; ++i)
        if (faces[i] >= nin + nd)
            return fail(SBG_ERR_INVALID, "face %zu names vertex %u of %zu", i / 3, faces[i], nin + nd);
    const size_t nall = nin + nd;
    TRY(ctx->adjXyz.ensure(3 * nin + 3));
    TRY(ctx->adjAll.ensure(3 * nall + 3));
    TRY(ctx->adjGall.ensure(3 * nall + 3));
    TRY(ctx->adjGradD.ensure(3 * nd + 3));
    TRY(ctx->adjVol.ensure(1));
    TRY(ctx->adjSrc.ensure(5 * nd + 5));
    TRY(ctx->adjFaces.ensure(3 * nf + 3));
    hipStream_t s = ctx->stream;
    if (nin)
        HIP_TRY(hipMemcpyAsync(ctx->adjXyz.p, xyz_in, sizeof(double) * 3 * nin, hipMemcpyHostToDevice, s));
    if (nd)
        HIP_TRY(hipMemcpyAsync(ctx->adjSrc.p, src, sizeof(uint32_t) * 5 * nd, hipMemcpyHostToDevice, s));
    if (nf)
        HIP_TRY(hipMemcpyAsync(ctx->adjFaces.p, faces, sizeof(uint32_t) * 3 * nf, hipMemcpyHostToDevice, s));
    const bool wantGrad = grad_in != nullptr;
    if (wantGrad && grad_derived && nd)
        HIP_TRY(hipMemcpyAsync(ctx->adjGradD.p, grad_derived, sizeof(double) * 3 * nd, hipMemcpyHostToDevice, s));
    launch_fill_u32(reinterpret_cast<uint32_t *>(ctx->adjVol.p), 0u, 2, s);
    if (wantGrad)
        launch_fill_u32(reinterpret_cast<uint32_t *>(ctx->adjGall.p), 0u, 6 * nall, s);
    launch_adjoint(ctx->adjXyz.p, (uint32_t)nin, ctx->adjSrc.p, (uint32_t)nd, ctx->adjFaces.p, (uint32_t)nf,
                   vol_weight, (wantGrad && grad_derived) ? ctx->adjGradD.p : nullptr, ctx->adjAll.p, ctx->adjVol.p,
                   wantGrad ? ctx->adjGall.p : nullptr, s);
    TRY(check_launch("adjoint"));
    if (derived_xyz && nd)
        HIP_TRY(hipMemcpyAsync(derived_xyz, ctx->adjAll.p + 3 * nin, sizeof(double) * 3 * nd, hipMemcpyDeviceToHost,
                               s));
    if (volume)
        HIP_TRY(hipMemcpyAsync(volume, ctx->adjVol.p, sizeof(double), hipMemcpyDeviceToHost, s));
    if (wantGrad && nin)
        HIP_TRY(hipMemcpyAsync(grad_in, ctx->adjGall.p, sizeof(double) * 3 * nin, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SBG_OK;
}

int sbg_calibrate_copy(sbg_context ctx, size_t bytes, int iters, double *avg_ms)
{
    TRY(use(ctx));
    if (!avg_ms || iters <= 0 || bytes < 16)
        return fail(SBG_ERR_INVALID, "bad argument");
    const size_t n16 = bytes / 16;
    DevBuf<uint4> src, dst;
    TRY(src.ensure(n16));
    int rc = dst.ensure(n16);
    if (rc != SBG_OK) {
        src.release();
        return rc;
    }
    launch_fill_u32((uint32_t *)src.p, 1u, 4 * n16, ctx->stream);
    launch_copy16(src.p, dst.p, n16, ctx->stream);  // warm
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess)
        e = hipEventCreate(&e1);
    if (e == hipSuccess)
        e = hipEventRecord(e0, ctx->stream);
    for (int i = 0; e == hipSuccess && i < iters; ++i)
        launch_copy16(src.p, dst.p, n16, ctx->stream);
    if (e == hipSuccess)
        e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess)
        e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess)
        e = hipEventElapsedTime(&ms, e0, e1);
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    src.release();
    dst.release();
    if (e != hipSuccess)
        return fail(SBG_ERR_HIP, "calibration copy failed: %s", hipGetErrorString(e));
    *avg_ms = (double)ms / iters;
    return check_launch("copy");
}

}  // extern "C"

// the argument checks of the ray measurement aids
static int bench_rays_check(sbg_mesh self, sbg_mesh other, sbg_pairs p, int side, int iters)
{
    if (!self || !other || !p || iters <= 0)
        return fail(SBG_ERR_INVALID, "bad argument");
    if (self->ctx != other->ctx || self->ctx != p->ctx)
        return fail(SBG_ERR_INVALID, "handles belong to different contexts");
    if (!self->prepared || !other->prepared)
        return fail(SBG_ERR_STATE, "mesh not prepared");
    if (!p->tested && p->n)
        return fail(SBG_ERR_STATE, "sbg_tri_tri has not run on this pair list");
    if (side != 0 && side != 1)
        return fail(SBG_ERR_INVALID, "side must be 0 or 1");
    if (p->n && (side == 0 ? p->maxA : p->maxB) >= self->nt)
        return fail(SBG_ERR_INVALID, "pair list does not index this mesh on side %d", side);
    return use(self->ctx);
}

// their set-up: the queries of sbg_classify_triangles(self, other, p, side)
// in c->pts, other's grid built; *empty when there is nothing to cast
static int bench_rays_setup(sbg_mesh self, sbg_mesh other, sbg_pairs p, int side, bool &empty)
{
    empty = true;
    sbg_context c = self->ctx;
    const size_t nt = self->nt;
    if (!nt || !other->nt)
        return SBG_OK;
    TRY(c->flags.ensure(nt + 4));
    TRY(c->pts.ensure(4 * nt));
    TRY(c->bits.ensure(3 * nt + 4));
    TRY(c->ovf.ensure(4));
    TRY(c->rayList.ensure(ray_list_u32(nt)));
    TRY(c->workCounts.ensure(4));
    launch_fill_u32((uint32_t *)c->flags.p, 0u, (nt + 3) / 4, c->stream);
    launch_mark_intersected(p->pairs.p, p->status.p, (uint32_t)p->n, side, c->flags.p, c->stream);
    launch_centroids_ordered(self->tri64.p, c->flags.p, self->sortedIds, nullptr, (uint32_t)nt, c->pts.p, c->stream);
    TRY(ensure_grid(other, c->stream));
    empty = false;
    return SBG_OK;
}

extern "C" {

int sbg_bench_rays(sbg_mesh self, sbg_mesh other, sbg_pairs p, int side, int iters, double *avg_ms,
                   uint64_t *work)
{
    if (!avg_ms)
        return fail(SBG_ERR_INVALID, "bad argument");
    *avg_ms = 0.0;
    if (work)
        for (int k = 0; k < 5; ++k)
            work[k] = 0;
    TRY(bench_rays_check(self, other, p, side, iters));
    MeshUse mu(self, other);
    bool empty = true;
    TRY(bench_rays_setup(self, other, p, side, empty));
    if (empty)
        return SBG_OK;
    sbg_context c = self->ctx;
    const size_t nt = self->nt;
    const GridSet gs = grid_set(other, c->tune.ray_precision == SBG_PRECISION_FP32_FILTERED);
    std::vector<hipEvent_t> ev(2 * (size_t)iters, nullptr);
    for (auto &e : ev)
        HIP_TRY(hipEventCreate(&e));
    for (int i = 0; i < iters; ++i) {
        launch_fill_u32(c->ovf.p, 0u, 4, c->stream);
        HIP_TRY(hipEventRecord(ev[2 * i], c->stream));
        launch_rays_xy_ordered(gs, other->tri64.p, c->pts.p, self->sortedIds, (uint32_t)nt, c->bits.p, c->rayList.p,
                               c->ovf.p, c->stream);
        HIP_TRY(hipEventRecord(ev[2 * i + 1], c->stream));
    }
    launch_fill_u32((uint32_t *)c->workCounts.p, 0u, 8, c->stream);
    launch_ray_work_ordered(gs, c->pts.p, self->sortedIds, (uint32_t)nt, c->workCounts.p, c->stream);
    uint64_t h[4] = {0, 0, 0, 0};
    uint32_t deferred = 0;  // rays the last timed launch handed to the wave kernel
    HIP_TRY(hipMemcpyAsync(h, c->workCounts.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(&deferred, c->ovf.p + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double total = 0.0;
    for (int i = 0; i < iters; ++i) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
        total += ms;
    }
    for (auto &e : ev)
        (void)hipEventDestroy(e);
    *avg_ms = total / iters;
    if (work) {
        for (int k = 0; k < 4; ++k)
            work[k] = h[k];
        work[4] = deferred;
    }
    return check_launch("bench rays");
}

int sbg_bench_rays_wave(sbg_mesh self, sbg_mesh other, sbg_pairs p, int side, int iters, double *avg_ms,
                        uint64_t *rays)
{
    if (!avg_ms)
        return fail(SBG_ERR_INVALID, "bad argument");
    *avg_ms = 0.0;
    if (rays)
        *rays = 0;
    TRY(bench_rays_check(self, other, p, side, iters));
    MeshUse mu(self, other);
    bool empty = true;
    TRY(bench_rays_setup(self, other, p, side, empty));
    if (empty)
        return SBG_OK;
    sbg_context c = self->ctx;
    const uint32_t nt = (uint32_t)self->nt;
    const GridSet gs = grid_set(other, c->tune.ray_precision == SBG_PRECISION_FP32_FILTERED);
    launch_fill_u32(c->ovf.p, 0u, 4, c->stream);
    launch_rays_xy_ordered(gs, other->tri64.p, c->pts.p, self->sortedIds, nt, c->bits.p, c->rayList.p, c->ovf.p,
                           c->stream);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; ++i)
        launch_rays_wave_ordered(gs, other->tri64.p, c->pts.p, self->sortedIds, nt, c->bits.p, c->rayList.p,
                                 c->ovf.p, c->stream);
    HIP_TRY(hipEventRecord(e1, c->stream));
    uint32_t deferred = 0;
    HIP_TRY(hipMemcpyAsync(&deferred, c->ovf.p + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / iters;
    if (rays)
        *rays = deferred;
    return check_launch("bench rays wave");
}

// Testing aid, run on the host: the ray kernels' FP32 certificate for n
// (point, axis, triangle) candidates, from the same 48-B records they read.
int sbg_ray_certify(const double *pts, const int32_t *axis, const double *tris9, size_t n, int32_t *verdict,
                    float *t, float *et)
{
    if (n && (!pts || !axis || !tris9 || !verdict || !t || !et))
        return fail(SBG_ERR_INVALID, "null argument");
    for (size_t i = 0; i < n; ++i) {
        if (axis[i] < 0 || axis[i] > 2)
            return fail(SBG_ERR_INVALID, "axis %d of candidate %zu", axis[i], i);
        float4 r[3];
        make_tri32(tris9 + 9 * i, tris9 + 9 * i + 3, tris9 + 9 * i + 6, r);
        const double *p = pts + 3 * i;
        const float4 p32 = make_float4((float)p[0], (float)p[1], (float)p[2], 0.0f);
        double pm = __builtin_fabs(p[0]);
        pm = __builtin_fabs(p[1]) > pm ? __builtin_fabs(p[1]) : pm;
        pm = __builtin_fabs(p[2]) > pm ? __builtin_fabs(p[2]) : pm;
        const RayCert c = ray_certificate(p32, fl_up(pm), axis[i], r);
        verdict[i] = c.verdict;
        t[i] = c.t;
        et[i] = c.et;
    }
    return SBG_OK;
}

// Testing aid, run on the host: hits_apart() of the ray kernels for n pairs.
int sbg_ray_hits_apart(const float *ta, const float *ea, const float *tb, const float *eb, size_t n, uint8_t *out)
{
    if (n && (!ta || !ea || !tb || !eb || !out))
        return fail(SBG_ERR_INVALID, "null argument");
    for (size_t i = 0; i < n; ++i)
        out[i] = hits_apart(ta[i], ea[i], tb[i], eb[i]) ? 1 : 0;
    return SBG_OK;
}

int sbg_bench_tri_tri(sbg_mesh a, sbg_mesh b, sbg_pairs p, int precision, int iters, double *avg_ms)
{
    if (!a || !b || !p || !avg_ms || iters <= 0)
        return fail(SBG_ERR_INVALID, "bad argument");
    if (precision != SBG_PRECISION_FP64 && precision != SBG_PRECISION_FP32_FILTERED)
        return fail(SBG_ERR_INVALID, "unknown precision %d", precision);
    sbg_context c = p->ctx;
    TRY(use(c));
    MeshUse mu(a, b);
    TRY(check_pairs_in_range(a, b, p));
    TRY(p->status.ensure(p->n));
    TRY(p->seg.ensure(6 * p->n));
    TRY(p->work.ensure(tritri_work_u32((uint32_t)p->n)));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, c->stream));
    for (int i = 0; i < iters; ++i)
        launch_tritri(a->tri64.p, b->tri64.p, a->tri32.p, b->tri32.p, p->pairs.p, (uint32_t)p->n, nullptr, precision,
                      p->status.p, p->seg.p, p->work.p, nullptr, c->stream);
    HIP_TRY(hipEventRecord(e1, c->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / iters;
    return check_launch("bench tri_tri");
}

}  // extern "C"

// ---- allocation-free classification ----------------------------------------
// sbg_classify_points_noalloc keeps every buffer its kernels use beyond the
// caller's inputs and outputs in one file-scope device array.  Each device
// holds its own copy at its own address, resolved once per device by
// hipGetSymbolAddress when a context is created there; host code uses only
// that resolved address (the array's name is a host-side stand-in).
namespace sbg {
namespace {
constexpr size_t kNoAllocMaxPoints = (size_t)1 << 16;
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
// layout for up to kNoAllocMaxPoints points: the points (3 doubles each),
// per-axis ray bits (3 n + 4), votes (n), per-point axis triples (3 n), four
// counters, and the deferred-ray lists (ray_list_u32(n) = 25 n u32)
constexpr size_t kNaPts = 0;
constexpr size_t kNaBits = align256(kNaPts + 24 * kNoAllocMaxPoints);
constexpr size_t kNaVotes = align256(kNaBits + 3 * kNoAllocMaxPoints + 4);
constexpr size_t kNaAxes = align256(kNaVotes + kNoAllocMaxPoints);
constexpr size_t kNaCounters = align256(kNaAxes + 3 * kNoAllocMaxPoints);
constexpr size_t kNaLists = align256(kNaCounters + 16);
constexpr size_t kNoAllocBytes = align256(kNaLists + 4 * 25 * kNoAllocMaxPoints);
constexpr int kNoAllocDevices = 64;

std::mutex g_scratchMu;
uint8_t *g_scratchPtr[kNoAllocDevices] = {};
}  // namespace
}  // namespace sbg

static __device__ __align__(256) unsigned char scratch[sbg::kNoAllocBytes];

namespace sbg {
int noalloc_attach(int device)
{
    if (device < 0 || device >= kNoAllocDevices)
        return fail(SBG_ERR_INVALID, "device %d beyond the %d the library tracks", device, kNoAllocDevices);
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (g_scratchPtr[device])
        return SBG_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SBG_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratchPtr[device] = static_cast<uint8_t *>(p);
    return SBG_OK;
}
}  // namespace sbg

extern "C" {

int sbg_classify_points_noalloc(sbg_mesh target, const double *pts, size_t n, uint8_t *axis_inside,
                                uint8_t *inside)
{
    if (!target || (n && !pts))
        return fail(SBG_ERR_INVALID, "null argument");
    if (!target->prepared)
        return fail(SBG_ERR_STATE, "mesh not prepared");
    if (n > kNoAllocMaxPoints)
        return fail(SBG_ERR_INVALID, "%zu points exceed the allocation-free limit of %zu", n, kNoAllocMaxPoints);
    sbg_context c = target->ctx;
    TRY(use(c));
    uint8_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratchMu);
        base = g_scratchPtr[c->device];
    }
    if (!base)
        return fail(SBG_ERR_STATE, "scratch array of device %d not resolved", c->device);
    if (!n)
        return SBG_OK;
    if (target->nt && !(target->gridReady && target->gridDensity == c->tune.grid_density))
        return fail(SBG_ERR_STATE, "the target's ray grid is not built: classify against it once with an "
                                   "allocating call (sbg_classify_points) first");
    const hipStream_t s = c->stream;
    MeshUse mu(target);
    if (target->gridStream && target->gridStream != s)
        HIP_TRY(hipStreamWaitEvent(s, target->gridDone, 0));
    double *dpts = reinterpret_cast<double *>(base + kNaPts);
    uint8_t *bits = base + kNaBits, *votes = base + kNaVotes, *axes = base + kNaAxes;
    uint32_t *counters = reinterpret_cast<uint32_t *>(base + kNaCounters);
    uint32_t *lists = reinterpret_cast<uint32_t *>(base + kNaLists);
    const bool allAxes = axis_inside != nullptr;
    uint32_t *host = c->pinned + kRayPinned;  // page-locked words made with the context
    host[0] = 0;
    HIP_TRY(hipMemcpyAsync(dpts, pts, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s));
    if (target->nt == 0) {
        launch_fill_u32(reinterpret_cast<uint32_t *>(bits), 0u, (3 * n + 3) / 4, s);
    } else {
        launch_fill_u32(counters, 0u, 4, s);  // kernels, not memsets: the call may be captured
        launch_rays(grid_set(target, c->tune.ray_precision == SBG_PRECISION_FP32_FILTERED), target->tri64.p, dpts,
                    nullptr, nullptr, (uint32_t)n, allAxes, bits, lists,
                    counters, s);
        HIP_TRY(hipMemcpyAsync(host, counters, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    }
    launch_vote(bits, nullptr, (uint32_t)n, votes, s);
    if (allAxes)
        launch_interleave_axes(bits, (uint32_t)n, axes, s);
    TRY(check_launch("allocation-free rays"));
    HIP_TRY(hipStreamSynchronize(s));
    if (host[0] & 2u) {
        target->gridReady = false;  // the next allocating call rebuilds it at its exact size
        return fail(SBG_ERR_STATE, "the target's ray grid was truncated: classify against it once with an "
                                   "allocating call (sbg_classify_points) first");
    }
    if (host[0])
        return fail(SBG_ERR_OVERFLOW, "a ray has more distinct hits than its key list holds");
    if (inside)
        HIP_TRY(hipMemcpy(inside, votes, n, hipMemcpyDeviceToHost));
    if (axis_inside)
        HIP_TRY(hipMemcpy(axis_inside, axes, 3 * n, hipMemcpyDeviceToHost));
    return SBG_OK;
}

}  // extern "C"
