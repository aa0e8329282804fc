// Host-side internals shared by the C-ABI translation units: error
// reporting, growable device buffers, the context record and stage timing.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "sbg_internal.h"
#include "sbgpu.h"

namespace sbg {

// records the message for sbg_last_error() and returns `code`
int fail(int code, const char *fmt, ...);

#define HIP_TRY(expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return fail(SBG_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

#define TRY(expr)              \
    do {                       \
        int r_ = (expr);       \
        if (r_ != SBG_OK)      \
            return r_;         \
    } while (0)

// growable device buffer (contents are not preserved on growth)
// page-locked host staging (hipHostMalloc), grown on demand: device-to-host
// copies into it are plain DMA on the stream and never block the host
struct PinnedBuf {
    uint8_t *p = nullptr;
    size_t cap = 0;
    int ensure(size_t n)
    {
        if (n <= cap && p)
            return SBG_OK;
        release();
        const size_t count = n ? n : 1;
        if (hipHostMalloc((void **)&p, count, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(SBG_ERR_OOM, "hipHostMalloc of %zu bytes failed", count);
        }
        cap = count;
        return SBG_OK;
    }
    void release()
    {
        if (p)
            (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    int ensure(size_t n)
    {
        if (n <= cap && p)
            return SBG_OK;
        release();
        const size_t count = n ? n : 1;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(SBG_ERR_OOM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        }
        cap = count;
        return SBG_OK;
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

constexpr int kMaxLevels = 256;
constexpr int kComponentsPinned = kMaxLevels + 7;  // sbg_components' repeated-half-edge flag
constexpr int kRayPinned = kMaxLevels + 8;      // 4 words per classification side
constexpr int kPinnedWords = kRayPinned + 8;
constexpr int kExpandGrid = 1024;
constexpr int kLevelsPerSync = 8;
// breadth-first levels before the persistent depth-first kernel takes the
// frontier, its wave count, and the initial / largest spill region per wave
constexpr int kBfsLevels = 24;
constexpr int kDfsWaves = 512;
// the single-pair search expands levels of at most kTopMax node pairs in one
// workgroup, then queues kExtraLevels grid-wide levels (no-ops once
// bfs_levels is reached) before the depth-first stage
constexpr uint32_t kTopMax = 2048;
constexpr int kExtraLevels = 8;
// longest run of one first triangle the quick pair sort orders by insertion
constexpr uint32_t kSortRunMax = 64;
constexpr uint32_t kSpillInit = 2048, kSpillMax = 1u << 15;
// sbg_mesh_create's uploads: host threads that fill page-locked staging (two
// slots each), the slot size, and the smallest share worth a thread of its own
constexpr int kUploadThreads = 8;
constexpr size_t kStageSlot = (size_t)2 << 20;
constexpr size_t kUploadShare = (size_t)2 << 20;
// the threads' copies alternate between this many streams, so the DMA of
// one waits for no other's
constexpr int kUploadStreams = 2;

}  // namespace sbg

using namespace sbg;

struct sbg_context_s {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;        // second stream (default priority): side 1 of sbg_classify_pair
    hipEvent_t auxReady = nullptr;    // orders the aux stream after the main one
    hipStream_t grids = nullptr;      // third stream (default priority): ray grids built at prepare time
    uint32_t *pinned = nullptr;  // host-visible counters (kPinnedWords): search, then two ray sides
    // mesh uploads: their own streams, page-locked staging of 2 x kUploadThreads
    // slots and the event after each slot's last copy (made on first use)
    hipStream_t upload[kUploadStreams] = {};
    PinnedBuf staging;
    hipEvent_t stageDone[kUploadThreads][2] = {};
    DevBuf<uint2> front[2];
    DevBuf<uint2> spill;              // depth-first spill regions
    uint32_t spillPerWave = kSpillInit;
    bool depthFirst = true;           // false after the spill regions proved too small
    sbg_tuning tune = {kBfsLevels, kDfsWaves, 2, 1, 2, SBG_PRECISION_FP32_FILTERED, 64};
    DevBuf<uint32_t> counters;
    size_t frontCap = 0, candCap = 0;
    // candidates expected from the next search (the last count plus a margin):
    // the pair sort is queued over that many padded slots before the host
    // reads the count, and redone on the exact count if more were found
    size_t sortHint = 0;
    DevBuf<uint32_t> hist;
    DevBuf<GdStats> stats;
    // classification scratch, two sets (side 0 and side 1 of sbg_classify_pair;
    // single calls and other scratch uses take set 0)
    DevBuf<double> pts, pts2;
    DevBuf<uint8_t> flags, bits, votes, flags2, bits2, votes2;
    DevBuf<uint32_t> ovf, rayList, ovf2, rayList2;
    PinnedBuf votesHost, votesHost2;  // vote readbacks of the two sets
    DevBuf<unsigned long long> workCounts;  // sbg_bench_rays
    DevBuf<unsigned long long> ccKeys;      // sbg_components: half-edge table
    DevBuf<uint32_t> ccOwners, ccParent, ccLabels;
    DevBuf<uint8_t> ccBlocked;
    DevBuf<uint32_t> heTris, heOpposite, heQueryOwner, heRepeat;  // sbg_half_edges
    DevBuf<unsigned long long> heQueries;
    DevBuf<double> rawA, rawB, rawSeg;
    DevBuf<uint8_t> rawStatus;
    DevBuf<double> adjXyz, adjAll, adjGradD, adjGall, adjVol;  // sbg_adjoint
    DevBuf<uint32_t> adjSrc, adjFaces;
    int refs = 1;  // the owner + every live mesh / pair list
    bool destroyed = false;
    bool timing = false;
    struct Ev {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Ev> events;
    double acc[SBG_STAGE_COUNT] = {0};
    // the buffers of the last destroyed pair list and meshes, handed to the
    // next ones
    struct sbg_pairs_s *sparePairs = nullptr;
    std::vector<struct sbg_mesh_s *> spareMeshes;
    // meshes prepared with their ray grid still to be queued on the grid
    // stream (queued after the next prepare of another mesh or at the next
    // pair search, so the host queues both prepares first)
    std::vector<struct sbg_mesh_s *> gridsPending;
    unsigned meshCount = 0;  // meshes made so far: even ones prepare on `stream`, odd ones on `aux`
};

namespace sbg {

// validates the handle and makes its device current
int use(sbg_context ctx);
// Meshes and pair lists hold a reference on their context, so destroying a
// context while they live only marks it; the last release frees it.
void context_release(sbg_context ctx);
void context_free(sbg_context ctx);

// HIP-event bracket of one stage (no-op unless timing is enabled)
struct Stage {
    sbg_context c;
    int stage;
    hipStream_t s;
    hipEvent_t a = nullptr;
    Stage(sbg_context ctx, int st, hipStream_t stream = nullptr) : c(ctx), stage(st), s(stream ? stream : ctx->stream)
    {
        if (c->timing && hipEventCreate(&a) == hipSuccess)
            (void)hipEventRecord(a, s);
        else
            a = nullptr;
    }
    ~Stage()
    {
        if (!a)
            return;
        hipEvent_t b = nullptr;
        if (hipEventCreate(&b) == hipSuccess) {
            (void)hipEventRecord(b, s);
            c->events.push_back({stage, a, b});
        } else {
            (void)hipEventDestroy(a);
        }
    }
};

int check_launch(const char *what);
int ceil_log2(size_t n);

}  // namespace sbg
