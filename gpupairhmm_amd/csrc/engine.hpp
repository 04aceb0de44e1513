// The engine object behind the C ABI (engine.cpp) and the serving loop
// (server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "gpupairhmm.h"
#include "host_batch.hpp"
#include "phmm_device.hpp"

namespace phmm {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct PinBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct Engine {
    int device = 0;
    phmm_options opt{};
    hipStream_t stream = nullptr;     // compute
    hipStream_t copy_stream = nullptr;
    std::vector<phmm_tile_config> tiles;
    // The one-template table of small batches (options.small_batch): the
    // classes of `tiles`, 16 rows per lane each (where 64 lanes hold them).
    std::vector<phmm_tile_config> tiles_small;
    const std::vector<phmm_tile_config>* act = &tiles;  // the current call's table
    void derive_small_tiles();
    float* d_ph2pr = nullptr;
    int num_cus = 256;
    std::string err;
    phmm_stats stats{};

    // scratch (grow-only)
    DevBuf rows, hapc, meta, gb_count, gb_start, read_perm, align_off, key_count, key_cursor,
        bin_range, align_range, smallints, items, supertasks, rescue_items, final_items, spill,
        hapflag;
    int32_t* h_bin_range = nullptr;   // pinned: 2*kMaxWorkBins work ranges, then 2*kMaxBins
                                      // alignment ranges
    int32_t* h_small = nullptr;       // pinned, small readbacks
    // event pairs around the forward kernels of timed calls
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_free, ev_pending;

    int64_t* d_rescued_total = nullptr;  // cumulative FP64 rescues

    // The per-call batch descriptor: one device copy, staged through a ring
    // of pinned slots (a slot is rewritten only after its copy ran).
    static constexpr int kDescRing = 16;
    DevBuf desc;
    BatchDesc* h_desc = nullptr;
    hipEvent_t desc_ev[kDescRing] = {};
    bool desc_used[kDescRing] = {};
    int desc_next = 0;
    // Calls share the scratch buffers: a call on another stream than the
    // previous one waits for it.
    hipEvent_t call_done = nullptr;
    hipStream_t last_stream = nullptr;
    bool call_recorded = false;
    // Captured graphs of the batched kernel sequence.  Kernel arguments hold
    // scratch pointers and the tile table, so every reallocation or table
    // change starts a new generation and the graphs are captured again.
    uint64_t generation = 0;
    struct Graphs {
        uint64_t gen = ~0ull;
        hipGraphExec_t all = nullptr, prep = nullptr, fwd = nullptr, rescue = nullptr;
    } graphs[2];  // [0]: tuned table, [1]: small-batch table
    hipStream_t cap_stream = nullptr;
    int64_t graph_captures = 0, graph_launches = 0;

    // host pipeline: slots of pinned staging + device copies
    struct Slot {
        PinBuf pin;
        DevBuf dev, scores;
        PinBuf pin_scores;
        hipEvent_t h2d_done = nullptr, compute_done = nullptr;
        bool busy = false;
        int64_t n_scores = 0;
    };
    Slot slots[2];

    ~Engine();
    int set_error(int code, const std::string& m)
    {
        err = m;
        return fail(code, m);
    }
    int check(hipError_t e, const char* what)
    {
        if (e == hipSuccess)
            return PHMM_OK;
        hipGetLastError();
        return set_error(e == hipErrorOutOfMemory ? PHMM_ENOMEM : PHMM_EDEVICE,
                         std::string(what) + ": " + hipGetErrorString(e));
    }
    int ensure(DevBuf& b, size_t bytes, const char* what);
    int ensure_pin(PinBuf& b, size_t bytes, const char* what);
    int ensure_zeroed(DevBuf& b, size_t bytes, const char* what);
    int score_device(const phmm_batch_view& v, float* d_scores, hipStream_t st);
    // The three stages of a call: binning, FP32 forward kernels, FP64
    // rescue.  `plain`: sizes read back to the host and exact launches;
    // else persistent grids that read their ranges from device memory.
    int enqueue_prep(const BinParams& bp, const BatchDesc& hd, bool fixed, hipStream_t st);
    int enqueue_fwd(const FwdParams& fp, const BinParams& bp, bool plain, hipStream_t st);
    int enqueue_rescue(const FwdParams& fp, const BinParams& bp, bool plain, hipStream_t st);
    enum { kStagePrep = 1, kStageFwd = 2, kStageRescue = 4 };
    int capture(int stages, const BinParams& bp, const FwdParams& fp, const BatchDesc& hd,
                hipGraphExec_t* out);
    void drop_graphs();
    int score_batch(const phmm_batch_view& v, float* out);
    int finish_timing();
    int take_error_flag();
    // Stage groups [g0, g1) of host view v into slot s as a self-contained
    // batch and enqueue copy-in, scoring and copy-out; s.compute_done fires
    // when s.pin_scores holds the scores.
    int launch_chunk(Slot& s, const phmm_batch_view& v, int g0, int g1);
    // Groups [g0, g1) of several views, concatenated in order into one slot.
    struct ChunkPart {
        const phmm_batch_view* v;
        int g0, g1;
    };
    int launch_parts(Slot& s, const ChunkPart* parts, int nparts);
    void free_slot(Slot& s);
};

extern const phmm_tile_config kDefaultTiles[];
extern const int kNumDefaultTiles;

}  // namespace phmm

struct phmm_engine;

namespace phmm {
// The engine behind a C handle and the mutex that serializes its users.
Engine* engine_of(phmm_engine* h, std::mutex** mu);

}  // namespace phmm
