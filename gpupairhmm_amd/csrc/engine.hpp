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
    hipStream_t stream2 = nullptr;    // compute, host batches' alternate pieces
    // the batch size that picks the tile table when >= 0 (the pieces of a
    // host chunk score on the table of the whole chunk)
    int64_t table_hint = -1;
    // launch the kernels one by one even in graph mode: a graph launched
    // behind a wait on another stream's event (a host piece's copy) holds
    // the host thread in hipGraphLaunch until the copy has landed (1.5-2 ms
    // per piece measured), which would stall the staging of the next piece
    // (set for the pieces of a pipelined host batch)
    bool eager_hint = false;
    std::vector<phmm_tile_config> tiles;
    // The one-template table of small batches (options.small_batch): the
    // classes of `tiles`, 16 rows per lane each (where 64 lanes hold them).
    std::vector<phmm_tile_config> tiles_small;
    const std::vector<phmm_tile_config>* act = &tiles;  // the current call's table
    // per-position gcp: the tile kernels with a gap continuation per row
    bool roweps() const { return opt.semantics == PHMM_SEM_GCP; }
    void derive_small_tiles();
    float* d_ph2pr = nullptr;
    int num_cus = 256;
    std::string err;
    phmm_stats stats{};

    // Batch-sized scratch of a call, in this order (grow-only buffers, or
    // carved from a caller's arena by the allocation-free entry point).
    enum ScratchBuf {
        kRows, kHapc, kMeta, kGbCount, kGbStart, kReadPerm, kAlignOff, kKeyCount, kKeyCursor,
        kBinRange, kItems, kAlignRange, kSuperTasks, kRescueItems, kFinalItems, kHapFlag,
        kRescueExps, kNumScratch
    };
    // An arena must leave at least this much for the spill arena (one block
    // of the FP64 wave kernel holds 32 bytes per column of the longest hap).
    static constexpr size_t kMinArenaSpill = (size_t)8 << 20;
    static constexpr int kDescRing = 16;
    struct Graphs {
        uint64_t gen = ~0ull;
        hipGraphExec_t all = nullptr, prep = nullptr, fwd = nullptr, rescue = nullptr;
    };
    // The device state of a call: batch-sized scratch, counters, the batch
    // descriptor and the captured graphs that hold their addresses.  Calls of
    // one context share it and run one after another (a call on another
    // stream than the previous one waits for it); calls of different
    // contexts may overlap on the device.  Context 0 serves the C ABI's
    // score calls; the serving loop drives further ones on its own streams.
    struct CallCtx {
        DevBuf sbuf[kNumScratch];
        DevBuf spill;                 // border rows of the one-wave kernels
        size_t spill_bytes = (size_t)256 << 20;
        // fixed-size state, allocated by init_ctx: counters (zeroed) and the
        // device copy of the batch descriptor, staged through a ring of
        // pinned slots (a slot is rewritten only after its copy ran)
        DevBuf smallints;
        DevBuf desc;
        BatchDesc* h_desc = nullptr;
        hipEvent_t desc_ev[kDescRing] = {};
        bool desc_used[kDescRing] = {};
        int desc_next = 0;
        int32_t* h_bin_range = nullptr;  // pinned: 2*kMaxWorkBins work ranges, then 2*kMaxBins
                                         // alignment ranges (plain mode)
        hipEvent_t call_done = nullptr;
        hipStream_t last_stream = nullptr;
        bool call_recorded = false;
        bool ready = false;
        // Kernel arguments hold scratch pointers and the tile table, so every
        // reallocation or table change starts a new generation and the graphs
        // are captured again.
        uint64_t generation = 0;
        Graphs graphs[2];  // [0]: tuned table, [1]: small-batch table
        // the current call's spill arena (spill, or the rest of an arena)
        void* call_spill = nullptr;
        size_t call_spill_bytes = 0;
    };
    static constexpr int kMaxCtx = 4;
    CallCtx ctxs[kMaxCtx];
    bool claimed[kMaxCtx] = {true};  // context 0: the C ABI's score calls
    // A free context (>= 1) for a serving loop, or -1; release returns it.
    int claim_ctx()
    {
        for (int k = 1; k < kMaxCtx; ++k)
            if (!claimed[k]) {
                claimed[k] = true;
                return k;
            }
        return -1;
    }
    void release_ctx(int k)
    {
        if (k > 0 && k < kMaxCtx)
            claimed[k] = false;
    }
    CallCtx* cur = &ctxs[0];  // the context of the call being enqueued
    int32_t* h_small = nullptr;       // pinned, small readbacks
    // events around the forward and rescue stages of timed calls
    struct StageEvents {
        hipEvent_t e[3];  // before the forward stage, after it, after the rescue
    };
    std::vector<StageEvents> ev_free, ev_pending;

    // cumulative [rescued, of them in FP64, finished by in-pass scaling]
    int64_t* d_rescued_total = nullptr;
    static constexpr int kTotals = 3;

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
        int ctx = 0;                   // call context and stream of the slot's piece
        hipStream_t st = nullptr;
        phmm_batch_view dv{};          // the staged batch on the device
        int64_t table_n = -1;
        bool eager = false;            // launch its kernels one by one (a pipelined piece)
    };
    static constexpr int kSlots = 3;
    Slot slots[kSlots];

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
    // Grows b to at least `bytes`; a reallocation starts a new generation
    // of context c (its captured graphs hold the old address).
    int ensure(CallCtx& c, DevBuf& b, size_t bytes, const char* what);
    // Grows b to at least `bytes` (geometrically; contents not kept).
    int grow(DevBuf& b, size_t bytes, const char* what);
    int ensure_pin(PinBuf& b, size_t bytes, const char* what);
    // Allocates context k's fixed-size state (counters, descriptor slot and
    // ring, events) once; context 0 at engine creation.
    int init_ctx(int k);
    // Bytes of each batch-sized scratch buffer for view v with NB read bins
    // and NW work bins, and the (NB, NW) of the table a call of v uses.
    static void scratch_sizes(const phmm_batch_view& v, int NB, int NW, size_t out[kNumScratch]);
    // A stream on a hardware queue of its own (a CU mask naming every CU):
    // streams otherwise share the process's few queues, where kernels of
    // different streams run one after another.  Default priority.
    int create_queue_stream(hipStream_t* out);
    // options.range_rescue 1: the FP32 tiles scale per lane in the pass
    // itself; 2: a second, range-extended FP32 pass over the rescue lists
    bool scaled_pass() const { return opt.range_rescue == 1 && opt.fp64_rescue && !roweps(); }
    bool range_rescue() const { return opt.range_rescue == 2 && opt.fp64_rescue && !roweps(); }
    void bins_for(const phmm_batch_view& v, int* NB, int* NW) const;
    // Device bytes context k holds (scratch, spill, counters).
    size_t ctx_bytes(int k) const;
    // Grows context k's scratch to hold a call of v (spill included).
    int reserve_ctx(int k, const phmm_batch_view& v);
    // arena != null: every batch-sized buffer is carved from the arena_bytes
    // at arena (no allocation; the kernels launch one by one, untimed).
    // ctx: the call context (0 for the C ABI's score calls).
    int score_device(const phmm_batch_view& v, float* d_scores, hipStream_t st,
                     char* arena = nullptr, size_t arena_bytes = 0, int ctx = 0);
    // The three stages of a call: binning, FP32 forward kernels, FP64
    // rescue.  `plain`: sizes read back to the host and exact launches;
    // else persistent grids that read their ranges from device memory.
    int enqueue_prep(const BinParams& bp, const BatchDesc& hd, bool fixed, hipStream_t st);
    int enqueue_fwd(const FwdParams& fp, const BinParams& bp, bool plain, hipStream_t st);
    int enqueue_rescue(const FwdParams& fp, const BinParams& bp, bool plain, hipStream_t st);
    enum { kStagePrep = 1, kStageFwd = 2, kStageRescue = 4 };
    int capture(int stages, const BinParams& bp, const FwdParams& fp, const BatchDesc& hd,
                hipGraphExec_t* out);
    void drop_graphs(CallCtx& c);
    void free_ctx(CallCtx& c);
    int score_batch(const phmm_batch_view& v, float* out);
    int finish_timing();
    int take_error_flag();
    // Groups [g0, g1) of several views, concatenated in order into one slot.
    struct ChunkPart {
        const phmm_batch_view* v;
        int g0, g1;
    };
    // Stage the parts into slot s as a self-contained batch and enqueue the
    // copy-in (s.h2d_done fires when it landed); table_n: the batch size that
    // picks the tile table (-1: the staged batch's own).
    int stage_parts(Slot& s, const ChunkPart* parts, int nparts, int64_t table_n);
    // Enqueue the scoring of a staged slot (its context and stream, behind
    // the copy) and the copy-out (s.compute_done fires when s.pin_scores
    // holds the scores).
    int compute_slot(Slot& s);
    void free_slot(Slot& s);
};

extern const phmm_tile_config kDefaultTiles[];
extern const int kNumDefaultTiles;

}  // namespace phmm

struct phmm_engine;

namespace phmm {
// The engine behind a C handle and the mutex that serializes its users.
Engine* engine_of(phmm_engine* h, std::mutex** mu);

// Set-up of the allocation-free entry point (noalloc.hip), once per device
// with that device current: looks up the device address of its scratch.
int noalloc_init(int device);

}  // namespace phmm
