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
    float* d_ph2pr = nullptr;
    int num_cus = 256;
    std::string err;
    phmm_stats stats{};

    // scratch (grow-only)
    DevBuf rows, hapc, meta, gb_count, gb_start, read_perm, align_off, key_count, key_cursor,
        bin_range, smallints, items, rescue_list, final_list, spill;
    int32_t* h_bin_range = nullptr;   // pinned, 2*NB
    int32_t* h_small = nullptr;       // pinned, small readbacks
    // event pairs around the forward kernels of timed calls
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_free, ev_pending;

    int64_t* d_rescued_total = nullptr;  // cumulative FP64 rescues

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
    int score_batch(const phmm_batch_view& v, float* out);
    int finish_timing();
    int take_error_flag();
    // Stage groups [g0, g1) of host view v into slot s as a self-contained
    // batch and enqueue copy-in, scoring and copy-out; s.compute_done fires
    // when s.pin_scores holds the scores.
    int launch_chunk(Slot& s, const phmm_batch_view& v, int g0, int g1);
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
