// The allocation-free entry point, phmm_score_device_noalloc: the batch-sized
// scratch of a call (packed read rows, hap codes, bins, work lists, rescue
// lists, the one-wave kernels' spill arena) comes from `scratch`, a
// file-scope device array of PHMM_NOALLOC_SCRATCH_BYTES that every GPU the
// library runs on holds until the process ends.  The engine's fixed-size
// state (counters, descriptor slot, phred table) was allocated when the
// engine was created.  Such a call runs the same kernels as
// phmm_score_device, launched one by one (no graph capture, no kernel
// timing), so its scores are bit-identical to the batched or plain call's.
//
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (two engines, or two streams of one engine's callers) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.hpp"
#include "engine.hpp"

static __device__ __align__(256) unsigned char scratch[PHMM_NOALLOC_SCRATCH_BYTES];

namespace phmm {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char* g_scratch[kMaxDevices] = {};  // device address of `scratch` per GPU

}  // namespace

// Called by phmm_engine_create with `device` current.  In host code the
// array's name is only a host-side stand-in: kernels and copies must use the
// device address kept here.
int noalloc_init(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(PHMM_EINVAL, "noalloc_init: device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device])
        return PHMM_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return fail(PHMM_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                      hipGetErrorString(e));
    }
    g_scratch[device] = static_cast<char*>(p);
    return PHMM_OK;
}

}  // namespace phmm

extern "C" int phmm_score_device_noalloc(phmm_engine* h, const phmm_batch_view* b,
                                         float* d_scores, void* hip_stream)
{
    if (!h || !b)
        return phmm::fail(PHMM_EINVAL, "phmm_score_device_noalloc: null argument");
    std::mutex* mu = nullptr;
    phmm::Engine* e = phmm::engine_of(h, &mu);
    std::lock_guard<std::mutex> lk(*mu);
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> g(phmm::g_mu);
        base = e->device >= 0 && e->device < phmm::kMaxDevices ? phmm::g_scratch[e->device]
                                                                : nullptr;
    }
    if (!base)
        return e->set_error(PHMM_EDEVICE,
                            "phmm_score_device_noalloc: no scratch address for the engine's device");
    if (hipSetDevice(e->device) != hipSuccess) {
        (void)hipGetLastError();
        return e->set_error(PHMM_EDEVICE, "phmm_score_device_noalloc: hipSetDevice failed");
    }
    hipStream_t st = hip_stream ? static_cast<hipStream_t>(hip_stream) : e->stream;
    return e->score_device(*b, d_scores, st, base, PHMM_NOALLOC_SCRATCH_BYTES);
}
