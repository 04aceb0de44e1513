// The engine: device memory, work-list construction, kernel dispatch, FP64
// rescue, the chunked host pipeline, and the C ABI around them.
//
// Mirrors GpuPairHMMWorker (gpuworker.cuh:381-973): the constructor uploads
// the phred table and creates streams; processBatch... becomes
// Engine::score_batch (chunking by transfer_batchsize at record granularity,
// copies overlapped with compute) on top of Engine::score_device (the device
// part: binning, per-bin forward kernels, rescue).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "engine.hpp"

namespace phmm {

// Default tile table for gfx950, emitted by gpupairhmm_amd/autotune.py.
#include "tuning_gfx950.inc"

static thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}
const char* last_error() { return g_last_error.c_str(); }

// Reads and clears the device error flags (one per call context) the
// multi-tile and rescue kernels raise when the spill arena cannot hold one
// alignment.
int Engine::take_error_flag()
{
    int32_t flag = 0;
    for (CallCtx& c : ctxs) {
        if (!c.smallints.p)
            continue;
        int32_t* d_flag = static_cast<int32_t*>(c.smallints.p) + 2;
        int32_t f = 0;
        int rc;
        if ((rc = check(hipMemcpy(&f, d_flag, sizeof f, hipMemcpyDeviceToHost), "error flag")))
            return rc;
        if (f) {
            const int32_t zero = 0;
            hipMemcpy(d_flag, &zero, sizeof zero, hipMemcpyHostToDevice);
            flag |= f;
        }
    }
    if (!flag)
        return PHMM_OK;
    return set_error(PHMM_ERANGE, "multi-tile/rescue spill arena too small for a batch (flag " +
                                      std::to_string(flag) + ")");
}

Engine::~Engine()
{
    hipSetDevice(device);
    if (stream)
        hipStreamSynchronize(stream);
    if (stream2)
        hipStreamSynchronize(stream2);
    if (copy_stream)
        hipStreamSynchronize(copy_stream);
    for (auto& s : slots)
        free_slot(s);
    for (CallCtx& c : ctxs)
        free_ctx(c);
    if (cap_stream)
        hipStreamDestroy(cap_stream);
    if (d_rescued_total)
        hipFree(d_rescued_total);
    if (h_small)
        hipHostFree(h_small);
    if (d_ph2pr)
        hipFree(d_ph2pr);
    for (auto* v : {&ev_free, &ev_pending})
        for (auto& ev : *v)
            for (hipEvent_t e : ev.e)
                hipEventDestroy(e);
    if (stream)
        hipStreamDestroy(stream);
    if (copy_stream)
        hipStreamDestroy(copy_stream);
    if (stream2)
        hipStreamDestroy(stream2);
}

int Engine::create_queue_stream(hipStream_t* out)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0) {
        std::vector<uint32_t> mask((size_t)(prop.multiProcessorCount + 31) / 32, 0xffffffffu);
        if (hipExtStreamCreateWithCUMask(out, (uint32_t)mask.size(), mask.data()) == hipSuccess)
            return PHMM_OK;
    }
    (void)hipGetLastError();
    return check(hipStreamCreateWithFlags(out, hipStreamNonBlocking), "stream");
}

void Engine::free_ctx(CallCtx& c)
{
    drop_graphs(c);
    for (DevBuf& b : c.sbuf)
        if (b.p)
            hipFree(b.p);
    for (DevBuf* b : {&c.smallints, &c.spill, &c.desc})
        if (b->p)
            hipFree(b->p);
    if (c.h_desc)
        hipHostFree(c.h_desc);
    if (c.h_bin_range)
        hipHostFree(c.h_bin_range);
    for (auto& ev : c.desc_ev)
        if (ev)
            hipEventDestroy(ev);
    if (c.call_done)
        hipEventDestroy(c.call_done);
    c = CallCtx{};
}

int Engine::ensure(CallCtx& c, DevBuf& b, size_t bytes, const char* what)
{
    if (bytes <= b.cap && b.p)
        return PHMM_OK;
    ++c.generation;
    return grow(b, bytes, what);
}

int Engine::grow(DevBuf& b, size_t bytes, const char* what)
{
    if (bytes <= b.cap && b.p)
        return PHMM_OK;
    size_t nb = std::max<size_t>(bytes, 256);
    if (b.p) {
        hipDeviceSynchronize();  // in-flight kernels may still use it
        hipFree(b.p);
        nb = std::max(nb, b.cap + b.cap / 2);  // grow geometrically
        b.p = nullptr;
        b.cap = 0;
    }
    nb = (nb + 255) & ~(size_t)255;
    if (hipMalloc(&b.p, nb) != hipSuccess) {
        hipGetLastError();
        b.p = nullptr;
        return set_error(PHMM_ENOMEM, std::string("device allocation failed: ") + what + " (" +
                                          std::to_string(nb) + " bytes)");
    }
    b.cap = nb;
    return PHMM_OK;
}

int Engine::init_ctx(int k)
{
    if (k < 0 || k >= kMaxCtx)
        return set_error(PHMM_EINVAL, "call context out of range");
    CallCtx& c = ctxs[k];
    if (c.ready)
        return PHMM_OK;
    int rc;
    if ((rc = ensure(c, c.smallints, sizeof(int32_t) * (256 + kHmaxParts + kRetrySlots),
                     "counters")) ||
        (rc = check(hipMemset(c.smallints.p, 0, c.smallints.cap), "memset")) ||
        (rc = ensure(c, c.desc, sizeof(BatchDesc), "batch descriptor")) ||
        (rc = check(hipHostMalloc(reinterpret_cast<void**>(&c.h_desc),
                                  sizeof(BatchDesc) * kDescRing, hipHostMallocDefault),
                    "pinned descriptors")) ||
        (rc = check(hipHostMalloc(reinterpret_cast<void**>(&c.h_bin_range),
                                  2 * sizeof(int32_t) * (kMaxWorkBins + kMaxBins),
                                  hipHostMallocDefault),
                    "pinned")))
        return rc;
    for (int i = 0; i < kDescRing; ++i)
        if ((rc = check(hipEventCreateWithFlags(&c.desc_ev[i], hipEventDisableTiming), "event")))
            return rc;
    if ((rc = check(hipEventCreateWithFlags(&c.call_done, hipEventDisableTiming), "event")))
        return rc;
    c.ready = true;
    return PHMM_OK;
}

void Engine::bins_for(const phmm_batch_view& v, int* NB, int* NW) const
{
    // the table a call of v uses (see score_device)
    const bool small =
        roweps() || (opt.small_batch > 0 && v.num_alignments <= opt.small_batch);
    const int ntiles = (int)(small ? tiles_small : tiles).size();
    *NB = 2 * ntiles + 2;
    *NW = *NB + 2 * ntiles;
}

size_t Engine::ctx_bytes(int k) const
{
    const CallCtx& c = ctxs[k];
    size_t n = c.spill.cap + c.smallints.cap + c.desc.cap;
    for (const DevBuf& b : c.sbuf)
        n += b.cap;
    return n;
}

int Engine::reserve_ctx(int k, const phmm_batch_view& v)
{
    int rc;
    if ((rc = init_ctx(k)))
        return rc;
    CallCtx& c = ctxs[k];
    int NB, NW;
    bins_for(v, &NB, &NW);
    size_t need[kNumScratch];
    scratch_sizes(v, NB, NW, need);
    for (int i = 0; i < kNumScratch; ++i)
        if ((rc = ensure(c, c.sbuf[i], need[i], "scratch")))
            return rc;
    return ensure(c, c.spill, c.spill_bytes, "spill");
}

void Engine::scratch_sizes(const phmm_batch_view& v, int NB, int NW, size_t out[kNumScratch])
{
    const size_t G = (size_t)v.num_groups, nA = (size_t)std::max<int64_t>(v.num_alignments, 1);
    out[kRows] = sizeof(uint32_t) * (size_t)std::max<int64_t>(v.read_bytes, 1);
    out[kHapc] = (size_t)std::max<int64_t>(v.hap_bytes, 4) + 16;
    out[kMeta] = sizeof(ReadMeta) * (size_t)std::max(v.num_reads, 1);
    out[kGbCount] = sizeof(int32_t) * G * (size_t)(NB + 1);
    out[kGbStart] = out[kGbCount];
    out[kReadPerm] = sizeof(int32_t) * (size_t)std::max(v.num_reads, 1);
    out[kAlignOff] = sizeof(int64_t) * (G + 1);
    out[kKeyCount] = sizeof(int32_t) * (size_t)NW * kKeysPerBin;
    out[kKeyCursor] = out[kKeyCount];
    out[kBinRange] = sizeof(int32_t) * 2 * (size_t)NW;
    out[kItems] = sizeof(WorkItem) * nA;
    out[kAlignRange] = sizeof(int32_t) * (2 + kAlignParts) * (size_t)NB;
    // each super-task covers at least one (read, hap) alignment and no two
    // cover the same one: at most num_alignments of them
    out[kSuperTasks] = sizeof(SuperTask) * nA;
    out[kRescueItems] = sizeof(WorkItem) * nA;
    out[kFinalItems] = sizeof(WorkItem) * nA;
    out[kHapFlag] = (size_t)std::max(v.num_haps, 1);
    // per-lane magnitudes of rescued alignments: 16 bytes per rescue-list slot
    out[kRescueExps] = (size_t)kReKnownLanes * nA;
}

int Engine::ensure_pin(PinBuf& b, size_t bytes, const char* what)
{
    if (bytes <= b.cap && b.p)
        return PHMM_OK;
    if (b.p) {
        hipHostFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t nb = std::max<size_t>(bytes + bytes / 4, 4096);
    if (hipHostMalloc(&b.p, nb, hipHostMallocDefault) != hipSuccess) {
        hipGetLastError();
        b.p = nullptr;
        return set_error(PHMM_ENOMEM, std::string("pinned allocation failed: ") + what);
    }
    b.cap = nb;
    return PHMM_OK;
}

// Adds the elapsed times of every pending event triple (waits for them):
// forward stage e[0] -> e[1], FP64 rescue stage e[1] -> e[2].
int Engine::finish_timing()
{
    for (auto& ev : ev_pending) {
        float ms = 0.f, ms2 = 0.f;
        if (hipEventSynchronize(ev.e[2]) == hipSuccess &&
            hipEventElapsedTime(&ms, ev.e[0], ev.e[1]) == hipSuccess &&
            hipEventElapsedTime(&ms2, ev.e[1], ev.e[2]) == hipSuccess) {
            stats.kernel_seconds += ms * 1e-3;
            stats.rescue_seconds += ms2 * 1e-3;
        } else {
            hipGetLastError();
        }
        ev_free.push_back(ev);
    }
    ev_pending.clear();
    return PHMM_OK;
}

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// True when p lies in page-locked host memory (hipHostMalloc'd), which a
// copy engine reads directly.
static bool host_pinned(const void* p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

// The device part of a batch.  `v` holds device pointers (counts on the
// host) and must be self-contained (reads_prefix[0] == haps_prefix[0] == 0).
int Engine::score_device(const phmm_batch_view& v, float* d_scores, hipStream_t st, char* arena,
                         size_t arena_bytes, int ctx)
{
    if (v.num_groups < 0 || v.num_reads < 0 || v.num_haps < 0 || v.num_alignments < 0)
        return set_error(PHMM_EINVAL, "score_device: negative counts");
    if (v.num_alignments > INT32_MAX)
        return set_error(PHMM_ERANGE, "score_device: more than 2^31-1 alignments in one batch");
    if (v.num_alignments == 0 || v.num_groups == 0)
        return PHMM_OK;
    if (!d_scores || !v.read_lengths || !v.hap_lengths || !v.read_offsets || !v.hap_offsets ||
        !v.num_reads_per_group || !v.num_haps_per_group || !v.reads_prefix || !v.haps_prefix ||
        (v.num_reads > 0 && (!v.read_data || !v.base_quals || !v.ins_quals || !v.del_quals ||
                             !v.gcp_quals)) ||
        (v.num_haps > 0 && !v.hap_data))
        return set_error(PHMM_EINVAL, "score_device: null device pointer in batch view");
    // the encoders read sequence and quality bytes as dwords: every sequence
    // is padded to a multiple of 4 bytes (4 - len % 4, as the parser does)
    auto misaligned = [](const void* q) { return ((uintptr_t)q & 3) != 0; };
    if ((v.read_bytes & 3) || (v.hap_bytes & 3) || misaligned(v.read_data) ||
        misaligned(v.base_quals) || misaligned(v.ins_quals) || misaligned(v.del_quals) ||
        misaligned(v.hap_data) || (opt.semantics == PHMM_SEM_GCP && misaligned(v.gcp_quals)))
        return set_error(PHMM_EINVAL,
                         "score_device: sequence arrays must be 4-byte aligned with padded lengths "
                         "(byte counts multiples of 4)");
    // batched calls of small batches use the one-template table (plain mode
    // keeps the per-class launches of the tuned table); per-position gcp
    // always does: its tile kernels (a gap continuation register per row)
    // exist for 16 rows per lane
    const int64_t table_n = table_hint >= 0 ? table_hint : v.num_alignments;
    act = (roweps() || (opt.small_batch > 0 && table_n <= opt.small_batch)) ? &tiles_small : &tiles;
    const std::vector<phmm_tile_config>& tiles = *act;
    const int ntiles = (int)tiles.size();
    const int NB = 2 * ntiles + 2;  // + multi-tile bin + exact bin
    const int NW = NB + 2 * ntiles; // + an item bin per tile bin
    const int G = v.num_groups;
    const int64_t nA = v.num_alignments;
    int rc;
    if ((rc = init_ctx(ctx)))
        return rc;
    CallCtx& c = ctxs[ctx];
    cur = &c;
    static const char* const kScratchName[kNumScratch] = {
        "rows", "hap codes", "read meta", "group bins", "group bins", "read perm",
        "align offsets", "keys", "keys", "bin ranges", "work items", "alignment ranges",
        "super-tasks", "rescue list", "final list", "hap flags", "rescue maxima"};
    size_t need[kNumScratch];
    scratch_sizes(v, NB, NW, need);
    void* sp[kNumScratch];
    if (arena) {
        // the allocation-free form: every buffer at a 256-byte boundary of
        // the arena, the spill arena takes the rest
        size_t o = 0;
        for (int i = 0; i < kNumScratch; ++i) {
            sp[i] = arena + o;
            o = align256(o + need[i]);
        }
        if (o + kMinArenaSpill > arena_bytes)
            return set_error(PHMM_ERANGE,
                             "score_device: the batch needs " + std::to_string(o + kMinArenaSpill) +
                                 " bytes of device scratch; the arena holds " +
                                 std::to_string(arena_bytes));
        c.call_spill = arena + o;
        c.call_spill_bytes = arena_bytes - o;
    } else {
        for (int i = 0; i < kNumScratch; ++i) {
            if ((rc = ensure(c, c.sbuf[i], need[i], kScratchName[i])))
                return rc;
            sp[i] = c.sbuf[i].p;
        }
        if ((rc = ensure(c, c.spill, c.spill_bytes, "spill")))
            return rc;
        c.call_spill = c.spill.p;
        c.call_spill_bytes = c.spill.cap;
    }
    if (c.call_recorded && c.last_stream != st &&
        (rc = check(hipStreamWaitEvent(st, c.call_done, 0), "event wait")))
        return rc;

    BatchDesc hd{};
    hd.read_data = v.read_data;
    hd.hap_data = v.hap_data;
    hd.bq = v.base_quals;
    hd.iq = v.ins_quals;
    hd.dq = v.del_quals;
    hd.gcp = v.gcp_quals;
    hd.read_len = v.read_lengths;
    hd.hap_len = v.hap_lengths;
    hd.read_off = v.read_offsets;
    hd.hap_off = v.hap_offsets;
    hd.nreads = v.num_reads_per_group;
    hd.nhaps = v.num_haps_per_group;
    hd.rprefix = v.reads_prefix;
    hd.hprefix = v.haps_prefix;
    hd.scores = d_scores;
    hd.G = G;
    hd.nR = v.num_reads;
    hd.nH = v.num_haps;
    // super-task segments: about 32k tasks for the batch (16 per resident
    // wave), so long or few haps fall back to one hap per task
    {
        const double mean_hl = v.num_haps > 0 ? (double)v.hap_bytes / v.num_haps : 0.0;
        const double cols = (double)nA * mean_hl;
        hd.col_budget = (int32_t)std::min(8192.0, std::max(256.0, cols / (8.0 * 32768.0)));
    }
    hd.read_bytes = v.read_bytes;
    hd.hap_bytes = v.hap_bytes;
    hd.num_alignments = nA;
    {
        const int k = c.desc_next;
        c.desc_next = (c.desc_next + 1) % kDescRing;
        if (c.desc_used[k] && (rc = check(hipEventSynchronize(c.desc_ev[k]), "descriptor slot")))
            return rc;
        c.h_desc[k] = hd;
        if ((rc = check(hipMemcpyAsync(c.desc.p, &c.h_desc[k], sizeof(BatchDesc),
                                       hipMemcpyHostToDevice, st),
                        "descriptor copy")) ||
            (rc = check(hipEventRecord(c.desc_ev[k], st), "event")))
            return rc;
        c.desc_used[k] = true;
    }

    // counters: [0] max hap length, [1] final-list count, [2] error flag,
    // [4 .. 4+NB) per-bin rescue counts, [128 .. 128+kTaskCounters) task
    // counters of the set launches, [256 .. 256+kHmaxParts) partial maxima,
    // then kRetrySlots: per-bin retry counts and (last) the exps unit counter
    static_assert(4 + kMaxBins <= 128 && 128 + kTaskCounters <= 256 && kMaxBins < kRetrySlots,
                  "counter layout");
    int32_t* sm = static_cast<int32_t*>(c.smallints.p);
    BinParams bp{};
    bp.bd = static_cast<const BatchDesc*>(c.desc.p);
    bp.NB = NB;
    bp.NW = NW;
    bp.rows = static_cast<uint32_t*>(sp[kRows]);
    bp.hapc = static_cast<uint8_t*>(sp[kHapc]);
    bp.meta = static_cast<ReadMeta*>(sp[kMeta]);
    bp.gb_count = static_cast<int32_t*>(sp[kGbCount]);
    bp.gb_start = static_cast<int32_t*>(sp[kGbStart]);
    bp.read_perm = static_cast<int32_t*>(sp[kReadPerm]);
    bp.align_off = static_cast<int64_t*>(sp[kAlignOff]);
    bp.key_count = static_cast<int32_t*>(sp[kKeyCount]);
    bp.key_cursor = static_cast<int32_t*>(sp[kKeyCursor]);
    bp.bin_range = static_cast<int32_t*>(sp[kBinRange]);
    bp.hmax = sm + 0;
    bp.hmax_part = sm + 256;
    bp.final_count = sm + 1;
    bp.error_flag = sm + 2;
    bp.rescue_counts = sm + 4;
    bp.retry_counts = sm + 256 + kHmaxParts;
    bp.task_ctr = sm + 128;
    bp.hapflag = static_cast<uint8_t*>(sp[kHapFlag]);
    bp.items = static_cast<WorkItem*>(sp[kItems]);
    bp.st = static_cast<SuperTask*>(sp[kSuperTasks]);
    bp.align_range = static_cast<int32_t*>(sp[kAlignRange]);
    bp.align_part = bp.align_range + 2 * NB;
    bp.ph2pr = d_ph2pr;
    bp.ntiles = ntiles;
    for (int i = 0; i < ntiles; ++i) {
        bp.tiles[i] = tiles[(size_t)i].max_read_len;
        bp.lanes[i] = tiles[(size_t)i].lanes;
    }
    bp.semantics = opt.semantics;
    bp.allow_fast = 1;

    FwdParams fp{};
    fp.bd = bp.bd;
    fp.rows = bp.rows;
    fp.hapc = bp.hapc;
    fp.meta = bp.meta;
    fp.read_perm = bp.read_perm;
    fp.hapflag = bp.hapflag;
    fp.ph2pr = d_ph2pr;
    fp.rescue_items = static_cast<WorkItem*>(sp[kRescueItems]);
    fp.rescue_counts = bp.rescue_counts;
    fp.final_items = static_cast<WorkItem*>(sp[kFinalItems]);
    fp.final_count = bp.final_count;
    // the retry lists take the super-task array, which only the forward
    // stage reads (SuperTask is at least as large as WorkItem)
    static_assert(sizeof(SuperTask) >= sizeof(WorkItem), "retry lists in the super-task array");
    fp.retry_items = static_cast<WorkItem*>(sp[kSuperTasks]);
    fp.retry_counts = bp.retry_counts;
    fp.rescue_exps = static_cast<uint8_t*>(sp[kRescueExps]);
    // the range-extended pass pays off where the rescue is throughput-bound
    // (the tuned table); a small batch's rescue costs one sweep's latency per
    // stage, so a second stage would only add one
    fp.re_on = range_rescue() && act != &tiles_small ? 1 : 0;
    fp.scaled_total = reinterpret_cast<unsigned long long*>(d_rescued_total + 2);
    fp.rescue_min = opt.rescue_min_sum;
    fp.rescue_on = opt.fp64_rescue;
    fp.semantics = opt.semantics;

    const bool timed = opt.time_kernels != 0 && !arena;
    StageEvents evp{};
    if (timed) {
        if (ev_free.empty()) {
            for (hipEvent_t& e : evp.e)
                if ((rc = check(hipEventCreate(&e), "event")))
                    return rc;
        } else {
            evp = ev_free.back();
            ev_free.pop_back();
        }
    }
    const bool plain = opt.mode == PHMM_MODE_PLAIN;
    if (plain || !opt.graphs || arena || eager_hint) {
        if ((rc = enqueue_prep(bp, hd, false, st)))
            return rc;
        if (plain) {
            // one host round trip per batch for the per-bin sizes (the
            // reference syncs per chunk for its per-partition counts,
            // gpuworker.cuh:852-854)
            if ((rc = check(hipMemcpyAsync(c.h_bin_range, bp.bin_range, sizeof(int32_t) * 2 * NW,
                                           hipMemcpyDeviceToHost, st),
                            "bin range readback")) ||
                (rc = check(hipMemcpyAsync(c.h_bin_range + 2 * kMaxWorkBins, bp.align_range,
                                           sizeof(int32_t) * 2 * NB, hipMemcpyDeviceToHost, st),
                            "bin range readback")) ||
                (rc = check(hipStreamSynchronize(st), "bin range sync")))
                return rc;
        }
        if (timed)
            hipEventRecord(evp.e[0], st);
        if ((rc = enqueue_fwd(fp, bp, plain, st)))
            return rc;
        if (timed)
            hipEventRecord(evp.e[1], st);
        if (opt.fp64_rescue && (rc = enqueue_rescue(fp, bp, plain, st)))
            return rc;
        if (timed)
            hipEventRecord(evp.e[2], st);
    } else {
        // batched: replay captured graphs -- one for the whole call, or one
        // per stage when the forward kernels are timed
        Graphs& graphs = c.graphs[act == &tiles_small ? 1 : 0];
        if (graphs.gen != c.generation) {
            drop_graphs(c);
            c.graphs[0].gen = c.graphs[1].gen = c.generation;
        }
        auto launch = [&](int stages, hipGraphExec_t* ex) -> int {
            int r2;
            if (!*ex && (r2 = capture(stages, bp, fp, hd, ex)))
                return r2;
            ++graph_launches;
            return check(hipGraphLaunch(*ex, st), "graph launch");
        };
        if (!timed) {
            if ((rc = launch(kStagePrep | kStageFwd | kStageRescue, &graphs.all)))
                return rc;
        } else {
            if ((rc = launch(kStagePrep, &graphs.prep)))
                return rc;
            hipEventRecord(evp.e[0], st);
            if ((rc = launch(kStageFwd, &graphs.fwd)))
                return rc;
            hipEventRecord(evp.e[1], st);
            if (opt.fp64_rescue && (rc = launch(kStageRescue, &graphs.rescue)))
                return rc;
            hipEventRecord(evp.e[2], st);
        }
    }
    if (timed)
        ev_pending.push_back(evp);
    if ((rc = check(hipEventRecord(c.call_done, st), "event")))
        return rc;
    c.last_stream = st;
    c.call_recorded = true;
    stats.alignments += nA;
    return PHMM_OK;
}

int Engine::enqueue_prep(const BinParams& bp, const BatchDesc& hd, bool fixed, hipStream_t st)
{
    return check(run_binning(bp, hd, fixed, act == &tiles_small, st), "binning kernels");
}

// Pass 1: FP32 tile kernels over the tile bins' super-tasks (results below
// the rescue threshold go to the bin's rescue list) and the one-wave kernel
// over the multi-tile and exact bins' items.  Batched: one persistent launch
// per kernel template covering all of its bins, larger tiles first.  Plain:
// one launch per non-empty bin after the host has read the bin sizes back
// (the reference's per-partition launches).
int Engine::enqueue_fwd(const FwdParams& fp, const BinParams& bp, bool plain, hipStream_t st)
{
    const std::vector<phmm_tile_config>& tiles = *act;
    const int ntiles = (int)tiles.size();
    const int mb = 2 * ntiles;  // multi-tile bin
    const int eb = mb + 1;      // exact bin
    float* spill32 = static_cast<float*>(cur->call_spill);
    const int64_t n32 = (int64_t)(cur->call_spill_bytes / sizeof(float));
    const int32_t* h_bin_range = cur->h_bin_range;
    const int32_t* h_align = h_bin_range + 2 * kMaxWorkBins;
    const int NB = bp.NB;  // item bin of tile bin b: NB + b
    int rc;
    int ctr = 0;
    auto tile_set = [&](BinSet& bs) {
        bs.st = bp.st;
        bs.items = bp.items;
        bs.item_bin0 = NB;
        bs.range = bp.bin_range;
        bs.counts = bp.bin_range + 1;
        bs.count_stride = 2;
        bs.seg = bp.align_range;
    };
    if (plain) {
        for (int b = 2 * ntiles - 1; b >= 0; --b) {
            const phmm_tile_config& tc = tiles[(size_t)(b / 2)];
            const int W = 64 / tc.lanes;
            const int tasks = h_bin_range[2 * b + 1] + (h_bin_range[2 * (NB + b) + 1] + W - 1) / W;
            if (tasks <= 0)
                continue;
            if (b % 2 == 0)
                stats.fast_alignments += h_align[2 * b + 1];
            BinSet bs;
            bs.n = 2;
            bs.bin[0] = b;
            bs.bin[1] = NB + b;
            bs.lanes[0] = bs.lanes[1] = tc.lanes;
            tile_set(bs);
            if (ctr >= kTaskCounters - kF64Sets - kReSets)
                return set_error(PHMM_ERANGE, "too many set launches");
            bs.next_task = bp.task_ctr + ctr++;
            const bool scaled = scaled_pass() && tile_re_supported(tc.rows);
            const int grid = std::min(
                tasks, num_cus * tile_set_blocks_per_cu(tc.rows, (b % 2) == 0, false, roweps(),
                                                        scaled));
            FwdParams fpl = fp;
            fpl.re_on = fp.re_on && tile_re_supported(tc.rows);
            if ((rc = check(scaled ? launch_fwd_tile_set_scaled(tc.rows, (b % 2) == 0, fpl, bs,
                                                                grid, st)
                                   : launch_fwd_tile_set(tc.rows, (b % 2) == 0, fpl, bs, grid,
                                                         st, roweps()),
                            "forward kernel")))
                return rc;
        }
    } else {
        // templates (rows, form) in the order of their largest bin
        for (int b0 = 2 * ntiles - 1; b0 >= 0; --b0) {
            const int rows = tiles[(size_t)(b0 / 2)].rows;
            const bool fast = (b0 % 2) == 0;
            bool seen = false;
            for (int b = b0 + 1; b < 2 * ntiles && !seen; ++b)
                seen = tiles[(size_t)(b / 2)].rows == rows && ((b % 2) == 0) == fast;
            if (seen)
                continue;
            // the template's super-task bins, larger tiles first, then their
            // item bins (one sweep per task: they fill the launch's tail)
            BinSet bs;
            for (int pass = 0; pass < 2; ++pass)
                for (int b = b0; b >= 0; --b)
                    if (tiles[(size_t)(b / 2)].rows == rows && ((b % 2) == 0) == fast) {
                        bs.bin[bs.n] = pass == 0 ? b : NB + b;
                        bs.lanes[bs.n] = tiles[(size_t)(b / 2)].lanes;
                        ++bs.n;
                    }
            tile_set(bs);
            if (ctr >= kTaskCounters - kF64Sets - kReSets)
                return set_error(PHMM_ERANGE, "too many set launches");
            bs.next_task = bp.task_ctr + ctr++;
            FwdParams fpl = fp;
            fpl.re_on = fp.re_on && tile_re_supported(rows);
            const bool scaled = scaled_pass() && tile_re_supported(rows);
            if ((rc = check(scaled ? launch_fwd_tile_set_scaled(rows, fast, fpl, bs, 0, st)
                                   : launch_fwd_tile_set(rows, fast, fpl, bs, 0, st, roweps()),
                            "forward kernel")))
                return rc;
        }
    }
    // the multi-tile bin and the exact bin: one wave per alignment
    if (plain) {
        const int n = h_bin_range[2 * mb + 1] + h_bin_range[2 * eb + 1];
        if (n <= 0)
            return PHMM_OK;
        stats.multitile_alignments += n;
    }
    BinSet ws;
    ws.n = 2;
    ws.bin[0] = mb;
    ws.bin[1] = eb;
    ws.lanes[0] = ws.lanes[1] = 64;
    ws.items = bp.items;
    ws.range = bp.bin_range;
    ws.counts = bp.bin_range + 1;
    ws.count_stride = 2;
    const int blocks =
        plain ? std::min(num_cus * 8, h_bin_range[2 * mb + 1] + h_bin_range[2 * eb + 1]) : num_cus * 8;
    return check(launch_fwd_wave_f32(fp, ws, bp.hmax, spill32, n32, bp.error_flag, blocks, st),
                 "multi-tile kernel");
}

// Pass 2 (range_rescue): the rescue lists of tile bins whose rows per lane
// have a range-extended instance are recomputed in FP32 with per-lane
// power-of-two scales, one set launch per (rows, form) template; what that
// pass cannot accept joins the bin's retry list.  Pass 3: FP64 recompute of
// the retry lists (and of the rescue lists of the other bins) with an FP64
// tile of at least the bin's size (tiles above 512 rows: one wave per
// alignment), one launch per FP64 template and list source; pass 4: FP64
// waves for anything still not finished.
int Engine::enqueue_rescue(const FwdParams& fp, const BinParams& bp, bool plain, hipStream_t st)
{
    const std::vector<phmm_tile_config>& tiles = *act;
    const int ntiles = (int)tiles.size();
    double* spill64 = static_cast<double*>(cur->call_spill);
    const int64_t n64 = (int64_t)(cur->call_spill_bytes / sizeof(double));
    const int32_t* h_align = cur->h_bin_range + 2 * kMaxWorkBins;
    int rc;
    auto re_bin = [&](int b) {
        const phmm_tile_config& tc = tiles[(size_t)(b / 2)];
        return fp.re_on != 0 && tile_re_supported(tc.rows);
    };
    auto empty_bin = [&](int b) { return plain && h_align[2 * b + 1] <= 0; };
    // statistics: [0] rescued, [1] of them recomputed in FP64.  The final
    // list holds, at this point, the one-wave FP32 kernel's results that
    // need FP64 (the FP64 tile pass adds to it later).
    int64_t* tot = d_rescued_total;
    {
        CountSums c;
        c.add(tot, fp.final_count, 1);
        c.add(tot + 1, fp.final_count, 1);
        if ((rc = check(launch_add_counts(c, st), "rescue count")))
            return rc;
    }
    if (fp.re_on) {
        int nre = 0;
        for (int b0 = 2 * ntiles - 1; b0 >= 0; --b0) {
            const int rows = tiles[(size_t)(b0 / 2)].rows;
            const bool fast = (b0 % 2) == 0;
            if (!re_bin(b0))
                continue;
            // one launch per template: at its largest range-extended bin
            bool seen = false;
            for (int b = b0 + 1; b < 2 * ntiles && !seen; ++b)
                seen = tiles[(size_t)(b / 2)].rows == rows && ((b % 2) == 0) == fast && re_bin(b);
            if (seen)
                continue;
            BinSet bs;
            for (int b = b0; b >= 0; --b)
                if (tiles[(size_t)(b / 2)].rows == rows && ((b % 2) == 0) == fast && re_bin(b) &&
                    !empty_bin(b)) {
                    bs.bin[bs.n] = b;
                    bs.lanes[bs.n] = tiles[(size_t)(b / 2)].lanes;
                    ++bs.n;
                }
            if (!bs.n)
                continue;
            bs.items = fp.rescue_items;
            bs.range = bp.align_range;
            bs.counts = fp.rescue_counts;
            bs.count_stride = 1;
            bs.seg = bp.align_range;
            if (nre >= kReSets)
                return set_error(PHMM_ERANGE, "too many range-extended set launches");
            bs.next_task = bp.task_ctr + (kTaskCounters - kF64Sets - kReSets + nre++);
            if ((rc = check(launch_fwd_tile_set_re(rows, fast, fp, bs, 0, st),
                            "range-extended kernel")))
                return rc;
        }
    }
    // FP64 tile sets by rows per lane, form and list (a bin's list keeps the
    // bin's form: FAST-eligible reads recompute in the FAST form).  Each bin
    // takes the rows per lane nr whose modelled time per alignment is least:
    // the tile's rows G*nr (G = lanes holding the bin's class) over the
    // measured slot rate of full nr-row tiles (kF64SlotRate), scaled by the
    // lanes a wave leaves idle (64 - (64/G)*G) and the G fill steps of a
    // 300-column hap.  The 10-row FAST tile (248 VGPRs) still runs two waves
    // per SIMD; the 10-row SAFE one would not, so SAFE stops at 8.  The
    // small-batch table takes the widest tile of each form for every bin: a
    // small batch's FP64 time is the tail of each set launch, so one set per
    // form beats a set per cost choice.  options.f64_rows forces one
    // rows-per-lane choice (the autotuner's measurement of each).
    static constexpr int kNr64[] = {4, 6, 8, 10};
    const bool one_set = act == &tiles_small;
    // per-position gcp: the 8-row FP64 tiles in both forms
    const int one_set_fast = roweps() ? kRowEpsRows64 : 10;
    BinSet set[4][2][2], wave[2];  // [rows index][SAFE, FAST][rescue list, retry list]
    for (int src = 0; src < 2; ++src) {
        for (auto& row : set)
            for (auto& form : row) {
                BinSet& x = form[src];
                x.items = src ? fp.retry_items : fp.rescue_items;
                x.range = bp.align_range;
                x.counts = src ? fp.retry_counts : fp.rescue_counts;
                x.count_stride = 1;
            }
        wave[src].items = src ? fp.retry_items : fp.rescue_items;
        wave[src].range = bp.align_range;
        wave[src].counts = src ? fp.retry_counts : fp.rescue_counts;
        wave[src].count_stride = 1;
    }
    for (int b = 2 * ntiles - 1; b >= 0; --b) {
        if (empty_bin(b))
            continue;
        const int src = re_bin(b) ? 1 : 0;
        const int tile = tiles[(size_t)(b / 2)].max_read_len + 1;  // rows the FP64 tile must hold
        const bool fast = b % 2 == 0;
        if (tile > 512) {
            wave[src].bin[wave[src].n] = b;
            wave[src].lanes[wave[src].n] = 64;
            ++wave[src].n;
            continue;
        }
        int best = -1;
        double best_cost = 0;
        for (int i = 0; i < 4; ++i) {
            const int nr = kNr64[i];
            const int G = (tile + nr - 1) / nr;
            if (G > 64 || (!fast && nr > 8))
                continue;
            if (opt.f64_rows > 0) {
                if (nr != opt.f64_rows && !(!fast && opt.f64_rows > 8 && nr == 8))
                    continue;
            } else if (roweps() || one_set) {
                if (nr != (fast ? one_set_fast : 8))
                    continue;
            }
            const double rate = kF64SlotRate[fast ? 1 : 0][i];
            if (!(rate > 0.0))
                continue;
            const int W = 64 / G;
            const double cost =
                (double)G * nr / (rate * ((double)(W * G) / 64.0) * (300.0 / (300.0 + G)));
            if (best < 0 || cost < best_cost) {
                best = i;
                best_cost = cost;
            }
        }
        if (best < 0) {  // no tile holds the class: one wave per alignment
            wave[src].bin[wave[src].n] = b;
            wave[src].lanes[wave[src].n] = 64;
            ++wave[src].n;
            continue;
        }
        BinSet& x = set[best][fast][src];
        x.bin[x.n] = b;
        x.lanes[x.n] = (tile + kNr64[best] - 1) / kNr64[best];
        ++x.n;
    }
    int nset = 0;
    for (int i = 3; i >= 0; --i)  // more rows per lane first
        for (int f = 0; f < 2; ++f)
            for (int src = 0; src < 2; ++src) {
                BinSet& x = set[i][f][src];
                if (!x.n)
                    continue;
                if (nset >= kF64Sets)
                    return set_error(PHMM_ERANGE, "too many FP64 set launches");
                x.next_task = bp.task_ctr + (kTaskCounters - kF64Sets + nset++);
                if ((rc = check(launch_fwd_tile_set_f64(kNr64[i], f == 1, fp, x, 0, st, roweps()),
                                "FP64 tile kernel")))
                    return rc;
            }
    for (int src = 0; src < 2; ++src)
        if (wave[src].n &&
            (rc = check(launch_fwd_wave_f64(fp, wave[src], bp.hmax, spill64, n64, bp.error_flag,
                                            num_cus * 8, st),
                        "FP64 wave kernel")))
            return rc;
    BinSet fin;
    fin.n = 1;
    fin.bin[0] = 0;
    fin.lanes[0] = 64;
    fin.items = fp.final_items;
    fin.counts = fp.final_count;
    fin.count_stride = 0;
    if ((rc = check(launch_fwd_wave_f64(fp, fin, bp.hmax, spill64, n64, bp.error_flag, num_cus * 8,
                                        st),
                    "FP64 final kernel")))
        return rc;
    // statistics: the rescue lists; in FP64: the retry lists and the rescue
    // lists of bins without a range-extended pass
    {
        CountSums c;
        c.add(tot, fp.rescue_counts, 2 * ntiles);
        c.add(tot + 1, fp.re_on ? fp.retry_counts : fp.rescue_counts, 2 * ntiles);
        if ((rc = check(launch_add_counts(c, st), "rescue count")))
            return rc;
    }
    if (!fp.re_on)
        return PHMM_OK;
    for (int b = 0; b < 2 * ntiles; ++b)  // tables with rows lacking the range-extended pass
        if (!re_bin(b) &&
            (rc = check(launch_add_count(tot + 1, fp.rescue_counts + b, 1, st), "rescue count")))
            return rc;
    return PHMM_OK;
}

// Captures the given stages of the batched sequence (fixed grids) on a
// private stream and instantiates the graph.
int Engine::capture(int stages, const BinParams& bp, const FwdParams& fp, const BatchDesc& hd,
                    hipGraphExec_t* out)
{
    int rc;
    if (!cap_stream &&
        (rc = check(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking), "stream")))
        return rc;
    if ((rc = check(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeRelaxed),
                    "begin capture")))
        return rc;
    int r2 = PHMM_OK;
    if (stages & kStagePrep)
        r2 = enqueue_prep(bp, hd, true, cap_stream);
    if (!r2 && (stages & kStageFwd))
        r2 = enqueue_fwd(fp, bp, false, cap_stream);
    if (!r2 && (stages & kStageRescue) && opt.fp64_rescue)
        r2 = enqueue_rescue(fp, bp, false, cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cap_stream, &g);
    if (r2) {
        if (g)
            hipGraphDestroy(g);
        return r2;
    }
    if ((rc = check(e, "end capture")))
        return rc;
    const hipError_t e2 = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e2 != hipSuccess)
        *out = nullptr;
    ++graph_captures;
    return check(e2, "graph instantiate");
}

void Engine::derive_small_tiles()
{
    tiles_small = tiles;
    for (auto& t : tiles_small) {
        const int lanes = (t.max_read_len + 16) / 16;  // ceil((max_read_len + 1) / 16)
        if (lanes <= 64) {
            t.lanes = lanes;
            t.rows = 16;
        }
    }
}

void Engine::drop_graphs(CallCtx& c)
{
    Graphs* graphs = c.graphs;
    hipGraphExec_t* all[] = {&graphs[0].all, &graphs[0].prep, &graphs[0].fwd, &graphs[0].rescue,
                             &graphs[1].all, &graphs[1].prep, &graphs[1].fwd, &graphs[1].rescue};
    bool any = false;
    for (auto* x : all)
        any = any || *x;
    if (any)
        hipDeviceSynchronize();  // launched replays may still run
    for (auto* x : all)
        if (*x) {
            hipGraphExecDestroy(*x);
            *x = nullptr;
        }
}

void Engine::free_slot(Slot& s)
{
    if (s.dev.p)
        hipFree(s.dev.p);
    if (s.scores.p)
        hipFree(s.scores.p);
    if (s.pin.p)
        hipHostFree(s.pin.p);
    if (s.pin_scores.p)
        hipHostFree(s.pin_scores.p);
    if (s.h2d_done)
        hipEventDestroy(s.h2d_done);
    if (s.compute_done)
        hipEventDestroy(s.compute_done);
    s = Slot{};
}

// Stages groups of one or more host views, in order, into slot s as one
// self-contained batch (offsets and prefixes rebased) and enqueues the
// copy-in; s.h2d_done fires when it landed.  compute_slot then enqueues the
// scoring (on the slot's context and stream, behind the copy) and the
// copy-out; s.compute_done fires when s.pin_scores holds the scores.
int Engine::stage_parts(Slot& s, const ChunkPart* parts, int nparts, int64_t table_n)
{
    struct Span {
        int r0, r1, h0, h1;
        int64_t rb0, rb1, hb0, hb1;
    };
    std::vector<Span> sp((size_t)nparts);
    int64_t G = 0, nR = 0, nH = 0, rbytes = 0, hbytes = 0, nA = 0;
    for (int k = 0; k < nparts; ++k) {
        const phmm_batch_view& v = *parts[k].v;
        const int g0 = parts[k].g0, g1 = parts[k].g1;
        Span& x = sp[(size_t)k];
        x.r0 = v.reads_prefix[g0];
        x.r1 = v.reads_prefix[g1];
        x.h0 = v.haps_prefix[g0];
        x.h1 = v.haps_prefix[g1];
        x.rb0 = x.r1 > x.r0 ? v.read_offsets[x.r0] : 0;
        x.rb1 = x.r1 > x.r0 ? v.read_offsets[x.r1 - 1] + padded_length(v.read_lengths[x.r1 - 1]) : 0;
        x.hb0 = x.h1 > x.h0 ? v.hap_offsets[x.h0] : 0;
        x.hb1 = x.h1 > x.h0 ? v.hap_offsets[x.h1 - 1] + padded_length(v.hap_lengths[x.h1 - 1]) : 0;
        G += g1 - g0;
        nR += x.r1 - x.r0;
        nH += x.h1 - x.h0;
        rbytes += x.rb1 - x.rb0;
        hbytes += x.hb1 - x.hb0;
        for (int g = g0; g < g1; ++g)
            nA += (int64_t)v.num_reads_per_group[g] * v.num_haps_per_group[g];
    }
    if (G > INT32_MAX || nR > INT32_MAX || nH > INT32_MAX || rbytes > INT32_MAX || hbytes > INT32_MAX)
        return set_error(PHMM_ERANGE, "transfer chunk exceeds 32-bit offsets; lower transfer_batchsize");
    struct Layout {
        size_t read_data, hap_data, bq, iq, dq, gq, rl, hl, ro, ho, nr, nh, rp, hp, total;
    };
    Layout L{};
    {
        size_t o = 0;
        auto take = [&](size_t bytes) {
            const size_t at = o;
            o = align256(o + bytes);
            return at;
        };
        L.read_data = take((size_t)rbytes);
        L.bq = take((size_t)rbytes);
        L.iq = take((size_t)rbytes);
        L.dq = take((size_t)rbytes);
        L.gq = take((size_t)rbytes);
        L.hap_data = take((size_t)hbytes);
        L.rl = take(4 * (size_t)nR);
        L.hl = take(4 * (size_t)nH);
        L.ro = take(4 * (size_t)(nR + 1));
        L.ho = take(4 * (size_t)(nH + 1));
        L.nr = take(4 * (size_t)G);
        L.nh = take(4 * (size_t)G);
        L.rp = take(4 * (size_t)(G + 1));
        L.hp = take(4 * (size_t)(G + 1));
        L.total = o;
    }
    int rc;
    if ((rc = ensure_pin(s.pin, L.total, "staging")) || (rc = grow(s.dev, L.total, "chunk")) ||
        (rc = grow(s.scores, sizeof(float) * (size_t)std::max<int64_t>(nA, 1), "scores")) ||
        (rc = ensure_pin(s.pin_scores, sizeof(float) * (size_t)std::max<int64_t>(nA, 1),
                         "score staging")))
        return rc;
    char* P = static_cast<char*>(s.pin.p);
    int32_t* ro = reinterpret_cast<int32_t*>(P + L.ro);
    int32_t* ho = reinterpret_cast<int32_t*>(P + L.ho);
    int32_t* rp = reinterpret_cast<int32_t*>(P + L.rp);
    int32_t* hp = reinterpret_cast<int32_t*>(P + L.hp);
    // where each part lands in the chunk
    struct Base {
        int64_t rb, hb;
        int32_t r, h, g;
    };
    std::vector<Base> base((size_t)nparts);
    {
        Base cur{0, 0, 0, 0, 0};
        for (int k = 0; k < nparts; ++k) {
            base[(size_t)k] = cur;
            const Span& x = sp[(size_t)k];
            cur.rb += x.rb1 - x.rb0;
            cur.hb += x.hb1 - x.hb0;
            cur.r += x.r1 - x.r0;
            cur.h += x.h1 - x.h0;
            cur.g += parts[k].g1 - parts[k].g0;
        }
    }
    rp[0] = 0;
    hp[0] = 0;
    // the bulk bytes as slices of at most 2 MiB, spread over host threads
    // for large chunks; the per-sequence bookkeeping per part.  A caller's
    // array in pinned memory (the reference's batches are, common.cuh:
    // 115-116) is not staged: it is copied to the device directly.
    struct Copy {
        char* dst;
        const void* src;
        size_t n;
    };
    std::vector<Copy> copies;
    struct H2D {
        size_t off;       // in the chunk layout (device and staging alike)
        const void* src;  // a pinned caller array, or null: the staging copy
        size_t n;
    };
    std::vector<H2D> h2d;
    for (int k = 0; k < nparts; ++k) {
        const phmm_batch_view& v = *parts[k].v;
        const Span& x = sp[(size_t)k];
        const Base& bs = base[(size_t)k];
        const size_t rb = (size_t)(x.rb1 - x.rb0), hb = (size_t)(x.hb1 - x.hb0);
        // gap-continuation qualities are read only under the CPU and
        // per-position gcp semantics (the GPU kernel's eps is fixed)
        const size_t gb = opt.semantics != PHMM_SEM_GPU ? rb : 0;
        const Copy whole[6] = {{P + L.read_data + bs.rb, v.read_data + x.rb0, rb},
                               {P + L.bq + bs.rb, v.base_quals + x.rb0, rb},
                               {P + L.iq + bs.rb, v.ins_quals + x.rb0, rb},
                               {P + L.dq + bs.rb, v.del_quals + x.rb0, rb},
                               {P + L.gq + bs.rb, v.gcp_quals + x.rb0, gb},
                               {P + L.hap_data + bs.hb, v.hap_data + x.hb0, hb}};
        constexpr size_t kSlice = (size_t)2 << 20;
        for (const Copy& c : whole) {
            if (!c.n)
                continue;
            const size_t off = (size_t)(c.dst - P);
            if (host_pinned(c.src)) {
                h2d.push_back({off, c.src, c.n});
                continue;
            }
            h2d.push_back({off, nullptr, c.n});
            for (size_t o = 0; o < c.n; o += kSlice)
                copies.push_back({c.dst + o, static_cast<const char*>(c.src) + o,
                                  std::min(kSlice, c.n - o)});
        }
    }
    h2d.push_back({(size_t)L.rl, nullptr, (size_t)(L.total - L.rl)});  // lengths, offsets, groups
    auto bookkeeping = [&](int k) {
        const phmm_batch_view& v = *parts[k].v;
        const int g0 = parts[k].g0, g1 = parts[k].g1;
        const Span& x = sp[(size_t)k];
        const Base& bs = base[(size_t)k];
        std::memcpy(P + L.rl + 4 * (size_t)bs.r, v.read_lengths + x.r0, 4 * (size_t)(x.r1 - x.r0));
        std::memcpy(P + L.hl + 4 * (size_t)bs.h, v.hap_lengths + x.h0, 4 * (size_t)(x.h1 - x.h0));
        for (int i = x.r0; i < x.r1; ++i)
            ro[bs.r + (i - x.r0)] = (int32_t)(bs.rb + v.read_offsets[i] - x.rb0);
        for (int i = x.h0; i < x.h1; ++i)
            ho[bs.h + (i - x.h0)] = (int32_t)(bs.hb + v.hap_offsets[i] - x.hb0);
        std::memcpy(P + L.nr + 4 * (size_t)bs.g, v.num_reads_per_group + g0, 4 * (size_t)(g1 - g0));
        std::memcpy(P + L.nh + 4 * (size_t)bs.g, v.num_haps_per_group + g0, 4 * (size_t)(g1 - g0));
        for (int g = g0; g < g1; ++g) {
            rp[bs.g + (g - g0) + 1] = bs.r + (v.reads_prefix[g + 1] - x.r0);
            hp[bs.g + (g - g0) + 1] = bs.h + (v.haps_prefix[g + 1] - x.h0);
        }
    };
    const size_t bulk = 5 * (size_t)rbytes + (size_t)hbytes;
    const unsigned hw = std::thread::hardware_concurrency();
    const int nthreads = bulk >= ((size_t)16 << 20) ? (int)std::min(16u, hw > 1 ? hw : 1u) : 1;
    if (nthreads <= 1) {
        for (const Copy& c : copies)
            std::memcpy(c.dst, c.src, c.n);
        for (int k = 0; k < nparts; ++k)
            bookkeeping(k);
    } else {
        std::atomic<size_t> next_copy{0};
        std::atomic<int> next_part{0};
        auto work = [&]() {
            for (size_t i = next_copy++; i < copies.size(); i = next_copy++)
                std::memcpy(copies[i].dst, copies[i].src, copies[i].n);
            for (int k = next_part++; k < nparts; k = next_part++)
                bookkeeping(k);
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nthreads; ++t)
            pool.emplace_back(work);
        work();
        for (auto& th : pool)
            th.join();
    }
    ro[nR] = (int32_t)rbytes;
    ho[nH] = (int32_t)hbytes;
    if (!s.h2d_done) {
        if ((rc = check(hipEventCreateWithFlags(&s.h2d_done, hipEventDisableTiming), "event")) ||
            (rc = check(hipEventCreateWithFlags(&s.compute_done, hipEventDisableTiming), "event")))
            return rc;
    }
    // the copies in layout order, neighbouring staged ranges as one (the
    // gaps between arrays are alignment padding); under the GPU kernel's
    // semantics the gcp region was not staged and stays behind
    std::sort(h2d.begin(), h2d.end(), [](const H2D& a, const H2D& b) { return a.off < b.off; });
    {
        std::vector<H2D> merged;
        for (const H2D& c : h2d) {
            if (!merged.empty() && !c.src && !merged.back().src &&
                c.off <= merged.back().off + merged.back().n + 256)
                merged.back().n = c.off + c.n - merged.back().off;
            else
                merged.push_back(c);
        }
        h2d.swap(merged);
    }
    rc = PHMM_OK;
    for (const H2D& c : h2d)
        if ((rc = check(hipMemcpyAsync(static_cast<char*>(s.dev.p) + c.off,
                                       c.src ? c.src : static_cast<const void*>(P + c.off), c.n,
                                       hipMemcpyHostToDevice, copy_stream),
                        "H2D copy")))
            break;
    if (rc || (rc = check(hipEventRecord(s.h2d_done, copy_stream), "event")))
        return rc;
    char* D = static_cast<char*>(s.dev.p);
    phmm_batch_view& dv = s.dv;
    dv = phmm_batch_view{};
    dv.num_groups = (int32_t)G;
    dv.num_reads = (int32_t)nR;
    dv.num_haps = (int32_t)nH;
    dv.read_bytes = rbytes;
    dv.hap_bytes = hbytes;
    dv.num_alignments = nA;
    dv.read_data = reinterpret_cast<const uint8_t*>(D + L.read_data);
    dv.hap_data = reinterpret_cast<const uint8_t*>(D + L.hap_data);
    dv.base_quals = reinterpret_cast<const uint8_t*>(D + L.bq);
    dv.ins_quals = reinterpret_cast<const uint8_t*>(D + L.iq);
    dv.del_quals = reinterpret_cast<const uint8_t*>(D + L.dq);
    dv.gcp_quals = reinterpret_cast<const uint8_t*>(D + L.gq);
    dv.read_lengths = reinterpret_cast<const int32_t*>(D + L.rl);
    dv.hap_lengths = reinterpret_cast<const int32_t*>(D + L.hl);
    dv.read_offsets = reinterpret_cast<const int32_t*>(D + L.ro);
    dv.hap_offsets = reinterpret_cast<const int32_t*>(D + L.ho);
    dv.num_reads_per_group = reinterpret_cast<const int32_t*>(D + L.nr);
    dv.num_haps_per_group = reinterpret_cast<const int32_t*>(D + L.nh);
    dv.reads_prefix = reinterpret_cast<const int32_t*>(D + L.rp);
    dv.haps_prefix = reinterpret_cast<const int32_t*>(D + L.hp);
    s.table_n = table_n;
    s.n_scores = nA;
    s.busy = true;
    return PHMM_OK;
}

int Engine::compute_slot(Slot& s)
{
    hipStream_t st = s.st ? s.st : stream;
    int rc;
    if ((rc = check(hipStreamWaitEvent(st, s.h2d_done, 0), "event wait")))
        return rc;
    table_hint = s.table_n;
    eager_hint = s.eager;
    rc = score_device(s.dv, static_cast<float*>(s.scores.p), st, nullptr, 0, s.ctx);
    table_hint = -1;
    eager_hint = false;
    if (rc)
        return rc;
    if ((rc = check(hipMemcpyAsync(s.pin_scores.p, s.scores.p, sizeof(float) * (size_t)s.n_scores,
                                   hipMemcpyDeviceToHost, st),
                    "D2H copy")) ||
        (rc = check(hipEventRecord(s.compute_done, st), "event")))
        return rc;
    return PHMM_OK;
}

// Host batch in, host scores out.  Chunks close at the first record where
// the running alignment count reaches transfer_batchsize
// (gpuworker.cuh:659-669); chunk i+1 is staged and copied while chunk i
// computes.  A slot is restaged only after its previous chunk completed, so
// no copy can overtake the compute that reads the slot.
int Engine::score_batch(const phmm_batch_view& v, float* out)
{
    const std::string verr = validate_host_view(v);
    if (!verr.empty())
        return set_error(PHMM_EINVAL, "phmm_score_batch: " + verr);
    if (v.num_alignments == 0)
        return PHMM_OK;
    if (!out)
        return set_error(PHMM_EINVAL, "phmm_score_batch: null output");
    const auto t0 = std::chrono::steady_clock::now();
    const int64_t tb = opt.transfer_batchsize > 0 ? opt.transfer_batchsize : INT64_MAX;
    std::vector<std::pair<int, int>> chunks;
    {
        int g0 = 0;
        int64_t acc = 0;
        for (int g = 0; g < v.num_groups; ++g) {
            acc += (int64_t)v.num_reads_per_group[g] * v.num_haps_per_group[g];
            if (acc >= tb) {
                chunks.push_back({g0, g + 1});
                g0 = g + 1;
                acc = 0;
            }
        }
        if (g0 < v.num_groups)
            chunks.push_back({g0, v.num_groups});
    }
    // Pieces of a transfer chunk: a large chunk is staged, copied and scored
    // in up to four record-aligned pieces, so the copies of piece k+1 overlap
    // the kernels of piece k.  The first piece is half the size of the
    // others (its copy is the one nothing hides).  Every piece scores on the
    // tile table of its whole chunk, so the scores do not depend on the cut.
    struct Piece {
        int g0, g1;
        int64_t table_n;
    };
    std::vector<Piece> pieces;
    for (const auto& ch : chunks) {
        int64_t na = 0;
        for (int g = ch.first; g < ch.second; ++g)
            na += (int64_t)v.num_reads_per_group[g] * v.num_haps_per_group[g];
        const int n = (int)std::max<int64_t>(1, std::min<int64_t>(4, na / 250000));
        int g0 = ch.first;
        int64_t acc = 0;
        int k = 1;
        // cut k after a fraction (2k - 1) / (2n - 1) of the alignments
        for (int g = ch.first; g < ch.second && k < n; ++g) {
            acc += (int64_t)v.num_reads_per_group[g] * v.num_haps_per_group[g];
            if (acc * (2 * n - 1) >= na * (2 * k - 1)) {
                pieces.push_back({g0, g + 1, na});
                g0 = g + 1;
                ++k;
            }
        }
        if (g0 < ch.second)
            pieces.push_back({g0, ch.second, na});
    }
    // Alternate pieces run on a second call context and stream, each on a
    // hardware queue of its own, so a piece's kernels start while the
    // previous piece's last waves finish.  Without a free context (a server
    // holds them) both slots use context 0 and the engine's stream.
    int ctx2 = -1;
    if (pieces.size() > 1) {
        ctx2 = claim_ctx();
        if (ctx2 > 0 && !stream2 && create_queue_stream(&stream2) != PHMM_OK) {
            release_ctx(ctx2);
            ctx2 = -1;
        }
    }
    struct Release {
        Engine* e;
        int k;
        ~Release()
        {
            if (k > 0)
                e->release_ctx(k);
        }
    } release{this, ctx2};
    int rc;
    if ((rc = take_error_flag()))
        return rc;
    int64_t out_off[kSlots] = {};
    auto drain = [&](int si) -> int {
        Slot& s = slots[si];
        if (!s.busy)
            return PHMM_OK;
        int r2;
        if ((r2 = check(hipEventSynchronize(s.compute_done), "chunk sync")))
            return r2;
        std::memcpy(out + out_off[si], s.pin_scores.p, sizeof(float) * (size_t)s.n_scores);
        s.busy = false;
        return PHMM_OK;
    };
    // Software pipeline over a ring of slots: piece k is staged and its
    // copy enqueued before piece k-1's scoring is, so the copy engine always
    // has the next piece queued while the kernels of the previous one run.
    // Piece k scores on context/stream k % 2.
    int64_t next_off = 0;
    int pending = -1;  // staged piece whose scoring is not enqueued yet
    for (size_t ci = 0; ci < pieces.size(); ++ci) {
        const int si = (int)(ci % kSlots);
        if ((rc = drain(si)))
            return rc;
        Slot& s = slots[si];
        const bool alt = (ci % 2) == 1 && ctx2 > 0;
        s.ctx = alt ? ctx2 : 0;
        s.st = alt ? stream2 : stream;
        s.eager = pieces.size() > 1;
        const ChunkPart part{&v, pieces[ci].g0, pieces[ci].g1};
        if ((rc = stage_parts(s, &part, 1, pieces[ci].table_n)))
            return rc;
        out_off[si] = next_off;
        next_off += s.n_scores;
        if (pending >= 0 && (rc = compute_slot(slots[pending])))
            return rc;
        pending = si;
    }
    if (pending >= 0 && (rc = compute_slot(slots[pending])))
        return rc;
    for (int si = 0; si < kSlots; ++si)
        if ((rc = drain(si)))
            return rc;
    if ((rc = check(hipStreamSynchronize(stream), "final sync")) ||
        (ctx2 > 0 && (rc = check(hipStreamSynchronize(stream2), "final sync"))))
        return rc;
    finish_timing();
    if ((rc = take_error_flag()))
        return rc;
    stats.cells += phmm_batch_cells(&v);
    stats.engine_seconds +=
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return PHMM_OK;
}

}  // namespace phmm

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using phmm::Engine;

struct phmm_engine {
    Engine e;
    std::mutex mu;
};

struct phmm_device_batch {
    phmm_engine* eng = nullptr;
    void* mem = nullptr;
    phmm_batch_view view{};
};

namespace phmm {
Engine* engine_of(phmm_engine* h, std::mutex** mu)
{
    if (mu)
        *mu = &h->mu;
    return &h->e;
}
}  // namespace phmm

extern "C" int phmm_abi_version(void) { return PHMM_ABI_VERSION; }
extern "C" const char* phmm_last_error(void) { return phmm::last_error(); }

extern "C" void phmm_options_default(phmm_options* o)
{
    if (!o)
        return;
    std::memset(o, 0, sizeof *o);
    o->transfer_batchsize = 1000000;  // common.cuh:19
    o->mode = PHMM_MODE_BATCHED;
    o->semantics = PHMM_SEM_GPU;
    o->fp64_rescue = 1;
    o->rescue_min_sum = FLT_MIN;  // where the reference's FP32 sum underflows
    o->verbose = 0;
    o->graphs = 1;
    o->small_batch = 1 << 18;
    o->range_rescue = 1;
}

extern "C" int phmm_engine_create(int device, const phmm_options* opt, phmm_engine** out)
{
    if (!out)
        return phmm::fail(PHMM_EINVAL, "phmm_engine_create: null output");
    *out = nullptr;
    phmm_options o;
    if (opt)
        o = *opt;
    else
        phmm_options_default(&o);
    if (o.mode != PHMM_MODE_PLAIN && o.mode != PHMM_MODE_BATCHED)
        return phmm::fail(PHMM_EINVAL, "phmm_engine_create: unknown mode");
    if (o.semantics != PHMM_SEM_GPU && o.semantics != PHMM_SEM_CPU && o.semantics != PHMM_SEM_GCP)
        return phmm::fail(PHMM_EINVAL, "phmm_engine_create: unknown semantics");
    if (o.f64_rows != 0 && o.f64_rows != 4 && o.f64_rows != 6 && o.f64_rows != 8 && o.f64_rows != 10)
        return phmm::fail(PHMM_EINVAL, "phmm_engine_create: f64_rows must be 0, 4, 6, 8 or 10");
    if (o.range_rescue < 0 || o.range_rescue > 2)
        return phmm::fail(PHMM_EINVAL, "phmm_engine_create: range_rescue must be 0, 1 or 2");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        hipGetLastError();
        return phmm::fail(PHMM_EDEVICE, "phmm_engine_create: no HIP device");
    }
    if (device < 0 || device >= ndev)
        return phmm::fail(PHMM_EINVAL, "phmm_engine_create: device out of range");
    auto* h = new (std::nothrow) phmm_engine;
    if (!h)
        return phmm::fail(PHMM_ENOMEM, "phmm_engine_create: out of memory");
    Engine& e = h->e;
    e.device = device;
    e.opt = o;
    e.tiles.assign(phmm::kDefaultTiles, phmm::kDefaultTiles + phmm::kNumDefaultTiles);
    e.derive_small_tiles();
    int rc;
    auto bail = [&](int code) {
        const std::string m = e.err.empty() ? phmm::last_error() : e.err;
        delete h;
        return phmm::fail(code, m);
    };
    if ((rc = e.check(hipSetDevice(device), "hipSetDevice")))
        return bail(rc);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess)
        e.num_cus = prop.multiProcessorCount;
    else
        hipGetLastError();
    if ((rc = e.check(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking), "stream")) ||
        (rc = e.check(hipStreamCreateWithFlags(&e.copy_stream, hipStreamNonBlocking), "stream")) ||
        (rc = e.check(hipMalloc(&e.d_ph2pr, 128 * sizeof(float)), "ph2pr")) ||
        (rc = e.check(hipMalloc(&e.d_rescued_total, Engine::kTotals * sizeof(int64_t)), "counter")) ||
        (rc = e.check(hipMemset(e.d_rescued_total, 0, Engine::kTotals * sizeof(int64_t)), "counter")) ||
        (rc = e.check(hipHostMalloc(&e.h_small, 256, hipHostMallocDefault), "pinned")) ||
        (rc = e.init_ctx(0)) || (rc = phmm::noalloc_init(device)))
        return bail(rc);
    // ph2pr[q] = (float)10^(-q/10) (gpuworker.cuh:473-486)
    float table[128];
    for (int q = 0; q < 128; ++q)
        table[q] = (float)std::pow(10.0, (double)(-q) / 10.0);
    if ((rc = e.check(hipMemcpy(e.d_ph2pr, table, sizeof table, hipMemcpyHostToDevice), "ph2pr")))
        return bail(rc);
    *out = h;
    return PHMM_OK;
}

extern "C" void phmm_engine_destroy(phmm_engine* h) { delete h; }

extern "C" const char* phmm_engine_last_error(const phmm_engine* h)
{
    return h ? h->e.err.c_str() : phmm::last_error();
}

extern "C" int phmm_engine_stats(const phmm_engine* h, phmm_stats* out)
{
    if (!h || !out)
        return phmm::fail(PHMM_EINVAL, "phmm_engine_stats: null argument");
    auto* m = const_cast<phmm_engine*>(h);
    std::lock_guard<std::mutex> lk(m->mu);
    hipSetDevice(m->e.device);
    hipStreamSynchronize(m->e.stream);
    m->e.finish_timing();
    int64_t rescued[phmm::Engine::kTotals] = {0, 0, 0};
    if (m->e.d_rescued_total) {
        if (hipMemcpy(rescued, m->e.d_rescued_total, sizeof rescued, hipMemcpyDeviceToHost) !=
            hipSuccess)
            hipGetLastError();
    }
    *out = h->e.stats;
    out->rescued = rescued[0];
    out->f64_alignments = rescued[1];
    out->scaled_alignments = rescued[2];
    out->graph_launches = h->e.graph_launches;
    out->graph_captures = h->e.graph_captures;
    return PHMM_OK;
}

extern "C" int phmm_engine_reset_stats(phmm_engine* h)
{
    if (!h)
        return phmm::fail(PHMM_EINVAL, "phmm_engine_reset_stats: null engine");
    std::lock_guard<std::mutex> lk(h->mu);
    h->e.finish_timing();
    h->e.stats = phmm_stats{};
    h->e.graph_launches = 0;
    h->e.graph_captures = 0;
    if (h->e.d_rescued_total) {
        hipSetDevice(h->e.device);
        hipDeviceSynchronize();
        if (hipMemset(h->e.d_rescued_total, 0, phmm::Engine::kTotals * sizeof(int64_t)) != hipSuccess)
            hipGetLastError();
    }
    return PHMM_OK;
}

extern "C" int phmm_engine_set_tiles(phmm_engine* h, const phmm_tile_config* t, int n)
{
    if (!h || !t || n <= 0 || n > phmm::kMaxTiles)
        return phmm::fail(PHMM_EINVAL, "phmm_engine_set_tiles: bad arguments");
    int prev = 0;
    for (int i = 0; i < n; ++i) {
        const int tile = t[i].lanes * t[i].rows;
        if (t[i].lanes < 1 || t[i].lanes > 64 || !phmm::tile_rows_supported(t[i].rows) ||
            t[i].max_read_len <= prev || tile < t[i].max_read_len + 1)
            return phmm::fail(PHMM_EINVAL,
                              "phmm_engine_set_tiles: lanes 1..64, rows 4, 8, 12, 13, 16, 19, 20, "
                              "24, 28 or 32, max_read_len ascending and <= lanes*rows - 1");
        prev = t[i].max_read_len;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    h->e.tiles.assign(t, t + n);
    h->e.derive_small_tiles();
    for (auto& c : h->e.ctxs)
        ++c.generation;  // captured graphs hold the old table
    return PHMM_OK;
}

extern "C" int phmm_engine_get_tiles(const phmm_engine* h, phmm_tile_config* t, int cap, int* n)
{
    if (!h || !n)
        return phmm::fail(PHMM_EINVAL, "phmm_engine_get_tiles: null argument");
    *n = (int)h->e.tiles.size();
    for (int i = 0; i < *n && i < cap && t; ++i)
        t[i] = h->e.tiles[(size_t)i];
    return PHMM_OK;
}

extern "C" int phmm_score_batch(phmm_engine* h, const phmm_batch_view* b, float* out)
{
    if (!h || !b)
        return phmm::fail(PHMM_EINVAL, "phmm_score_batch: null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    hipSetDevice(h->e.device);
    return h->e.score_batch(*b, out);
}

extern "C" int phmm_score_device(phmm_engine* h, const phmm_batch_view* b, float* d_scores,
                                 void* hip_stream)
{
    if (!h || !b)
        return phmm::fail(PHMM_EINVAL, "phmm_score_device: null argument");
    std::lock_guard<std::mutex> lk(h->mu);
    hipSetDevice(h->e.device);
    hipStream_t st = hip_stream ? static_cast<hipStream_t>(hip_stream) : h->e.stream;
    return h->e.score_device(*b, d_scores, st);
}

extern "C" int phmm_engine_synchronize(phmm_engine* h)
{
    if (!h)
        return phmm::fail(PHMM_EINVAL, "phmm_engine_synchronize: null engine");
    std::lock_guard<std::mutex> lk(h->mu);
    hipSetDevice(h->e.device);
    int rc = h->e.check(hipDeviceSynchronize(), "synchronize");
    if (rc)
        return rc;
    h->e.finish_timing();
    return h->e.take_error_flag();
}

extern "C" int phmm_device_batch_upload(phmm_engine* h, const phmm_batch_view* v,
                                        phmm_device_batch** out)
{
    if (!h || !v || !out)
        return phmm::fail(PHMM_EINVAL, "phmm_device_batch_upload: null argument");
    *out = nullptr;
    const std::string verr = phmm::validate_host_view(*v);
    if (!verr.empty())
        return phmm::fail(PHMM_EINVAL, "phmm_device_batch_upload: " + verr);
    std::lock_guard<std::mutex> lk(h->mu);
    Engine& e = h->e;
    hipSetDevice(e.device);
    const int G = v->num_groups, nR = v->num_reads, nH = v->num_haps;
    size_t off[14];
    size_t o = 0;
    const size_t sizes[14] = {(size_t)v->read_bytes, (size_t)v->read_bytes, (size_t)v->read_bytes,
                              (size_t)v->read_bytes, (size_t)v->read_bytes, (size_t)v->hap_bytes,
                              4 * (size_t)nR,          4 * (size_t)nH,          4 * (size_t)(nR + 1),
                              4 * (size_t)(nH + 1),    4 * (size_t)G,           4 * (size_t)G,
                              4 * (size_t)(G + 1),     4 * (size_t)(G + 1)};
    const void* src[14] = {v->read_data,   v->base_quals,   v->ins_quals,           v->del_quals,
                           v->gcp_quals,   v->hap_data,     v->read_lengths,        v->hap_lengths,
                           v->read_offsets, v->hap_offsets, v->num_reads_per_group, v->num_haps_per_group,
                           v->reads_prefix, v->haps_prefix};
    for (int i = 0; i < 14; ++i) {
        off[i] = o;
        o = phmm::align256(o + sizes[i]);
    }
    auto* db = new (std::nothrow) phmm_device_batch;
    if (!db)
        return phmm::fail(PHMM_ENOMEM, "phmm_device_batch_upload: out of memory");
    if (hipMalloc(&db->mem, std::max<size_t>(o, 256)) != hipSuccess) {
        hipGetLastError();
        delete db;
        return e.set_error(PHMM_ENOMEM, "phmm_device_batch_upload: device allocation failed");
    }
    char* D = static_cast<char*>(db->mem);
    for (int i = 0; i < 14; ++i)
        if (sizes[i] && src[i] &&
            hipMemcpy(D + off[i], src[i], sizes[i], hipMemcpyHostToDevice) != hipSuccess) {
            hipGetLastError();
            hipFree(db->mem);
            delete db;
            return e.set_error(PHMM_EDEVICE, "phmm_device_batch_upload: copy failed");
        }
    phmm_batch_view dv = *v;
    dv.read_data = reinterpret_cast<const uint8_t*>(D + off[0]);
    dv.base_quals = reinterpret_cast<const uint8_t*>(D + off[1]);
    dv.ins_quals = reinterpret_cast<const uint8_t*>(D + off[2]);
    dv.del_quals = reinterpret_cast<const uint8_t*>(D + off[3]);
    dv.gcp_quals = reinterpret_cast<const uint8_t*>(D + off[4]);
    dv.hap_data = reinterpret_cast<const uint8_t*>(D + off[5]);
    dv.read_lengths = reinterpret_cast<const int32_t*>(D + off[6]);
    dv.hap_lengths = reinterpret_cast<const int32_t*>(D + off[7]);
    dv.read_offsets = reinterpret_cast<const int32_t*>(D + off[8]);
    dv.hap_offsets = reinterpret_cast<const int32_t*>(D + off[9]);
    dv.num_reads_per_group = reinterpret_cast<const int32_t*>(D + off[10]);
    dv.num_haps_per_group = reinterpret_cast<const int32_t*>(D + off[11]);
    dv.reads_prefix = reinterpret_cast<const int32_t*>(D + off[12]);
    dv.haps_prefix = reinterpret_cast<const int32_t*>(D + off[13]);
    db->eng = h;
    db->view = dv;
    *out = db;
    return PHMM_OK;
}

extern "C" int phmm_device_batch_view(const phmm_device_batch* db, phmm_batch_view* out)
{
    if (!db || !out)
        return phmm::fail(PHMM_EINVAL, "phmm_device_batch_view: null argument");
    *out = db->view;
    return PHMM_OK;
}

extern "C" void phmm_device_batch_free(phmm_device_batch* db)
{
    if (!db)
        return;
    if (db->mem) {
        hipDeviceSynchronize();
        hipFree(db->mem);
    }
    delete db;
}

extern "C" int phmm_device_alloc(phmm_engine* h, size_t bytes, void** out)
{
    if (!h || !out)
        return phmm::fail(PHMM_EINVAL, "phmm_device_alloc: null argument");
    hipSetDevice(h->e.device);
    if (hipMalloc(out, std::max<size_t>(bytes, 1)) != hipSuccess) {
        hipGetLastError();
        *out = nullptr;
        return phmm::fail(PHMM_ENOMEM, "phmm_device_alloc: device allocation failed");
    }
    return PHMM_OK;
}

extern "C" int phmm_device_free(phmm_engine* h, void* p)
{
    if (!h)
        return phmm::fail(PHMM_EINVAL, "phmm_device_free: null engine");
    if (p) {
        hipDeviceSynchronize();
        hipFree(p);
    }
    return PHMM_OK;
}

extern "C" int phmm_host_alloc_pinned(phmm_engine* h, size_t bytes, void** out)
{
    if (!h || !out)
        return phmm::fail(PHMM_EINVAL, "phmm_host_alloc_pinned: null argument");
    hipSetDevice(h->e.device);
    if (hipHostMalloc(out, std::max<size_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
        hipGetLastError();
        *out = nullptr;
        return phmm::fail(PHMM_ENOMEM, "phmm_host_alloc_pinned: pinned allocation failed");
    }
    return PHMM_OK;
}

extern "C" int phmm_host_free_pinned(phmm_engine* h, void* p)
{
    if (!h)
        return phmm::fail(PHMM_EINVAL, "phmm_host_free_pinned: null engine");
    if (p)
        hipHostFree(p);
    return PHMM_OK;
}

extern "C" int phmm_memcpy_d2h(phmm_engine* h, void* dst, const void* src, size_t bytes)
{
    if (!h || (bytes && (!dst || !src)))
        return phmm::fail(PHMM_EINVAL, "phmm_memcpy_d2h: null argument");
    hipSetDevice(h->e.device);
    return h->e.check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "phmm_memcpy_d2h");
}

extern "C" int phmm_memcpy_h2d(phmm_engine* h, void* dst, const void* src, size_t bytes)
{
    if (!h || (bytes && (!dst || !src)))
        return phmm::fail(PHMM_EINVAL, "phmm_memcpy_h2d: null argument");
    hipSetDevice(h->e.device);
    return h->e.check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "phmm_memcpy_h2d");
}
