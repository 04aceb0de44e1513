// Serving loop: continuous batching of active-region requests on one GPU.
//
// Admission (phmm_server_submit).  A request (a small host batch, usually
// one region = one record) is admitted when its batch image fits the device
// arena -- the HBM budget less the device memory the server's launch
// contexts hold (engine scratch, merge buffer, scores), counted at creation
// from the real allocations -- and the in-flight limit; otherwise submit
// returns PHMM_EBUSY.  An admitted request is staged in pinned memory and
// copied into its arena block, where it stays until its scores are
// returned.  The copy is issued at once when no copy is in flight, else by
// the next flush (at the latest when the worker merges the request), which
// moves every staged request with one copy per run contiguous in staging
// and arena: a burst of arrivals costs a few copy commands, not one each,
// and a launch waits only for the copies of what was staged before it.
//
// Launches.  A worker thread merges the admitted requests that wait (in
// arrival order, up to the per-launch caps) as soon as one of the server's
// launch contexts is free: a gather kernel assembles the merged batch from
// the arena on the device, then the engine scores it (binning, persistent
// forward kernels, FP64 rescue) and the scores return to pinned memory.
// Every context has its own stream (default priority) and its own engine
// scratch, so a new launch is queued on the device while others run and
// its kernels fill the compute units their tails leave idle: requests that
// arrive while others are on the GPU are admitted, merged and overlapped.
// A finished launch's scores are split per request, its arena blocks
// released, and each request's result carries its own status.
//
// Every alignment is scored by a kernel chosen from its own read length and
// coefficients only, so a request's scores are the same alone and in any
// mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "common.hpp"
#include "engine.hpp"
#include "gather.hpp"

namespace phmm {
namespace {

using Clock = std::chrono::steady_clock;

// First-in-first-out allocator over [0, cap) whose blocks may be released in
// any order: a block starts after the newest one, wrapping to 0 when the end
// is too short; released space returns once every older block is released.
struct Ring {
    size_t cap = 0, head = 0, live = 0;
    struct Block {
        size_t off, bytes;
        bool used;
    };
    std::deque<Block> blocks;

    bool alloc(size_t bytes, size_t* off)
    {
        if (bytes == 0 || bytes > cap)
            return false;
        size_t at;
        if (blocks.empty()) {
            at = 0;
        } else {
            const size_t tail = blocks.front().off;
            if (head > tail) {  // free: [head, cap) and [0, tail)
                if (cap - head >= bytes)
                    at = head;
                else if (tail >= bytes)
                    at = 0;
                else
                    return false;
            } else {  // wrapped: free [head, tail)
                if (tail - head >= bytes)
                    at = head;
                else
                    return false;
            }
        }
        blocks.push_back({at, bytes, true});
        head = at + bytes;
        live += bytes;
        *off = at;
        return true;
    }
    void release(size_t off)
    {
        for (Block& b : blocks)
            if (b.used && b.off == off) {
                b.used = false;
                live -= b.bytes;
                break;
            }
        while (!blocks.empty() && !blocks.front().used)
            blocks.pop_front();
        if (blocks.empty())
            head = 0;
    }
};

constexpr size_t kAlign = 256;
inline size_t round_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

struct Request {
    uint64_t tag = 0;
    int64_t nA = 0, cells = 0;
    int32_t G = 0, nR = 0, nH = 0;
    int64_t rbytes = 0, hbytes = 0;
    RequestLayout lay{};
    size_t off = 0, bytes = 0;  // arena block
    uint64_t seq = 0;           // submission order
    Clock::time_point t_submit;
};

struct Result {
    uint64_t tag = 0;
    std::vector<float> scores;
    double latency = 0.0;
    int rc = PHMM_OK;
    std::string msg;
};

// A launch context of the server: an engine call context, its stream, the
// merge buffer and the launch's scores.
struct LaunchCtx {
    int eng_ctx = -1;
    hipStream_t stream = nullptr;
    DevBuf buf, scores, d_reqs;
    PinBuf pin_scores, h_reqs;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<Request> reqs;
    int64_t n_scores = 0;
    Clock::time_point t_launch;
};

// Groups a launch may hold per request it may merge.
constexpr int kGroupsPerRequest = 16;
// Staged bytes that are copied without waiting for the flush in flight.
constexpr size_t kFlushBytes = 8 << 20;

}  // namespace
}  // namespace phmm

using phmm::Clock;
using phmm::LaunchCtx;
using phmm::Request;
using phmm::Result;

struct phmm_server {
    phmm_engine* handle = nullptr;
    phmm::Engine* eng = nullptr;
    std::mutex* eng_mu = nullptr;
    phmm_server_options opt{};
    int64_t max_groups = 0;
    // device arena of admitted requests and the pinned staging they pass
    // through (each its own ring); the copy stream moves staged images
    phmm::DevBuf arena;
    phmm::Ring arena_ring;
    phmm::PinBuf stage;
    phmm::Ring stage_ring;
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_mark = nullptr;  // recorded after every flush of staged copies
    // Staged requests whose copies are not issued yet (submission order):
    // staging offset, arena offset, bytes.  A flush issues one copy per run
    // contiguous in both rings, so a backlog costs a few copy commands, not
    // one per request.
    struct Staged {
        size_t soff, aoff, bytes;
    };
    std::vector<Staged> unflushed;
    size_t unflushed_bytes = 0;
    uint64_t next_seq = 0;
    std::atomic<uint64_t> flushed_seq{0};  // requests below it have their copies issued
    struct Flight {
        std::vector<size_t> soffs;  // staging blocks the copies read
        hipEvent_t ev;
    };
    std::deque<Flight> stage_pending;  // flushes in flight
    std::vector<hipEvent_t> event_pool;
    std::mutex submit_mu;  // staging, the copy stream and the fields above
    // flushes whose copies failed: [first, second) submission ranges whose
    // requests fail with copy_rc instead of launching (under mu)
    std::vector<std::pair<uint64_t, uint64_t>> copy_failed;
    int copy_rc = PHMM_OK;
    std::string copy_msg;
    int64_t fixed_bytes = 0;

    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request> queue;  // admitted, not launched
    int64_t queued_na = 0;      // alignments in the queue
    std::deque<Result> done;
    std::vector<LaunchCtx> ctxs;
    bool stop = false;
    phmm_server_stats st{};
    std::thread worker;

    ~phmm_server();
    int init();
    int create_launch_stream(hipStream_t* out);
    void run();
    bool retire(LaunchCtx& c, bool block);
    void launch(LaunchCtx& c, std::vector<Request>&& reqs);
    void fail_requests(std::vector<Request>& reqs, int rc, const std::string& msg);
    void reclaim_stage(bool block_one);
    void flush_copies();
    int check(hipError_t e, const char* what)
    {
        if (e == hipSuccess)
            return PHMM_OK;
        (void)hipGetLastError();
        return phmm::fail(e == hipErrorOutOfMemory ? PHMM_ENOMEM : PHMM_EDEVICE,
                          std::string("server: ") + what + ": " + hipGetErrorString(e));
    }
};

// A launch context's stream, on a hardware queue of its own: streams
// otherwise share the process's few queues, where kernels of different
// streams run one after another and launches could not overlap.  A stream
// created with a CU mask gets its own queue; the mask names every CU.
// Default priority, like every stream here.
int phmm_server::create_launch_stream(hipStream_t* out)
{
    const int rc = eng->create_queue_stream(out);
    return rc ? phmm::fail(rc, std::string("server: ") + eng->err) : PHMM_OK;
}

// Allocates the launch contexts at the per-launch caps (engine scratch
// included), then the arena: the budget less what the contexts hold.
int phmm_server::init()
{
    int rc;
    std::lock_guard<std::mutex> elk(*eng_mu);
    hipSetDevice(eng->device);
    const int64_t A = opt.max_batch_alignments;
    const int64_t B = opt.max_batch_bytes;
    max_groups = (int64_t)opt.max_batch_requests * phmm::kGroupsPerRequest;
    // a batch at every cap at once bounds the scratch of any merged launch
    phmm_batch_view capv{};
    capv.num_groups = (int32_t)std::min<int64_t>(max_groups, INT32_MAX);
    capv.read_bytes = B / 5;
    capv.hap_bytes = B;
    capv.num_reads = (int32_t)std::min<int64_t>(B / 20, INT32_MAX);
    capv.num_haps = (int32_t)std::min<int64_t>(B / 4, INT32_MAX);
    capv.num_alignments = A;
    ctxs.resize((size_t)opt.streams);
    for (LaunchCtx& c : ctxs) {
        c.eng_ctx = eng->claim_ctx();
        if (c.eng_ctx < 0)
            return phmm::fail(PHMM_EBUSY, "phmm_server_create: the engine has no free call context");
        eng->ctxs[c.eng_ctx].spill_bytes = (size_t)32 << 20;
        if ((rc = eng->reserve_ctx(c.eng_ctx, capv)))
            return rc;
        if ((rc = create_launch_stream(&c.stream)) ||
            (rc = check(hipEventCreateWithFlags(&c.done, hipEventDisableTiming), "event")) ||
            (rc = eng->grow(c.buf, (size_t)B, "merge buffer")) ||
            (rc = eng->grow(c.scores, sizeof(float) * (size_t)A, "launch scores")) ||
            (rc = eng->grow(c.d_reqs, sizeof(phmm::GatherReq) * (size_t)opt.max_batch_requests,
                            "merge list")) ||
            (rc = eng->ensure_pin(c.pin_scores, sizeof(float) * (size_t)A, "score staging")) ||
            (rc = eng->ensure_pin(c.h_reqs, sizeof(phmm::GatherReq) * (size_t)opt.max_batch_requests,
                                  "merge list staging")))
            return rc;
        fixed_bytes += (int64_t)(eng->ctx_bytes(c.eng_ctx) + c.buf.cap + c.scores.cap + c.d_reqs.cap);
    }
    if (fixed_bytes + (int64_t)(64 << 10) > opt.hbm_budget_bytes)
        return phmm::fail(PHMM_EINVAL,
                          "phmm_server_create: the HBM budget does not cover the launch contexts (" +
                              std::to_string(fixed_bytes) + " bytes) and 64 KiB of requests");
    const size_t cap = (size_t)(opt.hbm_budget_bytes - fixed_bytes) & ~(phmm::kAlign - 1);
    if ((rc = eng->grow(arena, cap, "request arena")))
        return rc;
    arena_ring.cap = cap;
    const size_t stage_cap = (size_t)std::min<int64_t>(std::max<int64_t>(2 * B, 64 << 20),
                                                        (int64_t)cap);
    if ((rc = eng->ensure_pin(stage, stage_cap, "request staging")))
        return rc;
    stage_ring.cap = stage_cap;
    if ((rc = check(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking), "stream")) ||
        (rc = check(hipEventCreateWithFlags(&copy_mark, hipEventDisableTiming), "event")))
        return rc;
    st.scratch_bytes = fixed_bytes;
    st.arena_bytes = (int64_t)cap;
    st.bytes_in_use = fixed_bytes;
    return PHMM_OK;
}

phmm_server::~phmm_server()
{
    hipSetDevice(eng->device);
    for (LaunchCtx& c : ctxs) {
        if (c.stream)
            hipStreamSynchronize(c.stream);
        for (phmm::DevBuf* b : {&c.buf, &c.scores, &c.d_reqs})
            if (b->p)
                hipFree(b->p);
        for (phmm::PinBuf* b : {&c.pin_scores, &c.h_reqs})
            if (b->p)
                hipHostFree(b->p);
        if (c.done)
            hipEventDestroy(c.done);
        if (c.stream)
            hipStreamDestroy(c.stream);
        if (c.eng_ctx >= 0) {
            std::lock_guard<std::mutex> elk(*eng_mu);
            eng->release_ctx(c.eng_ctx);
        }
    }
    if (copy_stream)
        hipStreamSynchronize(copy_stream);
    for (Flight& f : stage_pending)
        event_pool.push_back(f.ev);
    for (hipEvent_t e : event_pool)
        hipEventDestroy(e);
    if (copy_mark)
        hipEventDestroy(copy_mark);
    if (copy_stream)
        hipStreamDestroy(copy_stream);
    if (arena.p)
        hipFree(arena.p);
    if (stage.p)
        hipHostFree(stage.p);
}

// Returns staging blocks whose copies finished (waiting for the oldest flush
// when block_one).  Caller holds submit_mu.
void phmm_server::reclaim_stage(bool block_one)
{
    while (!stage_pending.empty()) {
        Flight& f = stage_pending.front();
        const hipError_t q = block_one ? hipEventSynchronize(f.ev) : hipEventQuery(f.ev);
        if (q == hipErrorNotReady)
            return;
        (void)hipGetLastError();
        for (size_t off : f.soffs)
            stage_ring.release(off);
        event_pool.push_back(f.ev);
        stage_pending.pop_front();
        block_one = false;
    }
}

// Issues the copies of every staged request into its arena block on the
// copy stream, then records copy_mark (which launches wait for).  A failure
// is kept in copy_rc: the worker fails those requests instead of launching
// them.  Caller holds submit_mu, not mu.
void phmm_server::flush_copies()
{
    if (unflushed.empty())
        return;
    hipSetDevice(eng->device);
    hipEvent_t ev = nullptr;
    if (!event_pool.empty()) {
        ev = event_pool.back();
        event_pool.pop_back();
    }
    Flight f;
    f.soffs.reserve(unflushed.size());
    int rc = ev ? PHMM_OK : check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
    for (size_t i = 0; i < unflushed.size() && !rc;) {
        size_t j = i + 1, bytes = unflushed[i].bytes;
        while (j < unflushed.size() &&
               unflushed[j].soff == unflushed[j - 1].soff + unflushed[j - 1].bytes &&
               unflushed[j].aoff == unflushed[j - 1].aoff + unflushed[j - 1].bytes)
            bytes += unflushed[j++].bytes;
        rc = check(hipMemcpyAsync(static_cast<char*>(arena.p) + unflushed[i].aoff,
                                  static_cast<const char*>(stage.p) + unflushed[i].soff, bytes,
                                  hipMemcpyHostToDevice, copy_stream),
                   "request copy");
        i = j;
    }
    if (!rc && !(rc = check(hipEventRecord(ev, copy_stream), "event")))
        rc = check(hipEventRecord(copy_mark, copy_stream), "event");
    for (const Staged& u : unflushed)
        f.soffs.push_back(u.soff);
    unflushed.clear();
    unflushed_bytes = 0;
    if (rc) {
        // nothing of this flush can be trusted to arrive: its staging is
        // free again and its requests fail in the worker
        (void)hipStreamSynchronize(copy_stream);
        (void)hipGetLastError();
        for (size_t off : f.soffs)
            stage_ring.release(off);
        if (ev)
            event_pool.push_back(ev);
        std::lock_guard<std::mutex> lk(mu);
        copy_failed.push_back({flushed_seq.load(std::memory_order_relaxed), next_seq});
        copy_rc = rc;
        copy_msg = phmm::last_error();
    } else {
        f.ev = ev;
        stage_pending.push_back(std::move(f));
    }
    flushed_seq.store(next_seq, std::memory_order_release);
}

void phmm_server::fail_requests(std::vector<Request>& reqs, int rc, const std::string& msg)
{
    const auto now = Clock::now();
    for (Request& r : reqs) {
        Result res;
        res.tag = r.tag;
        res.scores.assign((size_t)r.nA, std::nanf(""));
        res.latency = std::chrono::duration<double>(now - r.t_submit).count();
        res.rc = rc;
        res.msg = msg;
        arena_ring.release(r.off);
        st.inflight -= 1;
        st.completed += 1;
        st.failed += 1;
        done.push_back(std::move(res));
    }
    st.bytes_in_use = fixed_bytes + (int64_t)arena_ring.live;
    cv.notify_all();
}

// Merges `reqs` into context c's buffer on the device and enqueues their
// scoring; c.done fires when c.pin_scores holds the scores.
void phmm_server::launch(LaunchCtx& c, std::vector<Request>&& reqs)
{
    int64_t rb = 0, hb = 0, nA = 0, cells = 0;
    int32_t nR = 0, nH = 0, G = 0;
    auto* gr = static_cast<phmm::GatherReq*>(c.h_reqs.p);
    const char* abase = static_cast<const char*>(arena.p);
    for (size_t k = 0; k < reqs.size(); ++k) {
        const Request& r = reqs[k];
        phmm::GatherReq& q = gr[k];
        q.src = reinterpret_cast<const uint8_t*>(abase + r.off);
        q.lay = r.lay;
        q.rbytes = r.rbytes;
        q.hbytes = r.hbytes;
        q.rb_dst = rb;
        q.hb_dst = hb;
        q.nR = r.nR;
        q.nH = r.nH;
        q.G = r.G;
        q.r_dst = nR;
        q.h_dst = nH;
        q.g_dst = G;
        q.pad = 0;
        rb += r.rbytes;
        hb += r.hbytes;
        nR += r.nR;
        nH += r.nH;
        G += r.G;
        nA += r.nA;
        cells += r.cells;
    }
    const phmm::RequestLayout L = phmm::request_layout(rb, hb, nR, nH, G);
    char* D = static_cast<char*>(c.buf.p);
    phmm::GatherDst dst{};
    dst.read_data = reinterpret_cast<uint8_t*>(D + L.read_data);
    dst.bq = reinterpret_cast<uint8_t*>(D + L.bq);
    dst.iq = reinterpret_cast<uint8_t*>(D + L.iq);
    dst.dq = reinterpret_cast<uint8_t*>(D + L.dq);
    dst.gcp = reinterpret_cast<uint8_t*>(D + L.gcp);
    dst.hap_data = reinterpret_cast<uint8_t*>(D + L.hap_data);
    dst.read_len = reinterpret_cast<int32_t*>(D + L.read_len);
    dst.hap_len = reinterpret_cast<int32_t*>(D + L.hap_len);
    dst.read_off = reinterpret_cast<int32_t*>(D + L.read_off);
    dst.hap_off = reinterpret_cast<int32_t*>(D + L.hap_off);
    dst.nreads = reinterpret_cast<int32_t*>(D + L.nreads);
    dst.nhaps = reinterpret_cast<int32_t*>(D + L.nhaps);
    dst.rprefix = reinterpret_cast<int32_t*>(D + L.rprefix);
    dst.hprefix = reinterpret_cast<int32_t*>(D + L.hprefix);
    dst.read_bytes = rb;
    dst.hap_bytes = hb;
    dst.nR = nR;
    dst.nH = nH;
    dst.G = G;
    phmm_batch_view v{};
    v.num_groups = G;
    v.num_reads = nR;
    v.num_haps = nH;
    v.read_bytes = rb;
    v.hap_bytes = hb;
    v.num_alignments = nA;
    v.read_data = dst.read_data;
    v.hap_data = dst.hap_data;
    v.base_quals = dst.bq;
    v.ins_quals = dst.iq;
    v.del_quals = dst.dq;
    v.gcp_quals = dst.gcp;
    v.read_lengths = dst.read_len;
    v.hap_lengths = dst.hap_len;
    v.read_offsets = dst.read_off;
    v.hap_offsets = dst.hap_off;
    v.num_reads_per_group = dst.nreads;
    v.num_haps_per_group = dst.nhaps;
    v.reads_prefix = dst.rprefix;
    v.haps_prefix = dst.hprefix;
    int rc = PHMM_OK;
    std::string msg;
    {
        std::lock_guard<std::mutex> elk(*eng_mu);
        hipSetDevice(eng->device);
        if ((size_t)L.total > c.buf.cap || (size_t)nA * sizeof(float) > c.scores.cap)
            rc = phmm::fail(PHMM_ERANGE, "server: merged launch exceeds its buffers");
        else if ((rc = check(hipStreamWaitEvent(c.stream, copy_mark, 0), "wait for staging")) ||
                 (rc = check(hipMemcpyAsync(c.d_reqs.p, c.h_reqs.p,
                                            sizeof(phmm::GatherReq) * reqs.size(),
                                            hipMemcpyHostToDevice, c.stream),
                             "merge list copy")) ||
                 (rc = check(phmm::launch_gather(static_cast<const phmm::GatherReq*>(c.d_reqs.p),
                                                 (int)reqs.size(), dst, c.stream),
                             "gather kernel")) ||
                 (rc = eng->score_device(v, static_cast<float*>(c.scores.p), c.stream, nullptr, 0,
                                         c.eng_ctx)) ||
                 (rc = check(hipMemcpyAsync(c.pin_scores.p, c.scores.p, sizeof(float) * (size_t)nA,
                                            hipMemcpyDeviceToHost, c.stream),
                             "score copy")))
            ;
        if (!rc)
            rc = check(hipEventRecord(c.done, c.stream), "event");
        if (rc)
            msg = phmm::last_error();
    }
    std::lock_guard<std::mutex> lk(mu);
    st.launches += 1;
    if (rc) {
        fail_requests(reqs, rc, msg);
        return;
    }
    st.cells += cells;
    c.reqs = std::move(reqs);
    c.n_scores = nA;
    c.busy = true;
    c.t_launch = Clock::now();
    int64_t running = 0;
    for (const LaunchCtx& x : ctxs)
        running += x.busy ? 1 : 0;
    st.max_concurrent = std::max(st.max_concurrent, running);
}

// Completes context c's launch if it finished (or waits for it when block).
bool phmm_server::retire(LaunchCtx& c, bool block)
{
    if (!c.busy)
        return false;
    const hipError_t q = block ? hipEventSynchronize(c.done) : hipEventQuery(c.done);
    if (q == hipErrorNotReady)
        return false;
    const auto now = Clock::now();
    std::lock_guard<std::mutex> lk(mu);
    c.busy = false;
    if (q != hipSuccess) {
        (void)hipGetLastError();
        fail_requests(c.reqs, PHMM_EDEVICE,
                      std::string("serving launch failed: ") + hipGetErrorString(q));
        c.reqs.clear();
        return true;
    }
    const float* sc = static_cast<const float*>(c.pin_scores.p);
    int64_t off = 0;
    for (Request& r : c.reqs) {
        Result res;
        res.tag = r.tag;
        res.scores.assign(sc + off, sc + off + r.nA);
        off += r.nA;
        res.latency = std::chrono::duration<double>(now - r.t_submit).count();
        arena_ring.release(r.off);
        st.inflight -= 1;
        st.completed += 1;
        done.push_back(std::move(res));
    }
    c.reqs.clear();
    st.bytes_in_use = fixed_bytes + (int64_t)arena_ring.live;
    cv.notify_all();
    return true;
}

void phmm_server::run()
{
    for (;;) {
        bool progressed = false;
        for (LaunchCtx& c : ctxs)
            progressed |= retire(c, false);
        std::vector<Request> batch;
        LaunchCtx* freec = nullptr;
        bool any_busy = false, need_flush = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            for (LaunchCtx& c : ctxs) {
                any_busy |= c.busy;
                if (!c.busy && !freec)
                    freec = &c;
            }
            if (stop && queue.empty() && !any_busy)
                return;
            const uint64_t copied = flushed_seq.load(std::memory_order_acquire);
            if (freec && !queue.empty() && queue.front().seq >= copied) {
                need_flush = true;  // the next request's copy is not issued yet
            } else if (freec && !queue.empty()) {
                // merge waiting requests in arrival order, up to the caps:
                // the latency cap unless the backlog says the server is
                // overloaded
                const int64_t cap =
                    opt.latency_batch_alignments > 0 && queued_na < 4 * opt.max_batch_alignments
                        ? std::min(opt.latency_batch_alignments, opt.max_batch_alignments)
                        : opt.max_batch_alignments;
                int64_t na = 0, bytes = 0, groups = 0;
                while (!queue.empty() && (int)batch.size() < opt.max_batch_requests) {
                    const Request& r = queue.front();
                    if (r.seq >= copied)
                        break;
                    if (!batch.empty() && (na + r.nA > cap ||
                                           bytes + (int64_t)r.bytes > opt.max_batch_bytes ||
                                           groups + r.G > max_groups))
                        break;
                    na += r.nA;
                    bytes += (int64_t)r.bytes;
                    groups += r.G;
                    queued_na -= r.nA;
                    batch.push_back(std::move(queue.front()));
                    queue.pop_front();
                }
                st.queued = (int64_t)queue.size();
                if (!copy_failed.empty()) {
                    // requests of a flush whose copies failed are not launched
                    std::vector<Request> bad, good;
                    for (Request& r : batch) {
                        bool failed = false;
                        for (const auto& fr : copy_failed)
                            failed |= r.seq >= fr.first && r.seq < fr.second;
                        (failed ? bad : good).push_back(std::move(r));
                    }
                    if (!bad.empty())
                        fail_requests(bad, copy_rc, copy_msg);
                    batch = std::move(good);
                    progressed = true;
                }
            } else if (!progressed) {
                if (!any_busy) {
                    cv.wait(lk, [&] { return stop || !queue.empty(); });
                    continue;
                }
                if (freec) {
                    // a context is free: wake for a new arrival or, after a
                    // short while, to retire finished launches
                    cv.wait_for(lk, std::chrono::microseconds(20),
                                [&] { return stop || !queue.empty(); });
                    continue;
                }
            }
        }
        if (need_flush) {
            std::lock_guard<std::mutex> sl(submit_mu);
            reclaim_stage(false);
            flush_copies();
            continue;
        }
        if (!batch.empty()) {
            launch(*freec, std::move(batch));
            continue;
        }
        if (!progressed && !freec) {
            // every context busy: nothing can launch before one finishes
            LaunchCtx* oldest = nullptr;
            for (LaunchCtx& c : ctxs)
                if (c.busy && (!oldest || c.t_launch < oldest->t_launch))
                    oldest = &c;
            if (oldest)
                retire(*oldest, true);
        }
    }
}

extern "C" void phmm_server_options_default(phmm_server_options* o)
{
    if (!o)
        return;
    std::memset(o, 0, sizeof *o);
    o->hbm_budget_bytes = (int64_t)64 << 30;  // 64 GiB of request state
    o->max_inflight = 1 << 20;
    o->max_batch_requests = 4096;
    o->max_batch_alignments = 1 << 20;
    o->max_batch_bytes = (int64_t)64 << 20;
    o->streams = 2;
    o->latency_batch_alignments = 1 << 19;
}

extern "C" int phmm_server_create(phmm_engine* e, const phmm_server_options* opt,
                                  phmm_server** out)
{
    if (!e || !out)
        return phmm::fail(PHMM_EINVAL, "phmm_server_create: null argument");
    *out = nullptr;
    phmm_server_options o;
    if (opt)
        o = *opt;
    else
        phmm_server_options_default(&o);
    if (o.hbm_budget_bytes <= 0 || o.max_inflight <= 0 || o.max_batch_requests <= 0 ||
        o.max_batch_alignments <= 0 || o.max_batch_alignments > INT32_MAX ||
        o.max_batch_bytes < 4096 || o.max_batch_bytes > ((int64_t)1 << 31) || o.streams < 1 ||
        o.streams >= phmm::Engine::kMaxCtx)
        return phmm::fail(PHMM_EINVAL,
                          "phmm_server_create: budget, limits and caps must be > 0, "
                          "max_batch_bytes 4 KiB .. 2 GiB, streams 1 .. " +
                              std::to_string(phmm::Engine::kMaxCtx - 1));
    auto* s = new (std::nothrow) phmm_server;
    if (!s)
        return phmm::fail(PHMM_ENOMEM, "phmm_server_create: out of memory");
    s->handle = e;
    s->eng = phmm::engine_of(e, &s->eng_mu);
    s->opt = o;
    const int rc = s->init();
    if (rc) {
        const std::string m = phmm::last_error();
        delete s;
        return phmm::fail(rc, m);
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return PHMM_OK;
}

extern "C" void phmm_server_destroy(phmm_server* s)
{
    if (!s)
        return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    // the worker launches what is queued, retires it and exits
    if (s->worker.joinable())
        s->worker.join();
    delete s;
}

extern "C" int phmm_server_submit(phmm_server* s, const phmm_batch_view* req, uint64_t tag)
{
    if (!s || !req)
        return phmm::fail(PHMM_EINVAL, "phmm_server_submit: null argument");
    const std::string verr = phmm::validate_host_view(*req);
    if (!verr.empty())
        return phmm::fail(PHMM_EINVAL, "phmm_server_submit: " + verr);
    const phmm_batch_view& v = *req;
    Request r;
    r.tag = tag;
    r.nA = v.num_alignments;
    r.cells = phmm_batch_cells(&v);
    r.G = v.num_groups;
    r.nR = v.num_reads;
    r.nH = v.num_haps;
    r.rbytes = v.read_bytes;
    r.hbytes = v.hap_bytes;
    r.lay = phmm::request_layout(v.read_bytes, v.hap_bytes, v.num_reads, v.num_haps, v.num_groups);
    r.bytes = phmm::round_up((size_t)r.lay.total);
    if (r.nA > s->opt.max_batch_alignments || (int64_t)r.bytes > s->opt.max_batch_bytes ||
        r.G > s->max_groups)
        return phmm::fail(PHMM_ERANGE,
                          "phmm_server_submit: request larger than one launch (alignments, "
                          "max_batch_bytes or groups)");
    std::lock_guard<std::mutex> sl(s->submit_mu);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop)
            return phmm::fail(PHMM_EINVAL, "phmm_server_submit: server stopping");
        if (s->st.inflight >= s->opt.max_inflight || !s->arena_ring.alloc(r.bytes, &r.off)) {
            s->st.refused += 1;
            return phmm::fail(PHMM_EBUSY,
                              "phmm_server_submit: HBM budget or in-flight limit reached");
        }
        s->st.inflight += 1;
        s->st.bytes_in_use = s->fixed_bytes + (int64_t)s->arena_ring.live;
    }
    // stage the image in pinned memory; a flush copies it into the arena
    size_t soff = 0;
    s->reclaim_stage(false);
    while (!s->stage_ring.alloc(r.bytes, &soff)) {
        if (!s->unflushed.empty()) {
            s->flush_copies();  // staged blocks return only once copied
            continue;
        }
        if (s->stage_pending.empty()) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->arena_ring.release(r.off);
            s->st.inflight -= 1;
            s->st.bytes_in_use = s->fixed_bytes + (int64_t)s->arena_ring.live;
            return phmm::fail(PHMM_ERANGE, "phmm_server_submit: request larger than the staging");
        }
        s->reclaim_stage(true);
    }
    char* P = static_cast<char*>(s->stage.p) + soff;
    const phmm::RequestLayout& L = r.lay;
    const size_t rb = (size_t)v.read_bytes, hb = (size_t)v.hap_bytes;
    std::memcpy(P + L.read_data, v.read_data, rb);
    std::memcpy(P + L.bq, v.base_quals, rb);
    std::memcpy(P + L.iq, v.ins_quals, rb);
    std::memcpy(P + L.dq, v.del_quals, rb);
    std::memcpy(P + L.gcp, v.gcp_quals, rb);
    std::memcpy(P + L.hap_data, v.hap_data, hb);
    std::memcpy(P + L.read_len, v.read_lengths, 4 * (size_t)v.num_reads);
    std::memcpy(P + L.hap_len, v.hap_lengths, 4 * (size_t)v.num_haps);
    std::memcpy(P + L.read_off, v.read_offsets, 4 * (size_t)v.num_reads);
    std::memcpy(P + L.hap_off, v.hap_offsets, 4 * (size_t)v.num_haps);
    std::memcpy(P + L.nreads, v.num_reads_per_group, 4 * (size_t)v.num_groups);
    std::memcpy(P + L.nhaps, v.num_haps_per_group, 4 * (size_t)v.num_groups);
    std::memcpy(P + L.rprefix, v.reads_prefix, 4 * (size_t)(v.num_groups + 1));
    std::memcpy(P + L.hprefix, v.haps_prefix, 4 * (size_t)(v.num_groups + 1));
    r.seq = s->next_seq++;
    s->unflushed.push_back({soff, r.off, r.bytes});
    s->unflushed_bytes += r.bytes;
    r.t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->queued_na += r.nA;
        s->queue.push_back(std::move(r));
        s->st.submitted += 1;
        s->st.queued = (int64_t)s->queue.size();
        s->st.max_queue_depth = std::max<int64_t>(s->st.max_queue_depth, s->st.queued);
    }
    // copy at once when no flush is in flight; otherwise requests gather
    // behind the one in flight (the worker flushes what it is about to
    // merge) until kFlushBytes wait
    s->reclaim_stage(false);
    if (s->stage_pending.empty() || s->unflushed_bytes >= phmm::kFlushBytes)
        s->flush_copies();
    s->cv.notify_all();
    return PHMM_OK;
}

extern "C" int phmm_server_poll(phmm_server* s, int timeout_ms, uint64_t* tag, float* scores,
                                int64_t capacity, int64_t* n_scores, double* latency_seconds)
{
    if (!s || !tag || !n_scores)
        return phmm::fail(PHMM_EINVAL, "phmm_server_poll: null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty()) {
        if (timeout_ms > 0)
            s->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s->done.empty(); });
        else if (timeout_ms < 0)
            s->cv.wait(lk, [&] { return !s->done.empty(); });
    }
    if (s->done.empty())
        return PHMM_EAGAIN;
    Result& r = s->done.front();
    *n_scores = (int64_t)r.scores.size();
    *tag = r.tag;
    if (capacity < (int64_t)r.scores.size() || (!scores && !r.scores.empty()))
        return phmm::fail(PHMM_ERANGE, "phmm_server_poll: score buffer too small");
    if (!r.scores.empty())
        std::memcpy(scores, r.scores.data(), sizeof(float) * r.scores.size());
    if (latency_seconds)
        *latency_seconds = r.latency;
    // the status of this request's own launch (a failed launch fails only
    // the requests it held)
    const int rc = r.rc;
    const std::string msg = r.msg;
    s->done.pop_front();
    if (rc != PHMM_OK)
        return phmm::fail(rc, msg);
    return PHMM_OK;
}

extern "C" int phmm_server_stats_get(const phmm_server* s, phmm_server_stats* out)
{
    if (!s || !out)
        return phmm::fail(PHMM_EINVAL, "phmm_server_stats_get: null argument");
    auto* m = const_cast<phmm_server*>(s);
    std::lock_guard<std::mutex> lk(m->mu);
    *out = s->st;
    return PHMM_OK;
}
