// Serving loop: continuous batching of active-region requests.
//
// Requests (each a small host batch, usually one region = one record) are
// admitted while their estimated device footprint fits the HBM budget and the
// in-flight limit; otherwise submit returns PHMM_EBUSY.  A worker thread
// merges every admitted request that is waiting into one batch and enqueues
// it (copy-in, binning, forward kernels, FP64 rescue, copy-out) into one of
// two device slots, so a new merged batch is already queued on the GPU while
// the previous one runs; requests that arrive meanwhile join the next merge.
// When a slot's batch completes its scores are split back per request and
// the request's budget is released.  Every alignment is scored by a kernel
// chosen from its own read length and coefficients only, so a request's
// scores are the same alone and in any mix.
#include <chrono>
#include <climits>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

#include "common.hpp"
#include "engine.hpp"

namespace phmm {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t tag = 0;
    HostBatch batch;
    int64_t bytes = 0;
    Clock::time_point t_submit;
};

struct Result {
    uint64_t tag = 0;
    std::vector<float> scores;
    double latency = 0.0;
};

// Device bytes a request occupies while admitted (staging copy, scores and
// the engine's per-alignment work state).
int64_t request_bytes(const phmm_batch_view& v)
{
    return 9 * v.read_bytes + 2 * v.hap_bytes + 24 * v.num_alignments +
           24 * (int64_t)(v.num_reads + v.num_haps) + 64 * (int64_t)v.num_groups + 4096;
}

}  // namespace
}  // namespace phmm

using phmm::Clock;
using phmm::Request;
using phmm::Result;


struct phmm_server {
    phmm_engine* handle = nullptr;
    phmm::Engine* eng = nullptr;
    std::mutex* eng_mu = nullptr;
    phmm_server_options opt{};
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request> queue;
    std::deque<Result> done;
    struct Launch {
        std::vector<Request> reqs;
        int slot = 0;
    };
    std::deque<Launch> running;
    phmm::Engine::Slot slots[2];
    bool slot_busy[2] = {false, false};
    bool stop = false;
    int error = PHMM_OK;
    std::string error_msg;
    phmm_server_stats st{};
    std::thread worker;

    void run();
    bool complete_front(bool block);
};

// Retires the oldest running launch if it finished (or waits when block).
bool phmm_server::complete_front(bool block)
{
    Launch& L = running.front();
    phmm::Engine::Slot& s = slots[L.slot];
    hipError_t q = block ? hipEventSynchronize(s.compute_done) : hipEventQuery(s.compute_done);
    if (q == hipErrorNotReady)
        return false;
    const auto now = Clock::now();
    std::unique_lock<std::mutex> lk(mu);
    if (q != hipSuccess) {
        hipGetLastError();
        error = PHMM_EDEVICE;
        error_msg = std::string("serving launch failed: ") + hipGetErrorString(q);
    }
    const float* sc = static_cast<const float*>(s.pin_scores.p);
    int64_t off = 0;
    for (Request& r : L.reqs) {
        Result res;
        res.tag = r.tag;
        const int64_t n = r.batch.num_alignments;
        res.scores.assign(sc + off, sc + off + n);
        off += n;
        res.latency = std::chrono::duration<double>(now - r.t_submit).count();
        st.bytes_in_use -= r.bytes;
        st.inflight -= 1;
        st.completed += 1;
        done.push_back(std::move(res));
    }
    slot_busy[L.slot] = false;
    running.pop_front();
    cv.notify_all();
    return true;
}

void phmm_server::run()
{
    for (;;) {
        std::vector<Request> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            const bool can_launch = !queue.empty() && running.size() < 2;
            if (!can_launch && running.empty()) {
                cv.wait(lk, [&] { return stop || (!queue.empty() && running.size() < 2); });
                if (stop && queue.empty())
                    return;
            }
            if (!queue.empty() && running.size() < 2) {
                // merge waiting requests, up to the per-launch limits (at
                // least one request), so the next launch can be staged on
                // the host while this one computes
                const int cap = opt.max_batch_requests > 0 ? opt.max_batch_requests : 1 << 30;
                const int64_t acap =
                    opt.max_batch_alignments > 0 ? opt.max_batch_alignments : INT64_MAX;
                int64_t na = 0;
                while (!queue.empty() && (int)batch.size() < cap &&
                       (batch.empty() || na + queue.front().batch.num_alignments <= acap)) {
                    na += queue.front().batch.num_alignments;
                    batch.push_back(std::move(queue.front()));
                    queue.pop_front();
                }
                st.queued = (int64_t)queue.size();
            }
        }
        if (!batch.empty()) {
            Launch L;
            L.slot = slot_busy[0] ? 1 : 0;
            L.reqs = std::move(batch);
            // the requests are staged straight into the slot (no merged copy)
            std::vector<phmm_batch_view> views;
            std::vector<phmm::Engine::ChunkPart> parts;
            views.reserve(L.reqs.size());
            int64_t na = 0, cells = 0;
            for (auto& r : L.reqs) {
                views.push_back(r.batch.view());
                na += r.batch.num_alignments;
                cells += phmm_batch_cells(&views.back());
            }
            for (auto& v : views)
                parts.push_back({&v, 0, v.num_groups});
            int rc = PHMM_OK;
            {
                std::lock_guard<std::mutex> elk(*eng_mu);
                hipSetDevice(eng->device);
                if (na > 0)
                    rc = eng->launch_parts(slots[L.slot], parts.data(), (int)parts.size());
                else
                    rc = eng->check(hipEventRecord(slots[L.slot].compute_done, eng->stream), "event");
            }
            std::unique_lock<std::mutex> lk(mu);
            st.launches += 1;
            if (rc == PHMM_OK)
                st.cells += cells;
            if (rc != PHMM_OK) {
                error = rc;
                error_msg = eng->err;
                // fail the requests: return empty results so callers unblock
                for (Request& r : L.reqs) {
                    Result res;
                    res.tag = r.tag;
                    res.scores.assign((size_t)r.batch.num_alignments, __builtin_nanf(""));
                    st.bytes_in_use -= r.bytes;
                    st.inflight -= 1;
                    st.completed += 1;
                    done.push_back(std::move(res));
                }
                cv.notify_all();
                continue;
            }
            slot_busy[L.slot] = true;
            running.push_back(std::move(L));
            continue;
        }
        // nothing to launch now: retire finished work, briefly waiting for
        // new arrivals when both slots are not busy
        if (!running.empty()) {
            if (complete_front(false))
                continue;
            std::unique_lock<std::mutex> lk(mu);
            if (running.size() < 2 && queue.empty())
                cv.wait_for(lk, std::chrono::microseconds(50),
                            [&] { return stop || !queue.empty(); });
            else if (running.size() >= 2) {
                lk.unlock();
                complete_front(true);
            }
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
    o->max_batch_alignments = 4 << 20;
}

extern "C" int phmm_server_create(phmm_engine* e, const phmm_server_options* opt,
                                  phmm_server** out)
{
    if (!e || !out)
        return phmm::fail(PHMM_EINVAL, "phmm_server_create: null argument");
    auto* s = new (std::nothrow) phmm_server;
    if (!s)
        return phmm::fail(PHMM_ENOMEM, "phmm_server_create: out of memory");
    s->handle = e;
    s->eng = phmm::engine_of(e, &s->eng_mu);
    if (opt)
        s->opt = *opt;
    else
        phmm_server_options_default(&s->opt);
    if (s->opt.hbm_budget_bytes <= 0 || s->opt.max_inflight <= 0) {
        delete s;
        return phmm::fail(PHMM_EINVAL, "phmm_server_create: budget and max_inflight must be > 0");
    }
    for (auto& sl : s->slots) {
        if (hipEventCreateWithFlags(&sl.h2d_done, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&sl.compute_done, hipEventDisableTiming) != hipSuccess) {
            hipGetLastError();
            delete s;
            return phmm::fail(PHMM_EDEVICE, "phmm_server_create: event creation failed");
        }
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
    // drain: the worker exits once the queue is empty; retire running work
    if (s->worker.joinable())
        s->worker.join();
    while (!s->running.empty())
        s->complete_front(true);
    for (auto& sl : s->slots)
        s->eng->free_slot(sl);
    delete s;
}

extern "C" int phmm_server_submit(phmm_server* s, const phmm_batch_view* req, uint64_t tag)
{
    if (!s || !req)
        return phmm::fail(PHMM_EINVAL, "phmm_server_submit: null argument");
    const std::string verr = phmm::validate_host_view(*req);
    if (!verr.empty())
        return phmm::fail(PHMM_EINVAL, "phmm_server_submit: " + verr);
    const int64_t need = phmm::request_bytes(*req);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop)
            return phmm::fail(PHMM_EINVAL, "phmm_server_submit: server stopping");
        if (need > s->opt.hbm_budget_bytes)
            return phmm::fail(PHMM_ERANGE, "phmm_server_submit: request larger than the HBM budget");
        if (s->st.bytes_in_use + need > s->opt.hbm_budget_bytes ||
            s->st.inflight >= s->opt.max_inflight) {
            s->st.refused += 1;
            return phmm::fail(PHMM_EBUSY, "phmm_server_submit: HBM budget or in-flight limit reached");
        }
        s->st.bytes_in_use += need;
        s->st.inflight += 1;
    }
    Request r;
    r.tag = tag;
    r.bytes = need;
    r.t_submit = Clock::now();
    // copy the request (rebased to a self-contained batch)
    {
        phmm::HostBatch& b = r.batch;
        const phmm_batch_view& v = *req;
        b.read_data.assign(v.read_data, v.read_data + v.read_bytes);
        b.base_quals.assign(v.base_quals, v.base_quals + v.read_bytes);
        b.ins_quals.assign(v.ins_quals, v.ins_quals + v.read_bytes);
        b.del_quals.assign(v.del_quals, v.del_quals + v.read_bytes);
        b.gcp_quals.assign(v.gcp_quals, v.gcp_quals + v.read_bytes);
        b.hap_data.assign(v.hap_data, v.hap_data + v.hap_bytes);
        b.read_lengths.assign(v.read_lengths, v.read_lengths + v.num_reads);
        b.hap_lengths.assign(v.hap_lengths, v.hap_lengths + v.num_haps);
        b.read_offsets.assign(v.read_offsets, v.read_offsets + v.num_reads + 1);
        b.hap_offsets.assign(v.hap_offsets, v.hap_offsets + v.num_haps + 1);
        b.read_offsets[(size_t)v.num_reads] = (int32_t)v.read_bytes;
        b.hap_offsets[(size_t)v.num_haps] = (int32_t)v.hap_bytes;
        b.num_reads.assign(v.num_reads_per_group, v.num_reads_per_group + v.num_groups);
        b.num_haps.assign(v.num_haps_per_group, v.num_haps_per_group + v.num_groups);
        b.reads_prefix.assign(v.reads_prefix, v.reads_prefix + v.num_groups + 1);
        b.haps_prefix.assign(v.haps_prefix, v.haps_prefix + v.num_groups + 1);
        b.num_alignments = v.num_alignments;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->queue.push_back(std::move(r));
        s->st.submitted += 1;
        s->st.queued = (int64_t)s->queue.size();
        s->st.max_queue_depth = std::max<int64_t>(s->st.max_queue_depth, s->st.queued);
    }
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
    s->done.pop_front();
    if (s->error != PHMM_OK) {
        const int e = s->error;
        return phmm::fail(e, s->error_msg);
    }
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
