// gpuPairHMM -- command-line drop-in for the reference tool (main.cu:30-94).
//
// Same flags and semantics: --inputfile, --outputfile, --fileBatchsize,
// --transferBatchsize (<= 0: unlimited), --verbose, --queue_depth,
// --checkResults (accepted, as in the reference it has no effect).  Extra
// flags: --mode plain|batched, --semantics gpu|cpu|gcp, --no-rescue.
//
// Three threads joined by bounded queues, like the reference pipeline
// (main.cu:7-27): parser -> engine -> writer.  Scores are written as the
// reference writes them (output_writer_worker.cu:29-159); /dev/null skips
// formatting.  On exit it prints "Batch processing time: Xs (Y GCUPS)" over
// the engine time of all batches (gpuworker.cuh:418-428).
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <fstream>
#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gpupairhmm.h"
#include "host_batch.hpp"

namespace {

struct Options {
    bool checkResults = false;
    bool verbose = false;
    long long transferBatchsize = 1000000;  // common.cuh:19
    long long fileBatchsize = 10000000;     // common.cuh:20
    int queue_depth = 3;
    std::string inputfile, outputfile;
    int mode = PHMM_MODE_BATCHED;
    int semantics = PHMM_SEM_GPU;
    bool rescue = true;
};

struct Work {
    phmm::HostBatch batch;
    std::vector<float> scores;
};

template <class T>
class Queue {
    std::deque<T> q_;
    std::mutex m_;
    std::condition_variable cv_;

public:
    void push(T v)
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            q_.push_back(std::move(v));
        }
        cv_.notify_one();
    }
    T pop()
    {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return !q_.empty(); });
        T v = std::move(q_.front());
        q_.pop_front();
        return v;
    }
};

[[noreturn]] void die(const std::string& m)
{
    std::cerr << m << "\n";
    std::exit(EXIT_FAILURE);
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    for (int x = 1; x < argc; x++) {
        const std::string a = argv[x];
        auto next = [&]() -> const char* {
            if (x + 1 >= argc)
                die("missing value for " + a);
            return argv[++x];
        };
        if (a == "--inputfile")
            o.inputfile = next();
        else if (a == "--outputfile")
            o.outputfile = next();
        else if (a == "--transferBatchsize")
            o.transferBatchsize = std::atoll(next());
        else if (a == "--fileBatchsize")
            o.fileBatchsize = std::atoll(next());
        else if (a == "--queue_depth")
            o.queue_depth = std::atoi(next());
        else if (a == "--checkResults")
            o.checkResults = true;
        else if (a == "--verbose")
            o.verbose = true;
        else if (a == "--mode") {
            const std::string m = next();
            if (m == "plain")
                o.mode = PHMM_MODE_PLAIN;
            else if (m == "batched")
                o.mode = PHMM_MODE_BATCHED;
            else
                die("--mode must be plain or batched");
        } else if (a == "--semantics") {
            const std::string m = next();
            if (m == "gpu")
                o.semantics = PHMM_SEM_GPU;
            else if (m == "cpu")
                o.semantics = PHMM_SEM_CPU;
            else if (m == "gcp")
                o.semantics = PHMM_SEM_GCP;
            else
                die("--semantics must be gpu, cpu or gcp");
        } else if (a == "--no-rescue")
            o.rescue = false;
    }
    if (o.queue_depth < 1)
        o.queue_depth = 1;
    std::cout << "options.inputfile = " << o.inputfile << "\n";
    std::cout << "options.outputfile = " << o.outputfile << "\n";
    std::cout << "options.transferBatchsize = " << o.transferBatchsize << "\n";
    std::cout << "options.verbose = " << o.verbose << "\n";
    std::cout << "options.fileBatchsize = " << o.fileBatchsize << "\n";

    std::ifstream in(o.inputfile);
    if (!in)
        die("Could not open file" + o.inputfile);
    std::ofstream out(o.outputfile);
    if (!out)
        die("Could not open file" + o.outputfile);
    const bool no_output = (o.outputfile == "/dev/null");

    phmm_options eo;
    phmm_options_default(&eo);
    eo.transfer_batchsize = o.transferBatchsize;
    eo.mode = o.mode;
    eo.semantics = o.semantics;
    eo.fp64_rescue = o.rescue ? 1 : 0;
    eo.verbose = o.verbose;
    phmm_engine* eng = nullptr;
    if (phmm_engine_create(0, &eo, &eng) != PHMM_OK)
        die(std::string("engine: ") + phmm_last_error());

    Queue<std::unique_ptr<Work>> free_q, to_gpu, to_writer;
    for (int i = 0; i < o.queue_depth; ++i)
        free_q.push(std::make_unique<Work>());

    std::thread parser([&] {
        const long long fb = o.fileBatchsize > 0 ? o.fileBatchsize : (1LL << 62);
        std::vector<phmm::SequenceGroup> groups;
        long long n_align = 0;
        int64_t line_no = 0;
        std::string line;
        try {
            while (std::getline(in, line)) {
                ++line_no;
                if (line.empty())
                    continue;
                groups.push_back(phmm::parse_group(in, line, &line_no));
                n_align += (long long)groups.back().num_reads * groups.back().num_haps;
                if (n_align >= fb) {
                    auto w = free_q.pop();
                    w->batch.assign(groups);
                    groups.clear();
                    n_align = 0;
                    to_gpu.push(std::move(w));
                }
            }
            if (!groups.empty()) {
                auto w = free_q.pop();
                w->batch.assign(groups);
                to_gpu.push(std::move(w));
            }
        } catch (const std::exception& e) {
            die(e.what());
        }
        to_gpu.push(nullptr);
    });

    std::thread worker([&] {
        for (;;) {
            auto w = to_gpu.pop();
            if (!w)
                break;
            const phmm_batch_view v = w->batch.view();
            w->scores.assign((size_t)v.num_alignments, 0.f);
            if (phmm_score_batch(eng, &v, w->scores.data()) != PHMM_OK)
                die(std::string("gpu worker: ") + phmm_engine_last_error(eng));
            to_writer.push(std::move(w));
        }
        to_writer.push(nullptr);
    });

    std::thread writer([&] {
        std::string buf;
        for (;;) {
            auto w = to_writer.pop();
            if (!w)
                break;
            if (!no_output) {
                buf.clear();
                phmm::format_scores(w->batch.view(), w->scores.data(), buf);
                out.write(buf.data(), (std::streamsize)buf.size());
                if (!out)
                    die("error writing " + o.outputfile);
            }
            w->batch.clear();
            w->scores.clear();
            free_q.push(std::move(w));
        }
        out.flush();
    });

    parser.join();
    worker.join();
    writer.join();

    phmm_stats st;
    phmm_engine_stats(eng, &st);
    const double gcups = st.engine_seconds > 0 ? (double)st.cells / 1e9 / st.engine_seconds : 0.0;
    std::cout << "Batch processing time: " << st.engine_seconds << "s (" << gcups << " GCUPS)\n";
    if (o.verbose)
        std::cout << "alignments " << st.alignments << ", FP64 rescued " << st.rescued << "\n";
    phmm_engine_destroy(eng);
    return 0;
}
