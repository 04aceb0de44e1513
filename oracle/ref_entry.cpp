// C entry points around the reference's own CPU forward algorithm.
//
// Test infrastructure only.  oracle/Makefile compiles `align_host` and
// `generate_ph2pr` from the reference's source where it lies
// (gpuworker.cuh:473-486 and 1022-1098: pure ISO C++; the rest of that header
// needs the CUDA toolkit and is not built) into oracle/_ref/ref_funcs.o, and
// links it with this file.  This file holds only declarations of those two
// functions and extern "C" wrappers that tests call through ctypes.
#include <omp.h>

#include <cstddef>
#include <cstdint>
#include <vector>

using std::size_t;

std::vector<float> generate_ph2pr(int max_ph2pr);
double align_host(const uint8_t* hap_bases, const int hap_length, const uint8_t* read_bases,
                  const int read_length, const uint8_t* base_quals, const uint8_t* ins_quals,
                  const uint8_t* del_quals, const uint8_t gcp_qual, const float* ph2pr);

extern "C" {

// Writes the reference's 128-entry ph2pr table.
void ref_ph2pr(float* out)
{
    const std::vector<float> t = generate_ph2pr(128);
    for (int i = 0; i < 128; ++i)
        out[i] = t[i];
}

// A batch in the reference's layout, scored like processBatchCPU
// (gpuworker.cuh:975-1020): OpenMP over alignments, schedule(dynamic, 16),
// results read-major per group.  `which` lists global alignment indices
// (null: all); out[i] receives alignment which[i] (or i).
int ref_score_batch(int G, const int* nR, const int* nH, const int* rp, const int* hp,
                    const uint8_t* read_data, const uint8_t* hap_data, const uint8_t* bq,
                    const uint8_t* iq, const uint8_t* dq, const uint8_t* gq, const int* rl,
                    const int* hl, const int* ro, const int* ho, const int64_t* which,
                    int64_t n, int nthreads, double* out)
{
    static const std::vector<float> table = generate_ph2pr(128);
    std::vector<int64_t> ap((size_t)G + 1, 0);
    for (int g = 0; g < G; ++g)
        ap[(size_t)g + 1] = ap[(size_t)g] + (int64_t)nR[g] * nH[g];
    if (!which)
        n = ap[(size_t)G];
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i) {
        const int64_t a = which ? which[i] : i;
        int lo = 0, hi = G;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (ap[(size_t)mid + 1] > a)
                hi = mid;
            else
                lo = mid + 1;
        }
        const int64_t local = a - ap[(size_t)lo];
        const int h = (int)(local % nH[lo]), r = (int)(local / nH[lo]);
        const int read = rp[lo] + r, hap = hp[lo] + h;
        out[i] = align_host(hap_data + ho[hap], hl[hap], read_data + ro[read], rl[read],
                            bq + ro[read], iq + ro[read], dq + ro[read], gq[ro[read]],
                            table.data());
    }
    return 0;
}

// One alignment through the reference's align_host with its own table.
double ref_align_host(const uint8_t* hap, int hap_len, const uint8_t* read, int read_len,
                      const uint8_t* base_q, const uint8_t* ins_q, const uint8_t* del_q,
                      uint8_t gcp0)
{
    static const std::vector<float> table = generate_ph2pr(128);
    return align_host(hap, hap_len, read, read_len, base_q, ins_q, del_q, gcp0, table.data());
}
}
