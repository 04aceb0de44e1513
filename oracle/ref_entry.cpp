// C entry points around the reference's own CPU forward algorithm.
//
// Test infrastructure only.  oracle/Makefile compiles `align_host` and
// `generate_ph2pr` from the reference's source where it lies
// (gpuworker.cuh:473-486 and 1022-1098: pure ISO C++; the rest of that header
// needs the CUDA toolkit and is not built) into oracle/_ref/ref_funcs.o, and
// links it with this file.  This file holds only declarations of those two
// functions and extern "C" wrappers that tests call through ctypes.
#include <cstdint>
#include <vector>

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

// One alignment through the reference's align_host with its own table.
double ref_align_host(const uint8_t* hap, int hap_len, const uint8_t* read, int read_len,
                      const uint8_t* base_q, const uint8_t* ins_q, const uint8_t* del_q,
                      uint8_t gcp0)
{
    static const std::vector<float> table = generate_ph2pr(128);
    return align_host(hap, hap_len, read, read_len, base_q, ins_q, del_q, gcp0, table.data());
}
}
