// Device-side merge of serving requests.  Every admitted request sits in the
// server's device arena as a self-contained batch image (its arrays at the
// offsets of RequestLayout, offsets relative to the request).  A launch
// gathers the requests it merges into one self-contained batch in the
// launch buffer -- sequence and quality bytes appended, offsets and group
// prefixes rebased -- so the batch never returns through host memory.
#include "gather.hpp"

namespace phmm {

// Block (x = request, y = part) copies its part of the request's dwords and
// rebases its index arrays; x = 0, y = 0 also writes the merged arrays'
// closing entries.
__global__ __launch_bounds__(256) void k_gather(const GatherReq* reqs, GatherDst dst)
{
    const int k = blockIdx.x;
    const GatherReq q = reqs[k];
    const int nthr = (int)(blockDim.x * gridDim.y);
    const int tid = (int)(blockIdx.y * blockDim.x + threadIdx.x);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(q.src);
    // sequence and quality bytes: dword copies (padded lengths are multiples
    // of 4, and every array starts 4-byte aligned)
    const int64_t rw = q.rbytes >> 2, hw = q.hbytes >> 2;
    const int64_t rdst = q.rb_dst >> 2, hdst = q.hb_dst >> 2;
    uint32_t* const rd_dst[5] = {reinterpret_cast<uint32_t*>(dst.read_data),
                                 reinterpret_cast<uint32_t*>(dst.bq),
                                 reinterpret_cast<uint32_t*>(dst.iq),
                                 reinterpret_cast<uint32_t*>(dst.dq),
                                 reinterpret_cast<uint32_t*>(dst.gcp)};
    const int64_t rd_src[5] = {q.lay.read_data, q.lay.bq, q.lay.iq, q.lay.dq, q.lay.gcp};
#pragma unroll
    for (int a = 0; a < 5; ++a) {
        const uint32_t* s = src + (rd_src[a] >> 2);
        for (int64_t i = tid; i < rw; i += nthr)
            rd_dst[a][rdst + i] = s[i];
    }
    {
        const uint32_t* s = src + (q.lay.hap_data >> 2);
        uint32_t* d = reinterpret_cast<uint32_t*>(dst.hap_data);
        for (int64_t i = tid; i < hw; i += nthr)
            d[hdst + i] = s[i];
    }
    const int32_t* rl = reinterpret_cast<const int32_t*>(q.src + q.lay.read_len);
    const int32_t* hl = reinterpret_cast<const int32_t*>(q.src + q.lay.hap_len);
    const int32_t* ro = reinterpret_cast<const int32_t*>(q.src + q.lay.read_off);
    const int32_t* ho = reinterpret_cast<const int32_t*>(q.src + q.lay.hap_off);
    const int32_t* nr = reinterpret_cast<const int32_t*>(q.src + q.lay.nreads);
    const int32_t* nh = reinterpret_cast<const int32_t*>(q.src + q.lay.nhaps);
    const int32_t* rp = reinterpret_cast<const int32_t*>(q.src + q.lay.rprefix);
    const int32_t* hp = reinterpret_cast<const int32_t*>(q.src + q.lay.hprefix);
    for (int i = tid; i < q.nR; i += nthr) {
        dst.read_len[q.r_dst + i] = rl[i];
        dst.read_off[q.r_dst + i] = ro[i] + (int32_t)q.rb_dst;
    }
    for (int i = tid; i < q.nH; i += nthr) {
        dst.hap_len[q.h_dst + i] = hl[i];
        dst.hap_off[q.h_dst + i] = ho[i] + (int32_t)q.hb_dst;
    }
    for (int i = tid; i < q.G; i += nthr) {
        dst.nreads[q.g_dst + i] = nr[i];
        dst.nhaps[q.g_dst + i] = nh[i];
        dst.rprefix[q.g_dst + i] = rp[i] + q.r_dst;
        dst.hprefix[q.g_dst + i] = hp[i] + q.h_dst;
    }
    if (k == 0 && tid == 0) {
        dst.read_off[dst.nR] = (int32_t)dst.read_bytes;
        dst.hap_off[dst.nH] = (int32_t)dst.hap_bytes;
        dst.rprefix[dst.G] = dst.nR;
        dst.hprefix[dst.G] = dst.nH;
    }
}

hipError_t launch_gather(const GatherReq* d_reqs, int nreq, const GatherDst& dst, hipStream_t st)
{
    if (nreq <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)nreq, 8), dim3(256), 0, st, d_reqs, dst);
    return hipGetLastError();
}

}  // namespace phmm
