// Batch images of the serving loop: how a request sits in the device arena,
// and the device-side merge of several requests into one launch (gather.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gpupairhmm.h"

namespace phmm {

// Byte offsets of a self-contained batch's arrays in one contiguous image
// (each array 256-byte aligned): the five read arrays of read_bytes each,
// the hap bytes, then the index arrays.  `total` is the image size.
struct RequestLayout {
    int64_t read_data, bq, iq, dq, gcp, hap_data;
    int64_t read_len, hap_len, read_off, hap_off, nreads, nhaps, rprefix, hprefix;
    int64_t total;
};

inline RequestLayout request_layout(int64_t read_bytes, int64_t hap_bytes, int64_t nR, int64_t nH,
                                    int64_t G)
{
    RequestLayout L{};
    int64_t o = 0;
    auto take = [&](int64_t bytes) {
        const int64_t at = o;
        o = (o + bytes + 255) & ~(int64_t)255;
        return at;
    };
    L.read_data = take(read_bytes);
    L.bq = take(read_bytes);
    L.iq = take(read_bytes);
    L.dq = take(read_bytes);
    L.gcp = take(read_bytes);
    L.hap_data = take(hap_bytes);
    L.read_len = take(4 * nR);
    L.hap_len = take(4 * nH);
    L.read_off = take(4 * (nR + 1));
    L.hap_off = take(4 * (nH + 1));
    L.nreads = take(4 * G);
    L.nhaps = take(4 * G);
    L.rprefix = take(4 * (G + 1));
    L.hprefix = take(4 * (G + 1));
    L.total = o;
    return L;
}

// One request of a merge: its image in the arena and where its bytes,
// reads, haps and groups start in the merged batch.
struct GatherReq {
    const uint8_t* src;
    RequestLayout lay;
    int64_t rbytes, hbytes;  // the request's read and hap bytes
    int64_t rb_dst, hb_dst;  // their offsets in the merged arrays
    int32_t nR, nH, G;
    int32_t r_dst, h_dst, g_dst;
    int32_t pad;
};

// The merged batch (device pointers) and its totals.
struct GatherDst {
    uint8_t *read_data, *bq, *iq, *dq, *gcp, *hap_data;
    int32_t *read_len, *hap_len, *read_off, *hap_off, *nreads, *nhaps, *rprefix, *hprefix;
    int64_t read_bytes, hap_bytes;
    int32_t nR, nH, G;
};

// Merges the nreq requests d_reqs[] (device memory) into dst on stream st.
hipError_t launch_gather(const GatherReq* d_reqs, int nreq, const GatherDst& dst, hipStream_t st);

}  // namespace phmm
