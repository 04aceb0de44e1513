// Device-side preparation of a batch: base encoding, per-read bins and a
// counting sort of all alignments into a work list ordered by (bin,
// descending hap length).
//
// Replaces the reference's convert_DNA / partitionIndicesKernel /
// computeAlignmentsPerPartitionPerBatch plus its thrust/cub scans and the
// per-chunk host round trip (utility_kernels.cuh:6-102; gpuworker.cuh:747-888).
// The reference partitions reads by the smallest tile that holds them; here a
// bin is (tile class, arithmetic form), and alignments inside a bin are
// ordered by hap length so the groups of one wave sweep the same number of
// columns.  Within a record, every alignment of a hap shares the hap's
// length, so counts and positions are reserved per (hap, bin) and the
// atomics scale with haps x bins, not alignments.
//
// Views passed here are self-contained: reads_prefix[0] == 0,
// haps_prefix[0] == 0, and read/hap offsets index read_data/hap_data.
#include "phmm_device.hpp"

namespace phmm {



__global__ void k_clear(int32_t* p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Grid-stride loops throughout: a launch covers any batch size, so the
// same grids serve every call when the sequence is captured in a graph.
#define GRID_STRIDE(i, n)                                                              \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)(n); \
         i += (int64_t)gridDim.x * blockDim.x)

__global__ void k_encode_reads(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    GRID_STRIDE(i, b.read_bytes)
    {
        const uint32_t c = (uint32_t)base_code_sem(b.read_data[i], bp.semantics);
        const uint32_t q = min((uint32_t)b.bq[i], 127u);
        const uint32_t qi = min((uint32_t)b.iq[i], 127u);
        const uint32_t qd = min((uint32_t)b.dq[i], 127u);
        bp.rows[i] = c | (q << 8) | (qi << 16) | (qd << 24);
    }
}

__global__ void k_encode_haps(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    GRID_STRIDE(i, b.hap_bytes)
    bp.hapc[i] = (uint8_t)base_code_sem(b.hap_data[i], bp.semantics);
}

__global__ void k_hap_max(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    int m = 0;
    GRID_STRIDE(i, b.nH)
    m = max(m, b.hap_len[i]);
    for (int o = 32; o > 0; o >>= 1)
        m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0)
        atomicMax(bp.hmax, m);
}

// Per read (one wave each): its bin.  Tile class k is the first table row
// whose max_read_len covers the read (its tile has room for one more row);
// reads beyond the last row go to the multi-tile bin 2*ntiles.  Bin 2k uses
// the rescaled (FAST) arithmetic when the read's coefficients keep it in
// range, else bin 2k+1 (SAFE).  The lanes check rows lane, lane+64, ... and
// the wave combines their verdicts and the product of alpha.
__global__ __launch_bounds__(256) void k_read_meta(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    const int lane = threadIdx.x & 63;
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    for (int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < b.nR; i += nwaves) {
        const int L = b.read_len[i];
        ReadMeta m;
        m.nl = 1.0f;
        int k = bp.ntiles;
        for (int j = 0; j < bp.ntiles; ++j)
            if (L <= bp.tiles[j]) {
                k = j;
                break;
            }
        if (L <= 0 || k == bp.ntiles) {
            m.bin = L <= 0 ? -1 : 2 * bp.ntiles;
            if (lane == 0)
                bp.meta[i] = m;
            continue;
        }
        const int off = b.read_off[i];
        if (bp.semantics == 1) {
            bool exotic = false;
            for (int r = lane; r < L; r += 64)
                exotic = exotic || row_code(bp.rows[off + r]) == kExoticCode;
            if (__any(exotic)) {
                m.bin = 2 * bp.ntiles + 1;
                if (lane == 0)
                    bp.meta[i] = m;
                continue;
            }
        }
        float eps = kEpsGpu;
        if (bp.semantics == 1)
            eps = bp.ph2pr[min((int)b.gcp[off], 127)];
        bool ok = true;
        double nl = 1.0;
        for (int r = lane; r < L; r += 64) {
            const uint32_t w = bp.rows[off + r];
            const float delta = bp.ph2pr[row_iq(w)];
            const float sigma = bp.ph2pr[row_dq(w)];
            const float alpha = 1.0f - (delta + sigma);
            if (!(alpha > 0.f && alpha <= 1.f && delta > 0.f && sigma > 0.f)) {
                ok = false;
            } else if (r > 0) {
                const uint32_t wp = bp.rows[off + r - 1];
                const float pdelta = bp.ph2pr[row_iq(wp)];
                const float palpha = 1.0f - (pdelta + bp.ph2pr[row_dq(wp)]);
                if (!(eps * pdelta / (delta * palpha) <= 0.5f))
                    ok = false;
            }
            nl *= (double)alpha;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
            nl *= __shfl_xor(nl, o, 64);
        ok = __all(ok) && bp.allow_fast != 0 && eps <= 0.5f;
        if (ok && !(nl >= 1e-30))  // the start constant must stay a normal float
            ok = false;
        m.nl = ok ? (float)nl : 1.0f;
        m.bin = 2 * k + (ok ? 0 : 1);
        if (lane == 0)
            bp.meta[i] = m;
    }
}

// Exclusive prefix of R_g*H_g (the reference's resultOffsetsPerBatch,
// gpuworker.cuh:802-815), one block of 1024 threads.
__global__ __launch_bounds__(1024) void k_align_offsets(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    const int G = b.G;
    __shared__ int64_t wsum[16];
    __shared__ int64_t carry;
    if (threadIdx.x == 0)
        carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int base = 0; base < G; base += 1024) {
        const int g = base + threadIdx.x;
        int64_t v = g < G ? (int64_t)b.nreads[g] * b.nhaps[g] : 0;
        int64_t x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t y = __shfl_up(x, o, 64);
            if (lane >= o)
                x += y;
        }
        if (lane == 63)
            wsum[wv] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t s = 0;
            for (int k = 0; k < 16; ++k) {
                const int64_t t = wsum[k];
                wsum[k] = s;
                s += t;
            }
        }
        __syncthreads();
        const int64_t incl = x + wsum[wv] + carry;
        if (g < G)
            bp.align_off[g] = incl - v;
        __syncthreads();
        if (threadIdx.x == 1023)
            carry = incl;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        bp.align_off[G] = carry;
}

// Per group (grid-stride over groups, one block each): count its reads per
// bin, lay them out bin by bin.
__global__ __launch_bounds__(256) void k_group_bins(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    __shared__ int cnt[kMaxBins + 1];
    __shared__ int pos[kMaxBins + 1];
    const int NB1 = bp.NB + 1;  // slot NB: empty reads
    for (int g = blockIdx.x; g < b.G; g += gridDim.x) {
        for (int k = threadIdx.x; k < NB1; k += blockDim.x)
            cnt[k] = 0;
        __syncthreads();
        const int r0 = b.rprefix[g], nr = b.nreads[g];
        for (int r = threadIdx.x; r < nr; r += blockDim.x) {
            const int k = bp.meta[r0 + r].bin;
            atomicAdd(&cnt[k < 0 ? bp.NB : k], 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int s = 0;
            for (int k = 0; k < NB1; ++k) {
                pos[k] = s;
                bp.gb_start[(int64_t)g * NB1 + k] = s;
                bp.gb_count[(int64_t)g * NB1 + k] = cnt[k];
                s += cnt[k];
            }
        }
        __syncthreads();
        for (int r = threadIdx.x; r < nr; r += blockDim.x) {
            const int k = bp.meta[r0 + r].bin;
            const int at = atomicAdd(&pos[k < 0 ? bp.NB : k], 1);
            bp.read_perm[r0 + at] = r0 + r;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int key_of(int bin, int hl)
{
    return bin * kHapBuckets + (kHapBuckets - 1 - hap_bucket(hl));
}

// Reference CPU semantics: per hap (one wave each), whether it holds a byte
// other than A/C/G/T/N.
__global__ __launch_bounds__(256) void k_hap_flags(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    const int lane = threadIdx.x & 63;
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    for (int h = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); h < b.nH; h += nwaves) {
        const int hl = b.hap_len[h], off = b.hap_off[h];
        bool exotic = false;
        for (int c = lane; c < hl; c += 64)
            exotic = exotic || bp.hapc[off + c] == kExoticCode;
        exotic = __any(exotic);
        if (lane == 0)
            bp.hapflag[h] = exotic ? 1 : 0;
    }
}

// The bin an alignment of read bin k and hap h goes to: a hap with bytes
// other than A/C/G/T/N (CPU semantics) sends its tile-bin alignments to the
// exact bin.
__device__ __forceinline__ int bin_of(const BinParams& bp, int k, int h)
{
    return (bp.semantics == 1 && k < 2 * bp.ntiles && bp.hapflag[h]) ? 2 * bp.ntiles + 1 : k;
}

// Per group: add its alignment counts to the (bin, hap length) histogram.
__global__ __launch_bounds__(256) void k_count_keys(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    const int NB1 = bp.NB + 1;
    for (int g = blockIdx.x; g < b.G; g += gridDim.x) {
        const int h0 = b.hprefix[g], nh = b.nhaps[g];
        for (int h = threadIdx.x; h < nh; h += blockDim.x) {
            const int hl = b.hap_len[h0 + h];
            if (hl <= 0)
                continue;
            for (int k = 0; k < bp.NB; ++k) {
                const int n = bp.gb_count[(int64_t)g * NB1 + k];
                if (n > 0)
                    atomicAdd(&bp.key_count[key_of(bin_of(bp, k, h0 + h), hl)], n);
            }
        }
    }
}

// Exclusive scan of each bin's hap-length histogram (one block per bin):
// key_cursor holds positions relative to the bin's begin, the bin's count
// goes to bin_range[2*bin+1].
__global__ __launch_bounds__(1024) void k_scan_keys(BinParams bp)
{
    __shared__ int wsum[16];
    __shared__ int carry;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int bin = blockIdx.x; bin < bp.NB; bin += gridDim.x) {
        if (threadIdx.x == 0)
            carry = 0;
        __syncthreads();
        const int base0 = bin * kHapBuckets;
        for (int base = 0; base < kHapBuckets; base += 1024) {
            const int i = base + threadIdx.x;
            const int v = i < kHapBuckets ? bp.key_count[base0 + i] : 0;
            int x = v;
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(x, o, 64);
                if (lane >= o)
                    x += y;
            }
            if (lane == 63)
                wsum[wv] = x;
            __syncthreads();
            if (threadIdx.x == 0) {
                int s = 0;
                for (int k = 0; k < 16; ++k) {
                    const int t = wsum[k];
                    wsum[k] = s;
                    s += t;
                }
            }
            __syncthreads();
            const int incl = x + wsum[wv] + carry;
            if (i < kHapBuckets)
                bp.key_cursor[base0 + i] = incl - v;
            __syncthreads();
            if (threadIdx.x == 1023)
                carry = incl;
            __syncthreads();
        }
        if (threadIdx.x == 0)
            bp.bin_range[2 * bin + 1] = carry;
        __syncthreads();
    }
}

// Bin begins: exclusive scan of the bin counts (NB <= 65, one wave).
__global__ __launch_bounds__(64) void k_scan_bins(BinParams bp)
{
    int carry = 0;
    for (int base = 0; base < bp.NB; base += 64) {
        const int k = base + (int)threadIdx.x;
        const int v = k < bp.NB ? bp.bin_range[2 * k + 1] : 0;
        int x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if ((int)threadIdx.x >= o)
                x += y;
        }
        if (k < bp.NB)
            bp.bin_range[2 * k] = carry + x - v;
        carry += __shfl(x, 63, 64);
    }
}

// Per group: reserve each (hap, bin) run and write its items; alignments
// with an empty read or hap get -inf directly (align_host returns
// log10(0) - log10(c) for them).
__global__ __launch_bounds__(256) void k_scatter(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    const int NB1 = bp.NB + 1;
    for (int g = blockIdx.x; g < b.G; g += gridDim.x) {
        const int r0 = b.rprefix[g], nr = b.nreads[g];
        const int h0 = b.hprefix[g], nh = b.nhaps[g];
        const int64_t A = bp.align_off[g];
        for (int h = threadIdx.x; h < nh; h += blockDim.x) {
            const int hl = b.hap_len[h0 + h];
            if (hl <= 0) {
                for (int r = 0; r < nr; ++r)
                    b.scores[A + (int64_t)r * nh + h] = -INFINITY;
                continue;
            }
            for (int k = 0; k < NB1; ++k) {
                const int n = bp.gb_count[(int64_t)g * NB1 + k];
                if (n <= 0)
                    continue;
                const int st = bp.gb_start[(int64_t)g * NB1 + k];
                if (k == bp.NB) {
                    for (int j = 0; j < n; ++j) {
                        const int rd = bp.read_perm[r0 + st + j];
                        b.scores[A + (int64_t)(rd - r0) * nh + h] = -INFINITY;
                    }
                    continue;
                }
                const int kb = bin_of(bp, k, h0 + h);
                const int base = bp.bin_range[2 * kb] + atomicAdd(&bp.key_cursor[key_of(kb, hl)], n);
                for (int j = 0; j < n; ++j) {
                    const int rd = bp.read_perm[r0 + st + j];
                    WorkItem it;
                    it.read = rd;
                    it.hap = h0 + h;
                    it.out = (int32_t)(A + (int64_t)(rd - r0) * nh + h);
                    it.pad = 0;
                    bp.items[base + j] = it;
                }
            }
        }
    }
}

__global__ void k_add_count(int64_t* total, const int32_t* counts, int n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        int64_t s = 0;
        for (int i = 0; i < n; ++i)
            s += counts[i];
        *total += s;
    }
}

// *total += sum of counts[0..n)
hipError_t launch_add_count(int64_t* total, const int32_t* counts, int n, hipStream_t st)
{
    hipLaunchKernelGGL(k_add_count, dim3(1), dim3(64), 0, st, total, counts, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
hipError_t run_binning(const BinParams& bp, const BatchDesc& hd, bool fixed_grids, hipStream_t st)
{
    const int tpb = 256;
    // grids: sized to the batch, or fixed (graph capture) at a size that
    // still fills the chip
    auto grid = [&](int64_t n, int per_block, int cap) {
        if (fixed_grids)
            return cap;
        int64_t b = (n + per_block - 1) / per_block;
        return (int)(b < cap ? (b < 1 ? 1 : b) : cap);
    };
    hipLaunchKernelGGL(k_clear, dim3(grid((int64_t)bp.NB * kHapBuckets, tpb, 512)), dim3(tpb), 0,
                       st, bp.key_count, (int64_t)bp.NB * kHapBuckets);
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, st, bp.hmax, (int64_t)1);
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(128), 0, st, bp.rescue_counts, (int64_t)bp.NB);
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, st, bp.final_count, (int64_t)1);
    hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, st, bp.task_ctr, (int64_t)kTaskCounters);
    if (fixed_grids || hd.read_bytes > 0)
        hipLaunchKernelGGL(k_encode_reads, dim3(grid(hd.read_bytes, tpb, 4096)), dim3(tpb), 0, st, bp);
    if (fixed_grids || hd.hap_bytes > 0)
        hipLaunchKernelGGL(k_encode_haps, dim3(grid(hd.hap_bytes, tpb, 2048)), dim3(tpb), 0, st, bp);
    if (fixed_grids || hd.nH > 0)
        hipLaunchKernelGGL(k_hap_max, dim3(grid(hd.nH, tpb, 256)), dim3(tpb), 0, st, bp);
    if (fixed_grids || hd.nR > 0)  // one wave per read
        hipLaunchKernelGGL(k_read_meta, dim3(grid(hd.nR, tpb / 64, 4096)), dim3(tpb), 0, st, bp);
    if (bp.semantics == 1 && (fixed_grids || hd.nH > 0))  // one wave per hap
        hipLaunchKernelGGL(k_hap_flags, dim3(grid(hd.nH, tpb / 64, 4096)), dim3(tpb), 0, st, bp);
    hipLaunchKernelGGL(k_align_offsets, dim3(1), dim3(1024), 0, st, bp);
    if (fixed_grids || hd.G > 0) {
        hipLaunchKernelGGL(k_group_bins, dim3(grid(hd.G, 1, 8192)), dim3(256), 0, st, bp);
        hipLaunchKernelGGL(k_count_keys, dim3(grid(hd.G, 1, 8192)), dim3(256), 0, st, bp);
    }
    hipLaunchKernelGGL(k_scan_keys, dim3(bp.NB), dim3(1024), 0, st, bp);
    hipLaunchKernelGGL(k_scan_bins, dim3(1), dim3(64), 0, st, bp);
    if (fixed_grids || hd.G > 0)
        hipLaunchKernelGGL(k_scatter, dim3(grid(hd.G, 1, 8192)), dim3(256), 0, st, bp);
    return hipGetLastError();
}

}  // namespace phmm
