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

// Several counter arrays zeroed by one launch (a fill kernel rather than
// memsets, so the sequence can be captured and replayed as a graph).
struct ClearList {
    int32_t* p[7];
    int32_t n[7];
    int32_t count;
};
__global__ void k_clear_list(ClearList cl)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int k = 0; k < cl.count; ++k)
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cl.n[k]; i += stride)
            cl.p[k][i] = 0;
}

// Grid-stride loops throughout: a launch covers any batch size, so the
// same grids serve every call when the sequence is captured in a graph.
#define GRID_STRIDE(i, n)                                                              \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (int64_t)(n); \
         i += (int64_t)gridDim.x * blockDim.x)

// Sequence bytes are padded to multiples of 4 per read/hap (the reference's
// padding rule) and start 4-byte aligned, so both passes below load and
// encode 4 bytes per lane: dword loads of each input array.

// One wave per hap: its code bytes (dwords over the padded length), whether
// it holds a byte other than A/C/G/T within its length (the wildcard --
// reference GPU semantics: any such byte; CPU semantics: 'N' -- or, under
// CPU semantics, a byte that matches nothing but itself; alignments with
// such haps go to the exact bin, so the tile kernels' emission tables need
// only the four bases), and the longest hap (per-block maxima into
// kHmaxParts counters, folded by k_scan_bins).
__global__ __launch_bounds__(256) void k_hap_pass(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    __shared__ int wmax[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nwaves = gridDim.x * (blockDim.x >> 6);
    const uint32_t* hd = reinterpret_cast<const uint32_t*>(b.hap_data);
    uint32_t* out = reinterpret_cast<uint32_t*>(bp.hapc);
    int m = 0;
    const int h0 = blockIdx.x * (blockDim.x >> 6) + wv;
    // the next hap's length and offset are loaded one hap ahead
    int hl_next = h0 < b.nH ? b.hap_len[h0] : 0;
    int off_next = h0 < b.nH ? b.hap_off[h0] : 0;
    for (int h = h0; h < b.nH; h += nwaves) {
        const int hl = hl_next, d0 = off_next >> 2;
        if (h + nwaves < b.nH) {
            hl_next = b.hap_len[h + nwaves];
            off_next = b.hap_off[h + nwaves];
        }
        m = max(m, hl);
        const int nd = (max(hl, 0) >> 2) + 1;  // padded length: hl + 4 - hl % 4
        bool other = false;
        for (int d = lane; d < nd; d += 64) {
            const uint32_t w = hd[d0 + d];
            uint32_t c = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t code = (uint32_t)base_code_sem((uint8_t)(w >> (8 * k)), bp.semantics);
                c |= code << (8 * k);
                other = other | ((4 * d + k < hl) & (code >= 4u));
            }
            out[d0 + d] = c;
        }
        other = __any(other);
        if (lane == 0)
            bp.hapflag[h] = other ? 1 : 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        m = max(m, __shfl_xor(m, o, 64));
    if (lane == 0)
        wmax[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int bm = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k)
            bm = max(bm, wmax[k]);
        if (bm > 0)  // the parts start at 0: blocks without haps skip the atomic
            atomicMax(&bp.hmax_part[blockIdx.x & (kHmaxParts - 1)], bm);
    }
}

// Per read (16 lanes each, four reads per wave): its packed rows and its
// bin.  Lane sl of a read's slot takes rows 4d .. 4d+3 for d = sl, sl+16,
// ... over the read's padded length: one dword of each input array, one
// 16-byte store of the packed words (base code | base q << 8 | ins q << 16 |
// del q << 24, qualities clamped to 127).  Tile class k is the first table
// row whose max_read_len covers the read (its tile has room for one more
// row); reads beyond the last row go to the multi-tile bin 2*ntiles.  Bin 2k
// uses the rescaled (FAST) arithmetic when the read's coefficients keep it in
// range, else bin 2k+1 (SAFE); under per-position gcp every row's eps
// enters the check.  Reference CPU and per-position gcp semantics: a read
// holding a byte other than A/C/G/T/N goes to the exact bin 2*ntiles+1.  The slot
// combines its lanes' verdicts and the product of alpha with row-local
// reductions.
__global__ __launch_bounds__(256) void k_read_meta(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    __shared__ float ph2pr[128];
    for (int q = threadIdx.x; q < 128; q += blockDim.x)
        ph2pr[q] = bp.ph2pr[q];
    __syncthreads();
    const uint32_t* rd = reinterpret_cast<const uint32_t*>(b.read_data);
    const uint32_t* q0 = reinterpret_cast<const uint32_t*>(b.bq);
    const uint32_t* q1 = reinterpret_cast<const uint32_t*>(b.iq);
    const uint32_t* q2 = reinterpret_cast<const uint32_t*>(b.dq);
    const uint32_t* q3 = reinterpret_cast<const uint32_t*>(b.gcp);
    const bool row_gcp = bp.semantics == 2;  // per-position gcp: eps of every row
    uint4* rows4 = reinterpret_cast<uint4*>(bp.rows);
    const int sl = threadIdx.x & 15;
    const int nslots = gridDim.x * (blockDim.x >> 4);
    // every 16-lane slot runs the same number of iterations, so the row-local
    // reductions below never have a slot missing
    const int iters = (b.nR + nslots - 1) / nslots;
    const int i0 = blockIdx.x * (blockDim.x >> 4) + (threadIdx.x >> 4);
    // the next read's length and offset are loaded one read ahead
    int L_next = i0 < b.nR ? b.read_len[i0] : 0;
    int off_next = i0 < b.nR ? b.read_off[i0] : 0;
    for (int it = 0; it < iters; ++it) {
        const int i = i0 + it * nslots;
        const bool valid = i < b.nR;
        const int L = valid ? L_next : 0;
        const int off_i = off_next;
        if (i + nslots < b.nR) {
            L_next = b.read_len[i + nslots];
            off_next = b.read_off[i + nslots];
        }
        int k = bp.ntiles;
        for (int j = 0; j < bp.ntiles; ++j)
            if (L <= bp.tiles[j]) {
                k = j;
                break;
            }
        const bool tiled = valid && L > 0 && k < bp.ntiles;
        const int off = (valid && L > 0) ? off_i : 0;
        const int nd = (valid && L > 0) ? (L >> 2) + 1 : 0;  // dwords of the padded read
        float eps = kEpsGpu;
        if (tiled && bp.semantics == 1)
            eps = ph2pr[min((int)b.gcp[off], 127)];
        else if (row_gcp)
            eps = 0.f;  // the rows' own eps are checked below
        bool ok = true, exotic = false;
        double nl = 1.0;
        float pdelta = 0.f, palpha = 1.f;  // the row above this lane's first row
        // same trip count on the slot's lanes (nd is the read's)
        for (int base = 0; base < nd; base += 16) {
            const int d = base + sl;
            uint32_t w[4] = {0u, 0u, 0u, 0u};
            uint32_t gw = 0u;  // gcp bytes of the four rows (per-position gcp only)
            if (d < nd) {
                const int q = (off >> 2) + d;
                const uint32_t r = rd[q], a = q0[q], c = q1[q], e = q2[q];
                if (row_gcp)
                    gw = q3[q];
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    const uint32_t code = (uint32_t)base_code_sem((uint8_t)(r >> (8 * x)), bp.semantics);
                    const uint32_t qb = min((a >> (8 * x)) & 0xffu, 127u);
                    const uint32_t qi = min((c >> (8 * x)) & 0xffu, 127u);
                    const uint32_t qd = min((e >> (8 * x)) & 0xffu, 127u);
                    w[x] = code | (qb << 8) | (qi << 16) | (qd << 24);
                }
                rows4[q] = make_uint4(w[0], w[1], w[2], w[3]);
            }
            if (!tiled)  // uniform over the slot
                continue;
            float dl[4], al[4], sg[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                dl[x] = ph2pr[row_iq(w[x])];
                sg[x] = ph2pr[row_dq(w[x])];
                al[x] = 1.0f - (dl[x] + sg[x]);
            }
            // row 4d-1: the next lower lane's last row; lane 0 keeps the
            // previous step's lane 15 row
            const float up_delta = __shfl_up(dl[3], 1, 16);
            const float up_alpha = __shfl_up(al[3], 1, 16);
            if (sl > 0) {
                pdelta = up_delta;
                palpha = up_alpha;
            }
            // branch-free per row (bitwise logic and selects: no divergent
            // control flow inside the slot)
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const int r = 4 * d + x;
                const bool in = r < L;
                // the row's gap continuation (per-position gcp: its own, at most 1/2)
                const float er = row_gcp ? ph2pr[min((gw >> (8 * x)) & 0xffu, 127u)] : eps;
                const bool range_bad = !((al[x] > 0.f) & (al[x] <= 1.f) & (dl[x] > 0.f) &
                                         (sg[x] > 0.f) & (er <= 0.5f));
                // e_r = eps_r*delta_{r-1} / (delta_r*alpha_{r-1}) > 1/2
                const bool er_bad = (r > 0) & !(er * pdelta <= 0.5f * (dl[x] * palpha));
                ok = ok & !(in & (range_bad | er_bad));
                exotic = exotic | (in & (row_code(w[x]) == kExoticCode));
                nl *= in ? (double)al[x] : 1.0;
                pdelta = dl[x];
                palpha = al[x];
            }
            const float last_delta = __shfl(dl[3], 15, 16);
            const float last_alpha = __shfl(al[3], 15, 16);
            if (sl == 0) {
                pdelta = last_delta;
                palpha = last_alpha;
            }
        }
        // reductions over the 16 lanes of this read (xor offsets < 16 stay
        // inside the 16-lane slot)
        int okv = ok ? 1 : 0, exv = exotic ? 1 : 0;
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) {
            nl *= __shfl_xor(nl, o, 64);
            okv &= __shfl_xor(okv, o, 64);
            exv |= __shfl_xor(exv, o, 64);
        }
        if (!valid || sl != 0)
            continue;
        ReadMeta m;
        m.nl = 1.0f;
        if (L <= 0) {
            m.bin = -1;
        } else if (k == bp.ntiles) {
            m.bin = 2 * bp.ntiles;
        } else if (bp.semantics != 0 && exv) {
            m.bin = 2 * bp.ntiles + 1;
        } else {
            bool fast = okv && bp.allow_fast != 0 && eps <= 0.5f;
            if (fast && !(nl >= 1e-30))  // the start constant must stay a normal float
                fast = false;
            m.nl = fast ? (float)nl : 1.0f;
            m.bin = 2 * k + (fast ? 0 : 1);
        }
        bp.meta[i] = m;
    }
}

// Exclusive prefix of R_g*H_g (the reference's resultOffsetsPerBatch,
// gpuworker.cuh:802-815), one block of 1024 threads, four consecutive
// groups per thread and pass.
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
    for (int base = 0; base < G; base += 4096) {
        const int g0 = base + 4 * (int)threadIdx.x;
        int64_t p[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            p[u] = g0 + u < G ? (int64_t)b.nreads[g0 + u] * b.nhaps[g0 + u] : 0;
        const int64_t v = p[0] + p[1] + p[2] + p[3];
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
        int64_t run = incl - v;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (g0 + u < G)
                bp.align_off[g0 + u] = run;
            run += p[u];
        }
        __syncthreads();
        if (threadIdx.x == 1023)
            carry = incl;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        bp.align_off[G] = carry;
}

// Sort key of a unit of group g: bin, then descending hap length (or
// segment columns), then g's part -- the part only spreads the atomics of
// many groups on one length over kKeyParts counters.
__device__ __forceinline__ int key_of(int bin, int hl, int g)
{
    return bin * kKeysPerBin + (kHapBuckets - 1 - hap_bucket(hl)) * kKeyParts + (g & (kKeyParts - 1));
}

// A group's non-empty read bins (including the empty-read slot NB) in LDS:
// bin id, count and start in the group's read_perm.  Every wave of the
// block runs the build (a block of several waves writes the same values).
struct GroupBins {
    int n;
    int k[kMaxBins + 1], cnt[kMaxBins + 1], st[kMaxBins + 1];
};
__device__ __forceinline__ void group_bins_compact(const BinParams& bp, int g, GroupBins& gb)
{
    const int lane = threadIdx.x & 63;
    const int NB1 = bp.NB + 1;
    if (lane == 0)
        gb.n = 0;
    __syncthreads();
    for (int k0 = 0; k0 < NB1; k0 += 64) {  // same trip count on every lane
        const int k = k0 + lane;
        const int c = k < NB1 ? bp.gb_count[(int64_t)g * NB1 + k] : 0;
        const unsigned long long m = __ballot(c > 0);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        const int base = gb.n;
        if (c > 0) {
            gb.k[base + before] = k;
            gb.cnt[base + before] = c;
            gb.st[base + before] = bp.gb_start[(int64_t)g * NB1 + k];
        }
        __syncthreads();
        if (lane == 0)
            gb.n = base + __popcll(m);
        __syncthreads();
    }
}

// The block of a group: count the group's reads per bin (slot NB: empty
// reads), store the counts, starts and the bin-ordered read permutation for
// the scatter, and keep the non-empty bins in LDS (as group_bins_compact
// would read them back).
__device__ __forceinline__ void group_bins_build(const BinParams& bp, const BatchDesc& b, int g,
                                                 GroupBins& gb, int* cnt, int* pos)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const int NB1 = bp.NB + 1;
    for (int k = tid; k < NB1; k += nt)
        cnt[k] = 0;
    __syncthreads();
    const int r0 = b.rprefix[g], nr = b.nreads[g];
    for (int r = tid; r < nr; r += nt) {
        const int k = bp.meta[r0 + r].bin;
        atomicAdd(&cnt[k < 0 ? bp.NB : k], 1);
    }
    __syncthreads();
    if (tid == 0) {
        int s = 0, n = 0;
        for (int k = 0; k < NB1; ++k) {
            const int c = cnt[k];
            pos[k] = s;
            bp.gb_start[(int64_t)g * NB1 + k] = s;
            bp.gb_count[(int64_t)g * NB1 + k] = c;
            if (c > 0) {
                gb.k[n] = k;
                gb.cnt[n] = c;
                gb.st[n] = s;
                ++n;
            }
            s += c;
        }
        gb.n = n;
    }
    __syncthreads();
    for (int r = tid; r < nr; r += nt) {
        const int k = bp.meta[r0 + r].bin;
        const int at = atomicAdd(&pos[k < 0 ? bp.NB : k], 1);
        bp.read_perm[r0 + at] = r0 + r;
    }
    __syncthreads();
}

// Per-group hap summary, uniform across the block (every wave computes it
// and writes the same values): haps with a sequence and
// those of them that go to the exact bin (a byte other than A/C/G/T); the others are
// cut, in order, into segments of at most `budget` columns (at least one hap
// each) -- the super-tasks' hap ranges and their work (sort key).
constexpr int kMaxSegs = 128;
struct GroupHaps {
    int nonempty, other, nseg;
    int hs[kMaxSegs], he[kMaxSegs], work[kMaxSegs];
    int cols[64];         // staging: columns of 64 haps (0: not cut into segments)
    int info[kMaxSegs];   // haps < kMaxSegs: length (0 if empty), top bit: flagged
};
// Length (<= 0: empty) and flag of record hap h, from the LDS copy where
// group_haps made one.
__device__ __forceinline__ void hap_info(const BinParams& bp, const BatchDesc& b,
                                         const GroupHaps& gh, int h0, int h, int& hl, bool& flag)
{
    if (h < kMaxSegs) {
        const int v = gh.info[h];
        hl = v & 0x7fffffff;
        flag = v < 0;
    } else {
        hl = b.hap_len[h0 + h];
        flag = bp.hapflag[h0 + h] != 0;
    }
}
__device__ __forceinline__ void group_haps(const BinParams& bp, const BatchDesc& b, int h0, int nh,
                                           GroupHaps& gh)
{
    int ne = 0, ex = 0, tot = 0;
    for (int h = (int)(threadIdx.x & 63); h < nh; h += 64) {
        const int hl = b.hap_len[h0 + h];
        if (hl > 0) {
            ++ne;
            if (bp.hapflag[h0 + h] != 0)
                ++ex;
            else
                tot += hl;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        ne += __shfl_xor(ne, o, 64);
        ex += __shfl_xor(ex, o, 64);
        tot += __shfl_xor(tot, o, 64);
    }
    // greedy segments number at most 2*tot/budget + 1: a larger budget where
    // they would not fit the table
    const int budget = max(b.col_budget, (2 * tot + kMaxSegs - 2) / (kMaxSegs - 1));
    // lane 0 cuts greedily over haps staged 64 at a time in LDS (one
    // parallel load per chunk instead of a serial chain of global loads)
    int n = 0, cur = 0, start = 0;
    for (int c0 = 0; c0 < nh; c0 += 64) {
        const int h = c0 + (int)(threadIdx.x & 63);
        int cols = 0;
        if (h < nh) {
            const int hl = b.hap_len[h0 + h];
            const bool flag = hl > 0 && bp.hapflag[h0 + h] != 0;
            cols = (hl > 0 && !flag) ? hl : 0;
            if (h < kMaxSegs)
                gh.info[h] = (hl > 0 ? hl : 0) | (flag ? (int)0x80000000u : 0);
        }
        gh.cols[threadIdx.x & 63] = cols;
        __syncthreads();
        if ((threadIdx.x & 63) == 0) {
            for (int i = 0; i < min(64, nh - c0); ++i) {
                const int hl = gh.cols[i];
                if (hl == 0)
                    continue;
                if (cur > 0 && cur + hl > budget) {
                    gh.hs[n] = start;
                    gh.he[n] = c0 + i;
                    gh.work[n] = cur;
                    ++n;
                    cur = 0;
                }
                if (cur == 0)
                    start = c0 + i;
                cur += hl;
            }
        }
        __syncthreads();
    }
    if ((threadIdx.x & 63) == 0) {
        gh.nonempty = ne;
        gh.other = ex;
        if (cur > 0) {
            gh.hs[n] = start;
            gh.he[n] = nh;
            gh.work[n] = cur;
            ++n;
        }
        gh.nseg = n;
    }
    __syncthreads();
}

__device__ __forceinline__ bool is_tile_bin(const BinParams& bp, int k) { return k < 2 * bp.ntiles; }
// A tile bin's `reads` reads of a group split into chunks of W = 64/lanes
// reads: one super-task per full chunk and hap segment.  The n = reads % W
// reads left over become one-alignment items of the bin's item bin (packed
// W to a task across records, so no group of a wave idles for them).
struct ChunkPlan {
    int W, nfull, rem;
};
__device__ __forceinline__ ChunkPlan chunk_plan(const BinParams& bp, int k, int reads)
{
    ChunkPlan c;
    c.W = 64 / bp.lanes[k >> 1];
    c.nfull = reads / c.W;
    c.rem = reads - c.nfull * c.W;
    return c;
}
// The item bin of tile bin k (work-bin ids past the multi-tile and exact bins).
__device__ __forceinline__ int item_bin_of(const BinParams& bp, int k) { return bp.NB + k; }

// Records with many (hap, bin) runs -- haps x min(reads, 64) above
// kBigGroupRuns -- are counted and scattered by a block of four waves
// (k_count_keys_t<4>, k_scatter_t<4>), the others by one wave each: one wave
// handling a large record's runs in turn is latency-bound.
constexpr int kBigGroupRuns = 512;
__device__ __forceinline__ bool big_group(const BatchDesc& b, int g)
{
    return (int64_t)b.nhaps[g] * min(b.nreads[g], 64) > kBigGroupRuns;
}

// Orders LDS traffic between the lanes of one wave (the waves of a block
// run different numbers of hap chunks, so no block barrier applies there).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per group (NW = 1: one wave, the small groups; NW = 4: a block of four
// waves, the large ones): its reads per bin (group_bins_build), the per-bin
// alignment counts and the sort-key histogram.  A tile bin's full chunks get
// one key (bin, segment columns) per hap segment, counted in super-tasks;
// item bins, the multi-tile and the exact bins get one key (bin, hap length)
// per (hap, bin) run, counted in alignments.
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_count_keys_t(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    __shared__ GroupBins gb;
    __shared__ GroupHaps gh;
    __shared__ int cnt[kMaxBins + 1], pos[kMaxBins + 1];
    const int eb = 2 * bp.ntiles + 1;
    for (int g = blockIdx.x; g < b.G; g += gridDim.x) {
        if (big_group(b, g) != (NW > 1))  // uniform over the block
            continue;
        group_bins_build(bp, b, g, gb, cnt, pos);
        const int h0 = b.hprefix[g], nh = b.nhaps[g];
        group_haps(bp, b, h0, nh, gh);
        const int n = gb.n;
        // per bin entry: alignment counts, spread over kAlignParts counters
        // per bin (every group adds to its bins: one counter per bin would
        // serialize the atomics of all groups)
        const int part = g & (kAlignParts - 1);
        for (int e = threadIdx.x; e < n; e += 64 * NW) {
            const int k = gb.k[e], c = gb.cnt[e];
            if (k == bp.NB)
                continue;
            if (is_tile_bin(bp, k)) {
                const int tiled = gh.nonempty - gh.other;
                if (tiled > 0)
                    atomicAdd(&bp.align_part[k * kAlignParts + part], c * tiled);
                if (gh.other > 0)
                    atomicAdd(&bp.align_part[eb * kAlignParts + part], c * gh.other);
            } else if (gh.nonempty > 0) {
                atomicAdd(&bp.align_part[k * kAlignParts + part], c * gh.nonempty);
            }
        }
        // per (tile bin entry, hap segment): its super-tasks' key
        for (int x = threadIdx.x; x < n * gh.nseg; x += 64 * NW) {
            const int e = x / gh.nseg, sg = x - e * gh.nseg;
            const int k = gb.k[e];
            if (k != bp.NB && is_tile_bin(bp, k)) {
                const ChunkPlan c = chunk_plan(bp, k, gb.cnt[e]);
                if (c.nfull > 0)
                    atomicAdd(&bp.key_count[key_of(k, gh.work[sg], g)], c.nfull);
            }
        }
        // per (hap, bin) item run: the one-wave bins' alignments, and a tile
        // bin's left-over reads against each of its haps
        for (int x = threadIdx.x; x < nh * n; x += 64 * NW) {
            const int h = x / n, e = x - h * n;
            const int k = gb.k[e];
            int hl;
            bool other;
            hap_info(bp, b, gh, h0, h, hl, other);
            if (hl <= 0 || k == bp.NB)
                continue;
            if (is_tile_bin(bp, k) && !other) {
                const ChunkPlan c = chunk_plan(bp, k, gb.cnt[e]);
                if (c.rem > 0)
                    atomicAdd(&bp.key_count[key_of(item_bin_of(bp, k), hl, g)], c.rem);
                continue;
            }
            atomicAdd(&bp.key_count[key_of(is_tile_bin(bp, k) ? eb : k, hl, g)], gb.cnt[e]);
        }
        __syncthreads();
    }
}

// Exclusive scan of each bin's key histogram (one block per bin): a thread
// per hap-length bucket adds the bucket's kKeyParts counters, the block scans
// the buckets, and each counter's cursor is its bucket's start plus the
// counters before it in the bucket.  key_cursor holds positions relative to
// the bin's begin, the bin's count goes to bin_range[2*bin+1].
__global__ __launch_bounds__(1024) void k_scan_keys(BinParams bp)
{
    static_assert(kKeyParts % 4 == 0, "buckets load as int4");
    static_assert(kAlignParts <= 1024, "one part per thread");
    __shared__ int wsum[16];
    __shared__ int carry;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int bin = blockIdx.x; bin < bp.NW; bin += gridDim.x) {
        if (bin < bp.NB) {  // fold the bin's alignment-count parts into part 0
            int v = (int)threadIdx.x < kAlignParts ? bp.align_part[bin * kAlignParts + threadIdx.x] : 0;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1)
                v += __shfl_xor(v, o, 64);
            if (lane == 0)
                wsum[wv] = v;
            __syncthreads();
            if (threadIdx.x == 0) {
                int t = 0;
                for (int k = 0; k < 16; ++k)
                    t += wsum[k];
                bp.align_part[bin * kAlignParts] = t;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0)
            carry = 0;
        __syncthreads();
        const int base0 = bin * kKeysPerBin;
        for (int base = 0; base < kHapBuckets; base += 1024) {
            const int bk = base + threadIdx.x;
            int c[kKeyParts];
            int v = 0;
#pragma unroll
            for (int q = 0; q < kKeyParts / 4; ++q) {
                int4 x = make_int4(0, 0, 0, 0);
                if (bk < kHapBuckets)
                    x = reinterpret_cast<const int4*>(bp.key_count + base0 + bk * kKeyParts)[q];
                c[4 * q] = x.x;
                c[4 * q + 1] = x.y;
                c[4 * q + 2] = x.z;
                c[4 * q + 3] = x.w;
                v += x.x + x.y + x.z + x.w;
            }
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
            if (bk < kHapBuckets) {
                int run = incl - v;
#pragma unroll
                for (int q = 0; q < kKeyParts / 4; ++q) {
                    int4 y;
                    y.x = run;
                    run += c[4 * q];
                    y.y = run;
                    run += c[4 * q + 1];
                    y.z = run;
                    run += c[4 * q + 2];
                    y.w = run;
                    run += c[4 * q + 3];
                    reinterpret_cast<int4*>(bp.key_cursor + base0 + bk * kKeyParts)[q] = y;
                }
            }
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

// Bin begins: exclusive scans of the bin counts -- in work units
// (bin_range, every work bin) and in alignments (align_range, the bins of
// reads) -- one wave.
// Counts come from range[2k+1], or from cnt[k] (then also stored there).
__device__ __forceinline__ void scan_pairs(int32_t* range, int NB, const int* cnt)
{
    int carry = 0;
    for (int base = 0; base < NB; base += 64) {
        const int k = base + (int)threadIdx.x;
        const int v = k < NB ? (cnt ? cnt[k] : range[2 * k + 1]) : 0;
        int x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if ((int)threadIdx.x >= o)
                x += y;
        }
        if (k < NB) {
            range[2 * k] = carry + x - v;
            if (cnt)
                range[2 * k + 1] = v;
        }
        carry += __shfl(x, 63, 64);
    }
}

__global__ __launch_bounds__(64) void k_scan_bins(BinParams bp)
{
    __shared__ int tot[kMaxBins];
    {  // the longest hap, from the hap pass's per-block maxima
        int m = 0;
#pragma unroll
        for (int q = threadIdx.x; q < kHmaxParts; q += 64)
            m = max(m, bp.hmax_part[q]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
            m = max(m, __shfl_xor(m, o, 64));
        if (threadIdx.x == 0)
            *bp.hmax = m;
    }
    scan_pairs(bp.bin_range, bp.NW, nullptr);
    // each bin's alignment count (k_scan_keys folded the bin's parts into
    // part 0)
    for (int k = threadIdx.x; k < bp.NB; k += 64)
        tot[k] = bp.align_part[k * kAlignParts];
    __syncthreads();
    scan_pairs(bp.align_range, bp.NB, tot);
}

// One group's item runs of haps [hc, hc + hcn) (at most kRuns (hap, bin
// entry) pairs), written by one wave: lanes reserve the runs -- tile-bin
// pairs are covered by super-tasks -- and list the others compactly; then
// the wave writes each listed run, lanes over the run's reads.  Alignments
// with an empty read or hap get -inf directly (align_host returns
// log10(0) - log10(c) for them).
template <int kRuns, bool kWaveSync>
__device__ __forceinline__ void scatter_hap_chunk(const BinParams& bp, const BatchDesc& b,
                                                  const GroupBins& gb, const GroupHaps& gh, int g,
                                                  int r0, int h0, int nh, int64_t A, int hc,
                                                  int hcn, int* run_x, int* run_base, int* w_off,
                                                  int* w_st, int* w_h, int* w_base)
{
    constexpr int kSkip = -2;  // run covered by a super-task
    const int lane = threadIdx.x & 63;
    const int eb = 2 * bp.ntiles + 1;
    const int n = gb.n;
    auto sync = [] {
        if (kWaveSync)
            wave_sync();
        else
            __syncthreads();
    };
    int nrun = 0;  // uniform
    for (int x0 = 0; x0 < hcn * n; x0 += 64) {  // same trip count on every lane
        const int x = x0 + lane;
        int base = kSkip;
        if (x < hcn * n) {
            const int h = hc + x / n, e = x % n;
            const int k = gb.k[e];
            int hl;
            bool other;
            hap_info(bp, b, gh, h0, h, hl, other);
            base = -1;  // -inf
            if (hl > 0 && k != bp.NB) {
                if (is_tile_bin(bp, k) && !other) {
                    // super-tasks cover the full chunks; left-over reads
                    // become items of the bin's item bin
                    const ChunkPlan cp = chunk_plan(bp, k, gb.cnt[e]);
                    const int kb = item_bin_of(bp, k);
                    base = cp.rem == 0 ? kSkip
                                       : bp.bin_range[2 * kb] +
                                             atomicAdd(&bp.key_cursor[key_of(kb, hl, g)], cp.rem);
                } else {
                    const int kb = is_tile_bin(bp, k) ? eb : k;
                    base = bp.bin_range[2 * kb] + atomicAdd(&bp.key_cursor[key_of(kb, hl, g)], gb.cnt[e]);
                }
            }
        }
        const bool keep = base != kSkip;
        const unsigned long long m = __ballot(keep);
        if (keep) {
            const int pos = nrun + __popcll(m & ((1ull << lane) - 1ull));
            run_x[pos] = x;
            run_base[pos] = base;
        }
        nrun += __popcll(m);
    }
    sync();
    // write the listed runs 64 at a time, lanes over the runs' (run, read)
    // pairs: many runs are a read or two (left-over reads), and every pair
    // costs a dependent load of read_perm
    for (int q0 = 0; q0 < nrun; q0 += 64) {  // same trip count on every lane
        const int q = q0 + lane;
        int cnt = 0, st = 0, h = 0, base = 0;
        if (q < nrun) {
            const int x = run_x[q];
            const int e = x % n;
            h = hc + x / n;
            base = run_base[q];
            cnt = gb.cnt[e];
            st = gb.st[e];
            if (base >= 0 && is_tile_bin(bp, gb.k[e])) {
                int hl;
                bool other;
                hap_info(bp, b, gh, h0, h, hl, other);
                if (!other) {  // an item-bin run: the left-over reads only
                    const ChunkPlan cp = chunk_plan(bp, gb.k[e], cnt);
                    st += cp.nfull * cp.W;
                    cnt = cp.rem;
                }
            }
        }
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o)
                incl += y;
        }
        const int total = __shfl(incl, 63, 64);
        w_off[lane] = incl - cnt;
        w_st[lane] = st;
        w_h[lane] = h;
        w_base[lane] = base;
        sync();
        const int last = min(64, nrun - q0) - 1;
        for (int f = lane; f < total; f += 64) {
            int lo = 0, hi = last;  // the last run whose offset is <= f
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (w_off[mid] <= f)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int j = f - w_off[lo];
            const int rd = bp.read_perm[r0 + w_st[lo] + j];
            const int64_t out = A + (int64_t)(rd - r0) * nh + w_h[lo];
            if (w_base[lo] < 0) {
                b.scores[out] = -INFINITY;  // empty read or empty hap
            } else {
                WorkItem it;
                it.read = rd;
                it.hap = h0 + w_h[lo];
                it.out = (int32_t)out;
                it.pad = 0;
                bp.items[w_base[lo] + j] = it;
            }
        }
        sync();
    }
    sync();
}

// Write the work lists, per group.  Tile bins: the group's reads of the bin
// in full chunks of 64/lanes, one super-task per chunk and hap segment; the
// reads left over go to the bin's item bin.  Item bins, the multi-tile and
// the exact bins: one item per alignment, reserved per (hap, bin) run (haps
// in chunks that fit the LDS table).  NW = 1: one wave per group, the
// small groups (big_group).  NW = 4: a block of four waves per group, the
// others; its waves share the super-task pairs and take the hap
// chunks in turn.  Both launches stride over all groups and skip the other
// kind.
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_scatter_t(BinParams bp)
{
    const BatchDesc& b = *bp.bd;
    constexpr int kRuns = 256;
    __shared__ GroupBins gb;
    __shared__ GroupHaps gh;
    __shared__ int run_x[NW][kRuns], run_base[NW][kRuns];
    __shared__ int w_off[NW][64], w_st[NW][64], w_h[NW][64], w_base[NW][64];
    const int wv = threadIdx.x >> 6;
    for (int g = blockIdx.x; g < b.G; g += gridDim.x) {
        if (big_group(b, g) != (NW > 1))  // uniform over the block
            continue;
        group_bins_compact(bp, g, gb);
        const int r0 = b.rprefix[g];
        const int h0 = b.hprefix[g], nh = b.nhaps[g];
        const int64_t A = bp.align_off[g];
        group_haps(bp, b, h0, nh, gh);
        const int n = gb.n;
        // super-tasks of the tile bins: read chunks x hap segments, one
        // thread per (bin entry, segment) pair: it reserves the pair's list
        // slots and writes its tasks
        const int npairs = n * gh.nseg;
        for (int x = threadIdx.x; x < npairs; x += 64 * NW) {
            const int e = x / gh.nseg, sg = x - e * gh.nseg;
            const int k = gb.k[e];
            if (k == bp.NB || !is_tile_bin(bp, k))
                continue;
            const ChunkPlan cp = chunk_plan(bp, k, gb.cnt[e]);
            SuperTask t;
            t.h0 = h0;
            t.nh = nh;
            t.hs = gh.hs[sg];
            t.r0 = r0;
            t.aoff = (int32_t)A;
            if (cp.nfull > 0) {
                const int base = bp.bin_range[2 * k] +
                                 atomicAdd(&bp.key_cursor[key_of(k, gh.work[sg], g)], cp.nfull);
                t.nc = cp.W;
                t.he = gh.he[sg];
                for (int j = 0; j < cp.nfull; ++j) {
                    t.perm = r0 + gb.st[e] + j * cp.W;
                    bp.st[base + j] = t;
                }
            }
        }
        const int hchunk = n > 0 ? max(1, kRuns / n) : nh;
        for (int hc = wv * hchunk; hc < nh; hc += NW * hchunk)
            scatter_hap_chunk<kRuns, (NW > 1)>(bp, b, gb, gh, g, r0, h0, nh, A, hc,
                                               min(hchunk, nh - hc), run_x[wv], run_base[wv],
                                               w_off[wv], w_st[wv], w_h[wv], w_base[wv]);
        __syncthreads();  // the next group rebuilds gb and gh
    }
}

__global__ void k_add_counts(CountSums c)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        for (int k = 0; k < c.n; ++k) {
            int64_t s = 0;
            for (int i = 0; i < c.len[k]; ++i)
                s += c.counts[k][i];
            // calls on several streams may add at once
            atomicAdd(reinterpret_cast<unsigned long long*>(c.total[k]), (unsigned long long)s);
        }
    }
}

hipError_t launch_add_counts(const CountSums& c, hipStream_t st)
{
    if (c.n < 1 || c.n > CountSums::kMax)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_add_counts, dim3(1), dim3(64), 0, st, c);
    return hipGetLastError();
}

hipError_t launch_add_count(int64_t* total, const int32_t* counts, int n, hipStream_t st)
{
    CountSums c;
    c.n = 1;
    c.total[0] = total;
    c.counts[0] = counts;
    c.len[0] = n;
    return launch_add_counts(c, st);
}

// ---------------------------------------------------------------------------
hipError_t run_binning(const BinParams& bp, const BatchDesc& hd, bool fixed_grids, bool small,
                       hipStream_t st)
{
    const int tpb = 256;
    // grids: sized to the batch, or fixed (graph capture) at a size that
    // still fills the chip -- a quarter of it for the small-batch sequence,
    // whose batches are too small to use more and whose latency pays for
    // every idle wave dispatched
    auto grid = [&](int64_t n, int per_block, int cap) {
        if (fixed_grids)
            return small ? cap / 4 : cap;
        int64_t b = (n + per_block - 1) / per_block;
        return (int)(b < cap ? (b < 1 ? 1 : b) : cap);
    };
    ClearList cl{};
    cl.p[0] = bp.key_count;
    cl.n[0] = bp.NW * kKeysPerBin;
    cl.p[1] = bp.hmax_part;
    cl.n[1] = kHmaxParts;
    cl.p[2] = bp.rescue_counts;
    cl.n[2] = bp.NB;
    cl.p[3] = bp.final_count;
    cl.n[3] = 1;
    cl.p[4] = bp.task_ctr;
    cl.n[4] = kTaskCounters;
    cl.p[5] = bp.align_part;
    cl.n[5] = kAlignParts * bp.NB;
    cl.p[6] = bp.retry_counts;
    cl.n[6] = kRetrySlots;
    cl.count = 7;
    hipLaunchKernelGGL(k_clear_list, dim3(fixed_grids && small ? 128 : 512), dim3(tpb), 0, st, cl);
    if (fixed_grids || hd.nH > 0)  // one wave per hap
        hipLaunchKernelGGL(k_hap_pass, dim3(grid(hd.nH, tpb / 64, 4096)), dim3(tpb), 0, st, bp);
    if (fixed_grids || hd.nR > 0)  // 16 lanes per read: rows and bins
        hipLaunchKernelGGL(k_read_meta, dim3(grid(hd.nR, tpb / 16, 4096)), dim3(tpb), 0, st, bp);
    hipLaunchKernelGGL(k_align_offsets, dim3(1), dim3(1024), 0, st, bp);
    if (fixed_grids || hd.G > 0) {
        hipLaunchKernelGGL(k_count_keys_t<1>, dim3(grid(hd.G, 1, 8192)), dim3(64), 0, st, bp);
        hipLaunchKernelGGL(k_count_keys_t<4>, dim3(grid(hd.G, 1, 2048)), dim3(256), 0, st, bp);
    }
    hipLaunchKernelGGL(k_scan_keys, dim3(bp.NW), dim3(1024), 0, st, bp);
    hipLaunchKernelGGL(k_scan_bins, dim3(1), dim3(64), 0, st, bp);
    if (fixed_grids || hd.G > 0) {
        hipLaunchKernelGGL(k_scatter_t<1>, dim3(grid(hd.G, 1, 8192)), dim3(64), 0, st, bp);
        hipLaunchKernelGGL(k_scatter_t<4>, dim3(grid(hd.G, 1, 2048)), dim3(256), 0, st, bp);
    }
    return hipGetLastError();
}

}  // namespace phmm
