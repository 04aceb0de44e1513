// Forward-algorithm kernels for gfx950 (CDNA4, wave64).
//
// The recurrence is the reference's (align_host, gpuworker.cuh:1051-1094;
// GPU kernel pairhmm_kernels.cuh:1417-1763):
//   M[r][c] = lambda(r,c) * (alpha_r M[r-1][c-1] + beta (I[r-1][c-1] + D[r-1][c-1]))
//   I[r][c] = delta_r M[r-1][c] + eps I[r-1][c]
//   D[r][c] = sigma_r M[r][c-1] + eps D[r][c-1]
//   row 0: M = I = 0, D = (FLT_MAX/16)/H;  column 0 (r >= 1): 0
//   score = log10(sum_c M[R][c] + I[R][c]) - log10(FLT_MAX/16)
//
// k_fwd_tile -- the hot path.  A wave holds 64/G alignments ("groups") of G
// lanes; lane t of a group holds read rows [t*NR, t*NR+NR) in registers and
// the haplotype streams through the group one column per step (anti-diagonal
// wavefront).  Row-above values cross lanes with DPP wave_shr:1, a group's
// first lane injects the row-0 boundary.  Emission probabilities for the
// 5 hap letters x NR rows of each lane sit in LDS (one ds_read_b128 per
// 4 cells, conflict-free by lane interleaving).  Two arithmetic forms:
//
//  SAFE (6 VALU/cell): variables M, Ib = beta*I, Dh = D/sigma_r.
//    M  = lambda*(alpha M' + Ib' + beta sigma_{r-1} Dh')      fma fma mul
//    Ib = beta delta_r M^ + eps Ib^                            mul fma
//    Dh = M< + eps Dh<                                         fma
//    Magnitudes never exceed the reference's.
//
//  FAST (5 VALU/cell): row-rescaled variables with n_r = prod_{i<=r} alpha_i,
//    Mh = M/n_r, Ih = I/(delta_r n_{r-1}), Dh = D/(sigma_r n_r):
//    Mh = lambda*(Mh' + b_r Ih' + d_r Dh')                     fma fma mul
//    Ih = Mh^ + e_r Ih^                                        fma
//    Dh = Mh< + eps Dh<                                        fma
//    b_r = beta delta_{r-1}/(alpha_{r-1} alpha_r), d_r = beta sigma_{r-1}/alpha_r,
//    e_r = eps delta_{r-1}/(delta_r alpha_{r-1}).  The start constant is
//    lowered by n_R/2 (it cancels in the score) so that every rescaled value
//    stays below FLT_MAX/16; the read pass admits a read only when
//    0 < alpha <= 1, e_r <= 1/2, eps <= 1/2 and n_R >= 1e-3.
//
// Both forms use the reference's out-of-bounds-row trick for the result
// (pairhmm_kernels.cuh:1325-1356): the first row past the read turns the
// I-chain into (M+I) of the last row, later rows pass it through, and the
// lane holding row R accumulates it one column behind.
//
// k_fwd_wave -- one alignment per wave, any read length, FP32 or FP64, with
// the plain (unscaled) recurrence.  Reads longer than the largest tile are
// processed in 64*NR-row tiles whose last row is spilled to global memory
// (the reference's multi-tile kernel, pairhmm_kernels.cuh:2167-3272, keeps a
// border column the same way); the FP64 instance recomputes results whose
// FP32 sum fell below the rescue threshold.
#include "phmm_device.hpp"

namespace phmm {

__device__ __forceinline__ int shr1i(int v)
{
    // DPP wave_shr:1 -- lane i receives lane i-1; lane 0 receives 0.
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ float shr1f(float v)
{
    return __int_as_float(shr1i(__float_as_int(v)));
}
__device__ __forceinline__ double shr1d(double v)
{
    const int lo = shr1i(__double2loint(v));
    const int hi = shr1i(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = max(v, __shfl_xor(v, o, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ float ph2pr_of(const float* table, int q)
{
    return table[q > 127 ? 127 : q];
}

// ---------------------------------------------------------------------------
// k_fwd_tile
// ---------------------------------------------------------------------------
// One task: the W = 64/G alignments items[begin + task*W ...] of a bin.
template <int NR, bool FAST>
__device__ __forceinline__ void fwd_tile_task(const FwdParams& p, float4* tab, int begin,
                                              int count, int G, int task)
{
    constexpr int NQ = NR / 4;
    const int lane = threadIdx.x;
    const int W = 64 / G;
    const int grp = lane / G;
    const int t = lane - grp * G;
    const bool first = (t == 0);
    const int local = task * W + grp;
    const bool active = grp < W && local < count;

    int L = 0, HL = 0, roff = 0, hoff = 0, out = -1, idx = -1;
    float nl = 1.0f;
    if (active) {
        idx = begin + local;
        const WorkItem it = p.items[idx];
        out = it.out;
        L = p.read_len[it.read];
        HL = p.hap_len[it.hap];
        roff = p.read_off[it.read];
        hoff = p.hap_off[it.hap];
        if (FAST)
            nl = p.meta[it.read].nl;
    }
    float eps = kEpsGpu, beta = kBetaGpu;
    if (p.semantics == 1 && active && L > 0) {
        eps = ph2pr_of(p.ph2pr, p.gcp[roff]);
        beta = (float)(1.0 - (double)eps);
    }

    // ---- per-row coefficients and the emission table -------------------
    const int r0 = t * NR;
    float c0[NR], c1[NR], c2[NR], c3[NR];  // FAST: b d e -; SAFE: alpha w delta' e
    {
        float pdelta = 0.f, psigma = 0.f, palpha = 1.f;
        if (r0 >= 1 && r0 - 1 < L) {
            const uint32_t w = p.rows[roff + r0 - 1];
            pdelta = ph2pr_of(p.ph2pr, row_iq(w));
            psigma = ph2pr_of(p.ph2pr, row_dq(w));
            palpha = 1.0f - (pdelta + psigma);
        }
        float lm[NR], lx[NR];
        int code[NR];
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            const int ri = r0 + j;  // 0-based row; 1-based row ri+1
            if (ri < L) {
                const uint32_t w = p.rows[roff + ri];
                const float pb = ph2pr_of(p.ph2pr, row_bq(w));
                const float delta = ph2pr_of(p.ph2pr, row_iq(w));
                const float sigma = ph2pr_of(p.ph2pr, row_dq(w));
                const float alpha = 1.0f - (delta + sigma);
                code[j] = row_code(w);
                lm[j] = 1.0f - pb;
                lx[j] = pb / 3.0f;
                if (FAST) {
                    if (ri == 0) {
                        c0[j] = 0.f;
                        c1[j] = beta / alpha;
                        c2[j] = 0.f;
                    } else {
                        c0[j] = beta * pdelta / (palpha * alpha);
                        c1[j] = beta * psigma / alpha;
                        c2[j] = eps * pdelta / (delta * palpha);
                    }
                    c3[j] = 0.f;
                } else {
                    c0[j] = alpha;
                    c1[j] = ri == 0 ? beta : beta * psigma;
                    c2[j] = beta * delta;
                    c3[j] = eps;
                }
                pdelta = delta;
                psigma = sigma;
                palpha = alpha;
            } else {
                code[j] = 4;
                lm[j] = 0.f;
                lx[j] = 0.f;
                const bool firstoob = (ri == L);
                if (FAST) {
                    c0[j] = 0.f;
                    c1[j] = 0.f;
                    c2[j] = firstoob ? pdelta / palpha : 1.0f;
                    c3[j] = 0.f;
                } else {
                    c0[j] = 0.f;
                    c1[j] = 0.f;
                    c2[j] = firstoob ? beta : 0.f;
                    c3[j] = 1.0f;
                }
            }
        }
        // emission table: tab[(letter*NQ + q)*64 + lane] = rows 4q..4q+3
#pragma unroll
        for (int l = 0; l < 5; ++l) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                float v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int j = 4 * q + k;
                    const bool m = (l == code[j]) || (l == 4) || (code[j] == 4);
                    v[k] = m ? lm[j] : lx[j];
                }
                tab[(l * NQ + q) * 64 + lane] = make_float4(v[0], v[1], v[2], v[3]);
            }
        }
    }

    // ---- wavefront sweep -------------------------------------------------
    float M[NR], D[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        M[j] = 0.f;
        D[j] = 0.f;
    }
    float C0 = 0.f;
    if (active && HL > 0)
        C0 = FAST ? (kScale / (float)HL) * (nl * (1.0f / kFastKappa)) : kScale / (float)HL;
    float Ml = 0.f, Dl = first ? C0 : 0.f;
    float Mul = 0.f, Dul = Dl;
    float Iin = 0.f;
    float acc = 0.f;
    int letter = 4;

    const int Hmax = wave_max(HL);
    const int nsteps = Hmax + G;
    const uint32_t* hdw = reinterpret_cast<const uint32_t*>(p.hapc + hoff);
    const int hdw_n = active ? (HL >> 2) + 1 : 0;  // padded length / 4
    uint32_t wcur = (first && 0 < hdw_n) ? hdw[0] : 0x04040404u;
    uint32_t wnext = (first && 1 < hdw_n) ? hdw[1] : 0x04040404u;

    for (int s4 = 0; s4 < nsteps; s4 += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int sh = shr1i(letter);
            letter = first ? (int)((wcur >> (8 * u)) & 0xffu) : sh;
            const float4* trow = tab + letter * (NQ * 64) + lane;
            float mp = Mul, dp = Dul, chain = Iin;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const float4 lam4 = trow[q * 64];
                const float lam[4] = {lam4.x, lam4.y, lam4.z, lam4.w};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int j = 4 * q + k;
                    const float mo = M[j], dov = D[j];
                    float t1;
                    if (FAST) {
                        t1 = fmaf(c1[j], dp, mp);
                        t1 = fmaf(c0[j], chain, t1);
                        chain = fmaf(c2[j], chain, mp);
                    } else {
                        t1 = fmaf(c1[j], dp, chain);
                        t1 = fmaf(c0[j], mp, t1);
                        chain = fmaf(c2[j], mp, c3[j] * chain);
                    }
                    M[j] = lam[k] * t1;
                    D[j] = fmaf(eps, dov, mo);
                    mp = mo;
                    dp = dov;
                }
            }
            const int c = s4 + u - t;
            acc += (c >= 1 && c <= HL) ? chain : 0.f;
            const float sM = shr1f(M[NR - 1]);
            const float sD = shr1f(D[NR - 1]);
            const float sI = shr1f(chain);
            Mul = Ml;
            Dul = Dl;
            Ml = first ? 0.f : sM;
            Dl = first ? C0 : sD;
            Iin = first ? 0.f : sI;
        }
        const int k = (s4 >> 2) + 2;
        wcur = wnext;
        wnext = (first && k < hdw_n) ? hdw[k] : 0x04040404u;
    }

    if (active && t == L / NR) {
        float score, raw;
        if (FAST) {
            score = log10f(acc) - (kLog10Scale - kLog10FastKappa);
            raw = acc * nl;
        } else {
            score = log10f(acc) - (kLog10Scale + log10f(beta));
            raw = acc / beta;
        }
        p.scores[out] = score;
        if (p.rescue_on && !(raw >= p.rescue_min)) {
            const int k = atomicAdd(p.rescue_count, 1);
            p.rescue_list[k] = idx;
        }
    }
}

// Bin range from `range` (device {begin, count}) when non-null, else the
// host values; blocks loop over tasks so a fixed grid covers any count.
template <int NR, bool FAST>
__global__ __launch_bounds__(64) void k_fwd_tile(FwdParams p, const int32_t* range, int begin,
                                                 int count, int G)
{
    static_assert(NR % 4 == 0 && NR >= 4 && NR <= 32, "rows per lane");
    __shared__ float4 tab[5 * (NR / 4) * 64];
    if (range) {
        begin = range[0];
        count = range[1];
    }
    const int W = 64 / G;
    const int ntasks = (count + W - 1) / W;
    for (int task = blockIdx.x; task < ntasks; task += gridDim.x)
        fwd_tile_task<NR, FAST>(p, tab, begin, count, G, task);
}

// ---------------------------------------------------------------------------
// k_fwd_wave
// ---------------------------------------------------------------------------
template <typename T>
struct WaveScalar;
template <>
struct WaveScalar<float> {
    __device__ static float shr(float v) { return shr1f(v); }
    __device__ static float scale() { return kScale; }
};
template <>
struct WaveScalar<double> {
    __device__ static double shr(double v) { return shr1d(v); }
    __device__ static double scale() { return kScaleD; }
};

// Work source: items [*begin_ptr, *begin_ptr + *count_ptr) when list is
// null, else the item indices list[0 .. *count_ptr).  `spill` holds
// spill_elems values of T; each resident block uses 3*(hmax+1) of them.
template <typename T, int NR>
__global__ __launch_bounds__(64) void k_fwd_wave(FwdParams p, const int32_t* list,
                                                 const int32_t* begin_ptr,
                                                 const int32_t* count_ptr, const int32_t* hmax_ptr,
                                                 T* spill, int64_t spill_elems,
                                                 int32_t* error_flag)
{
    using S = WaveScalar<T>;
    constexpr int TR = 64 * NR;
    const int lane = threadIdx.x;
    const int n = *count_ptr;
    const int begin = list ? 0 : *begin_ptr;
    const int hmax = *hmax_ptr;
    const int64_t stride = 3 * (int64_t)(hmax + 1);
    const int64_t cap_blocks = spill_elems / stride;
    const int nblk = (int)(cap_blocks < (int64_t)gridDim.x ? cap_blocks : (int64_t)gridDim.x);
    if (nblk <= 0) {
        if (blockIdx.x == 0 && lane == 0 && n > 0)
            atomicOr(error_flag, 1);
        return;
    }
    if ((int)blockIdx.x >= nblk)
        return;
    T* sp = spill + (int64_t)blockIdx.x * stride;

    for (int w = blockIdx.x; w < n; w += nblk) {
        const int idx = list ? list[w] : begin + w;
        const WorkItem it = p.items[idx];
        const int L = p.read_len[it.read];
        const int HL = p.hap_len[it.hap];
        const int roff = p.read_off[it.read];
        const int hoff = p.hap_off[it.hap];
        if (L <= 0 || HL <= 0) {
            if (lane == 0)
                p.scores[it.out] = -INFINITY;
            continue;
        }
        if (HL > hmax) {
            if (lane == 0) {
                atomicOr(error_flag, 2);
                p.scores[it.out] = __int_as_float(0x7fc00000);
            }
            continue;
        }
        T eps = (T)kEpsGpu, beta = (T)kBetaGpu;
        if (p.semantics == 1) {
            const float e = ph2pr_of(p.ph2pr, p.gcp[roff]);
            eps = (T)e;
            beta = (T)1 - (T)e;
        }
        const T C0 = S::scale() / (T)HL;
        const int ntiles = (L + 1 + TR - 1) / TR;
        T acc = 0;
        for (int k = 0; k < ntiles; ++k) {
            const int r0 = k * TR + lane * NR;
            T ca[NR], cdl[NR], cs[NR], ce[NR], lm[NR], lx[NR];
            int code[NR];
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                const int ri = r0 + j;
                if (ri < L) {
                    const uint32_t w32 = p.rows[roff + ri];
                    const T pb = (T)ph2pr_of(p.ph2pr, row_bq(w32));
                    const T de = (T)ph2pr_of(p.ph2pr, row_iq(w32));
                    const T si = (T)ph2pr_of(p.ph2pr, row_dq(w32));
                    ca[j] = (T)1 - (de + si);
                    cdl[j] = de;
                    cs[j] = si;
                    ce[j] = eps;
                    lm[j] = (T)1 - pb;
                    code[j] = row_code(w32);
                    lx[j] = code[j] == 4 ? lm[j] : pb / (T)3;
                } else {
                    ca[j] = 0;
                    cdl[j] = ri == L ? (T)1 : (T)0;
                    cs[j] = 0;
                    ce[j] = 1;
                    lm[j] = 0;
                    lx[j] = 0;
                    code[j] = 4;
                }
            }
            T M[NR], D[NR];
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                M[j] = 0;
                D[j] = 0;
            }
            const bool top = (k == 0);
            T Ml = 0, Dl = (lane == 0 && top) ? C0 : (T)0, Mul = 0, Dul = Dl, Iin = 0;
            const bool last_tile = (k == ntiles - 1);
            const int tr = last_tile ? (L - k * TR) / NR : -1;
            const int nsteps = HL + 64;
            for (int s = 0; s < nsteps; ++s) {
                const int c = s - lane;  // this lane's 0-based column
                const int letter = (c >= 0 && c < HL) ? (int)p.hapc[hoff + c] : 4;
                T mp = Mul, dp = Dul, chain = Iin;
                if (lane == 0 && !top) {
                    // row above the tile at column c-1, from the previous tile
                    if (c >= 1) {
                        mp = sp[3 * (c - 1) + 0];
                        chain = sp[3 * (c - 1) + 1];
                        dp = sp[3 * (c - 1) + 2];
                    } else {
                        mp = chain = dp = 0;
                    }
                }
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    const T mo = M[j], dov = D[j];
                    const T lam = (letter == code[j] || letter == 4) ? lm[j] : lx[j];
                    M[j] = lam * (ca[j] * mp + beta * (chain + dp));
                    chain = cdl[j] * mp + ce[j] * chain;
                    D[j] = cs[j] * mo + eps * dov;
                    mp = mo;
                    dp = dov;
                }
                if (lane == tr && c >= 1 && c <= HL)
                    acc += chain;
                if (!last_tile && lane == 63) {
                    if (c >= 0 && c < HL) {
                        sp[3 * c + 0] = M[NR - 1];
                        sp[3 * c + 2] = D[NR - 1];
                    }
                    if (c >= 1 && c <= HL)
                        sp[3 * (c - 1) + 1] = chain;
                }
                const T sM = S::shr(M[NR - 1]);
                const T sD = S::shr(D[NR - 1]);
                const T sI = S::shr(chain);
                Mul = Ml;
                Dul = Dl;
                Ml = lane == 0 ? (T)0 : sM;
                Dl = lane == 0 ? (top ? C0 : (T)0) : sD;
                Iin = lane == 0 ? (T)0 : sI;
            }
            __syncthreads();  // spilled row visible to the next tile
        }
        // the result lane of the last tile owns acc
        const int tr = (L - (ntiles - 1) * TR) / NR;
        if (lane == tr)
            p.scores[it.out] = (float)(log10(acc) - log10(S::scale()));
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int NR>
static hipError_t launch_tile_nr(bool fast, const FwdParams& p, const int32_t* range, int begin,
                                 int count, int G, int max_blocks, hipStream_t st)
{
    int blocks = max_blocks;
    if (!range) {
        const int W = 64 / G;
        const int need = (count + W - 1) / W;
        if (need <= 0)
            return hipSuccess;
        blocks = max_blocks > 0 && max_blocks < need ? max_blocks : need;
    }
    if (blocks <= 0)
        return hipErrorInvalidValue;
    if (fast)
        hipLaunchKernelGGL((k_fwd_tile<NR, true>), dim3(blocks), dim3(64), 0, st, p, range, begin,
                           count, G);
    else
        hipLaunchKernelGGL((k_fwd_tile<NR, false>), dim3(blocks), dim3(64), 0, st, p, range, begin,
                           count, G);
    return hipGetLastError();
}

// max_blocks: grid cap (0: one block per task); with a device range the
// grid is max_blocks and the blocks loop over the bin's tasks.
hipError_t launch_fwd_tile(int NR, bool fast, const FwdParams& p, const int32_t* range, int begin,
                           int count, int G, int max_blocks, hipStream_t st)
{
    if (G < 1 || G > 64)
        return hipErrorInvalidValue;
    switch (NR) {
    case 4: return launch_tile_nr<4>(fast, p, range, begin, count, G, max_blocks, st);
    case 8: return launch_tile_nr<8>(fast, p, range, begin, count, G, max_blocks, st);
    case 12: return launch_tile_nr<12>(fast, p, range, begin, count, G, max_blocks, st);
    case 16: return launch_tile_nr<16>(fast, p, range, begin, count, G, max_blocks, st);
    case 20: return launch_tile_nr<20>(fast, p, range, begin, count, G, max_blocks, st);
    case 24: return launch_tile_nr<24>(fast, p, range, begin, count, G, max_blocks, st);
    case 28: return launch_tile_nr<28>(fast, p, range, begin, count, G, max_blocks, st);
    case 32: return launch_tile_nr<32>(fast, p, range, begin, count, G, max_blocks, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fwd_wave_f32(const FwdParams& p, const int32_t* list, const int32_t* begin_ptr,
                               const int32_t* count_ptr, const int32_t* hmax_ptr, float* spill,
                               int64_t spill_elems, int32_t* error_flag, int blocks,
                               hipStream_t st)
{
    hipLaunchKernelGGL((k_fwd_wave<float, 16>), dim3(blocks), dim3(64), 0, st, p, list, begin_ptr,
                       count_ptr, hmax_ptr, spill, spill_elems, error_flag);
    return hipGetLastError();
}

hipError_t launch_fwd_wave_f64(const FwdParams& p, const int32_t* list, const int32_t* begin_ptr,
                               const int32_t* count_ptr, const int32_t* hmax_ptr, double* spill,
                               int64_t spill_elems, int32_t* error_flag, int blocks,
                               hipStream_t st)
{
    hipLaunchKernelGGL((k_fwd_wave<double, 8>), dim3(blocks), dim3(64), 0, st, p, list, begin_ptr,
                       count_ptr, hmax_ptr, spill, spill_elems, error_flag);
    return hipGetLastError();
}

}  // namespace phmm
