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
// k_fwd_tile_set -- the hot path.  A wave holds 64/G alignments ("groups") of G
// lanes; lane t of a group holds read rows [t*NR, t*NR+NR) in registers and
// the haplotype streams through the group one column per step (anti-diagonal
// wavefront).  Row-above values cross lanes with DPP wave_shr:1, a group's
// first lane injects the row-0 boundary.  Emission probabilities for the
// 4 bases x NR rows of each lane sit in LDS (one ds_read_b128 per 4 cells,
// conflict-free by lane interleaving; haps holding any other byte go to the
// one-wave kernel).  Its work comes in two units: super-tasks (up to 64/G
// reads of one record against a segment of the record's haps: the wave sets
// the reads up once and sweeps the haps one after another) and items (one
// alignment per group: reads a record leaves over after full chunks, and the
// FP64 rescue lists).  Two arithmetic forms:
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
//    0 < alpha <= 1, e_r <= 1/2, eps <= 1/2 and n_R >= 1e-30.
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
    // DPP wave_shr:1 -- lane i receives lane i-1; lane 0 receives 0 (bound
    // control: an out-of-range source reads as zero).
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
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

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = min(v, __shfl_xor(v, o, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

// Setup-time division (once per task and row).  FP32: the hardware
// reciprocal times the numerator, within 2 ulp of the quotient.  The
// correctly rounded division is a chain of dependent instructions per row
// that stalls the wave's setup: with it the forward kernel runs about 3%
// slower on 150-base reads against 150-350-base haps.  FP64 (the rescue)
// keeps the correctly rounded quotient.
__device__ __forceinline__ float qdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double qdiv(double a, double b) { return a / b; }

__device__ __forceinline__ float ph2pr_of(const float* table, int q)
{
    return table[q > 127 ? 127 : q];
}

// ---------------------------------------------------------------------------
// k_fwd_tile_set
// ---------------------------------------------------------------------------
template <typename T>
struct TileTraits;
template <>
struct TileTraits<float> {
    using V = float4;  // one ds_read_b128 = 4 rows
    static constexpr int K = 4;
    __device__ static V pack(const float* v) { return make_float4(v[0], v[1], v[2], v[3]); }
    __device__ static void unpack(const V& x, float* v)
    {
        v[0] = x.x;
        v[1] = x.y;
        v[2] = x.z;
        v[3] = x.w;
    }
    __device__ static float shr(float x) { return shr1f(x); }
};
template <>
struct TileTraits<double> {
    using V = double2;  // one ds_read_b128 = 2 rows
    static constexpr int K = 2;
    __device__ static V pack(const double* v) { return make_double2(v[0], v[1]); }
    __device__ static void unpack(const V& x, double* v)
    {
        v[0] = x.x;
        v[1] = x.y;
    }
    __device__ static double shr(double x) { return shr1d(x); }
};

// Shape of a lane's emission table in LDS: per letter, NV vectors of K rows
// (the last may hold fewer), tab[(letter*NV + q)*64 + lane] -- one ds_read
// of 16 bytes per vector, conflict-free because consecutive lanes read
// consecutive 16-byte words.  A last vector holding a single row would be
// read as one 4-byte word at a 16-byte lane stride, four lanes to a bank,
// so such a row (kTail) sits in a table of scalars instead,
// stab[letter*64 + lane], at a 4-byte lane stride.
template <typename T, int NR>
struct TabShape {
    static constexpr int K = TileTraits<T>::K;
    static constexpr bool kTail = K > 1 && NR % K == 1;
    static constexpr int NV = kTail ? NR / K : (NR + K - 1) / K;
    static constexpr int kVecs = 4 * NV * 64;
    static constexpr int kWords =  // table size in units of the vector type
        kVecs + (kTail ? (4 * 64 * (int)sizeof(T) + (int)sizeof(typename TileTraits<T>::V) - 1) /
                             (int)sizeof(typename TileTraits<T>::V)
                       : 0);
};

// Per-lane state of a tile task that depends only on the read: its rows'
// coefficients (the emission values of the rows live in the LDS table).
// ROWEPS: every row has its own gap continuation eps_r (per-position gcp
// semantics), so the D chain takes eps_r from a per-row register (FAST: c3;
// SAFE: c4) instead of the lane's eps.
template <typename T, int NR, bool ROWEPS>
struct TileRead {
    T c0[NR], c1[NR], c2[NR], c3[NR];  // FAST: b d e eps_r; SAFE: alpha w delta' e
    T c4[ROWEPS ? NR : 1];             // SAFE with ROWEPS: eps_r
    T eps, beta;
    int L;
    float nl;
};

// Gap continuation of the D chain of row j.
template <typename T, int NR, bool FAST, bool ROWEPS>
__device__ __forceinline__ T d_eps(const TileRead<T, NR, ROWEPS>& R, int j)
{
    if constexpr (!ROWEPS)
        return R.eps;
    else if constexpr (FAST)
        return R.c3[j];
    else
        return R.c4[j];
}

// beta = 1 - eps, rounded as the reference's host code rounds it in FP32
// (float beta = 1.0 - ph2pr[q], gpuworker.cuh:1048).
template <typename T>
__device__ __forceinline__ T beta_of(T e)
{
    if constexpr (sizeof(T) == 4)
        return (T)(float)(1.0 - (double)e);
    else
        return (T)1 - e;
}

// This lane's rows of `read` (inactive lanes: an empty read): coefficients in
// R, emission values in the LDS table.
//
// ROWEPS (per-position gcp): row r (1-based) takes eps_r = ph2pr[gcp[r-1]]
// and beta_r = 1 - eps_r where the reference takes one eps and beta.  FAST:
// b_r = beta_r delta_{r-1}/(alpha_{r-1} alpha_r), d_r = beta_r sigma_{r-1} /
// alpha_r, e_r = eps_r delta_{r-1}/(delta_r alpha_{r-1}), and the D chain
// Dh = Mh< + eps_r Dh<.  SAFE carries Ib_r = beta_{r+1} I_r (the next row's
// beta, so row r+1's M adds it with weight 1): c2 = beta_{r+1} delta_r,
// c3 = eps_r beta_{r+1} / beta_r, and D chain Dh = M< + eps_r Dh<; the last
// row's beta_{L+1} is taken as beta_L, which the epilogue divides out.
template <typename T, int NR, bool FAST, bool ROWEPS>
__device__ __forceinline__ void tile_setup(const FwdParams& p, typename TileTraits<T>::V* tab,
                                           bool active, int read, int t,
                                           TileRead<T, NR, ROWEPS>& R)
{
    const BatchDesc& bd = *p.bd;
    using TT = TileTraits<T>;
    constexpr int K = TT::K;
    const int lane = threadIdx.x;
    int L = 0, roff = 0;
    float nl = 1.0f;
    if (active) {
        L = bd.read_len[read];
        roff = bd.read_off[read];
        if (FAST)
            nl = p.meta[read].nl;
    }
    T eps = (T)kEpsGpu, beta = (T)kBetaGpu;
    if (p.semantics != 0 && active && L > 0) {
        // reference CPU semantics: the read's first gcp quality; per-position
        // gcp: the last row's (the SAFE epilogue divides by beta_{L+1} = beta_L)
        const float e = ph2pr_of(p.ph2pr, bd.gcp[roff + (ROWEPS ? L - 1 : 0)]);
        eps = (T)e;
        beta = beta_of((T)e);
    }
    R.eps = eps;
    R.beta = beta;
    R.L = L;
    R.nl = nl;
    // per-position gcp: eps and beta of row index ri (0-based, < L)
    auto row_eps = [&](int ri) -> T { return (T)ph2pr_of(p.ph2pr, bd.gcp[roff + ri]); };

    // ---- per-row coefficients and the emission table -------------------
    const int r0 = t * NR;
    {
        T pdelta = 0, psigma = 0, palpha = 1;
        if (r0 >= 1 && r0 - 1 < L) {
            const uint32_t w = p.rows[roff + r0 - 1];
            pdelta = (T)ph2pr_of(p.ph2pr, row_iq(w));
            psigma = (T)ph2pr_of(p.ph2pr, row_dq(w));
            palpha = (T)1 - (pdelta + psigma);
        }
        T lm[NR], lx[NR];
        int code[NR];
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            const int ri = r0 + j;  // 0-based row; 1-based row ri+1
            if (ri < L) {
                const uint32_t w = p.rows[roff + ri];
                const T pb = (T)ph2pr_of(p.ph2pr, row_bq(w));
                const T delta = (T)ph2pr_of(p.ph2pr, row_iq(w));
                const T sigma = (T)ph2pr_of(p.ph2pr, row_dq(w));
                const T alpha = (T)1 - (delta + sigma);
                // gap continuation of this row, and (SAFE) the next row's beta
                T er = eps, br = beta, bn = beta;
                if constexpr (ROWEPS) {
                    er = row_eps(ri);
                    br = beta_of(er);
                    bn = ri + 1 < L ? beta_of(row_eps(ri + 1)) : beta;
                }
                code[j] = row_code(w);
                lm[j] = (T)1 - pb;
                lx[j] = qdiv(pb, (T)3);
                if (FAST) {
                    if (ri == 0) {
                        R.c0[j] = 0;
                        R.c1[j] = qdiv(br, alpha);
                        R.c2[j] = 0;
                    } else {
                        R.c0[j] = qdiv(br * pdelta, palpha * alpha);
                        R.c1[j] = qdiv(br * psigma, alpha);
                        R.c2[j] = qdiv(er * pdelta, delta * palpha);
                    }
                    R.c3[j] = ROWEPS ? er : (T)0;
                } else {
                    R.c0[j] = alpha;
                    R.c1[j] = ri == 0 ? br : br * psigma;
                    R.c2[j] = bn * delta;
                    // equal betas keep eps exactly (no rounding of bn / br)
                    R.c3[j] = bn == br ? er : qdiv(er * bn, br);
                    if constexpr (ROWEPS)
                        R.c4[j] = er;
                }
                pdelta = delta;
                psigma = sigma;
                palpha = alpha;
            } else {
                code[j] = 4;
                lm[j] = 0;
                lx[j] = 0;
                const bool firstoob = (ri == L);
                if (FAST) {
                    R.c0[j] = 0;
                    R.c1[j] = 0;
                    R.c2[j] = firstoob ? qdiv(pdelta, palpha) : (T)1;
                    R.c3[j] = 0;
                } else {
                    R.c0[j] = 0;
                    R.c1[j] = 0;
                    R.c2[j] = firstoob ? beta : (T)0;
                    R.c3[j] = 1;
                    if constexpr (ROWEPS)
                        R.c4[j] = 0;
                }
            }
        }
        // emission table: tab[(letter*NV + q)*64 + lane] = rows K*q .. K*q+K-1
        // for the four bases (haps with any other byte go to the one-wave
        // kernel; a read's 'N' matches every base), a single last row in the
        // scalar table
        using Shape = TabShape<T, NR>;
        constexpr int NV = Shape::NV;
#pragma unroll
        for (int l = 0; l < 4; ++l) {
#pragma unroll
            for (int q = 0; q < NV; ++q) {
                T v[K];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int j = K * q + k;
                    if (j < NR) {
                        const bool m = (l == code[j]) || (code[j] == 4);
                        v[k] = m ? lm[j] : lx[j];
                    } else {
                        v[k] = 0;  // past the lane's rows: never used
                    }
                }
                tab[(l * NV + q) * 64 + lane] = TT::pack(v);
            }
            if constexpr (Shape::kTail) {
                constexpr int j = NR - 1;
                const bool m = (l == code[j]) || (code[j] == 4);
                reinterpret_cast<T*>(tab + Shape::kVecs)[l * 64 + lane] = m ? lm[j] : lx[j];
            }
        }
    }

}

// One sweep of this group's read (R) against a hap of length HL at hoff;
// the group's result lane writes the score of `it`.  T = float is the first
// pass: results whose FP32 sum (in the reference's units) is below
// rescue_min are appended to the bin's rescue list (rescue_items[rescue_base
// + n], n counted in rescue_counts[bin]).  T = double is the FP64 recompute
// of those lists (start constant 2^1000/H); a result whose sum is not far
// above double's floor goes to the final list.
//
// RE (T = float): the range-extended recompute of the rescue lists.  Lane t
// scales its values by 2^E_t (FwdParams: kReShift and the rule for E_t) --
// the mask multiply that injects the boundary on a group's first lane
// carries the factor 2^(E_t - E_{t-1}) on the others, so the sweep itself is
// unchanged -- and an accepted sum a is exactly a * 2^(kReShift - E_R) of
// the FP32 pass's units.  The magnitude of a lane's last row is read from
// its accumulator: every lane adds up the chain it passes down (only the
// result lane's sum is a score), so a lane above the result lane holds the
// sum over the columns of its last row's outgoing chain -- at no cost in
// the sweep.  The first pass keeps its ilogb for the recompute's scales,
// the recompute checks them with it.
// Returns, for the scaled instances (DS), the number of the wave's
// alignments that fell below the rescue threshold and were finished by the
// scaling (uniform across the wave); 0 otherwise.
// GE: the reference GPU semantics, whose gap continuation is the constant
// kEpsGpu for every read: the D chain's multiplier is then wave-uniform and
// sits in a scalar register, so those FMAs read two vector operands, not
// three (about 1.4% of the sweep's time on gfx950).
template <typename T, int NR, bool FAST, bool ROWEPS, bool RE = false, bool DS = false,
          bool GE = false>
__device__ __forceinline__ int tile_sweep(const FwdParams& p, const typename TileTraits<T>::V* tab,
                                          const TileRead<T, NR, ROWEPS>& R, bool active, int G,
                                          int t, int HL, int hoff, const WorkItem& it,
                                          int rescue_base, int bin)
{
    const BatchDesc& bd = *p.bd;
    using TT = TileTraits<T>;
    using V = typename TT::V;
    constexpr int K = TT::K;
    constexpr bool F32 = sizeof(T) == 4;
    const int lane = threadIdx.x;
    const bool first = (t == 0);
    const T beta = R.beta;
    const int L = R.L;
    const float nl = R.nl;
    constexpr bool REC = F32 && !ROWEPS && re_rows(NR) && !DS;  // record per-lane magnitudes
    static_assert(!RE || REC, "the range-extended pass is an FP32 pass of the one-eps tiles");
    static_assert(!DS || (F32 && !ROWEPS && !RE && re_rows(NR)),
                  "in-pass scaling: FP32 one-eps tiles of the range-extended rows");
    // the group's lane that forms the sum, the group's first lane and its
    // lanes as a mask (computed where used: nothing extra stays live in
    // the sweep)
    auto res_lane = [&]() { return L / NR; };
    auto group_mask = [&]() {
        return G >= 64 ? ~0ull : (((1ull << G) - 1ull) << (lane - t));
    };

    // ---- wavefront sweep -------------------------------------------------
    T M[NR], D[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        M[j] = 0;
        D[j] = 0;
    }
    T C0 = 0;
    if (active && HL > 0) {
        if (F32)
            C0 = FAST ? (T)((kScale / (float)HL) * (nl * (1.0f / kFastKappa))) : (T)(kScale / (float)HL);
        else
            C0 = FAST ? (T)((kStartF64 / (double)HL) * ((double)nl / (double)kFastKappa))
                      : (T)(kStartF64 / (double)HL);
    }
    // boundary injection: a group's first lane takes row 0 (M = I = 0,
    // D = C0) instead of its left neighbour's last row -- as a multiply by a
    // 0/1 lane mask, which issues at the full FP32 rate where a select does not
    T nf = first ? (T)0 : (T)1;
    int re_E = 0, re_dE = 0;
    bool re_bad = false;
    if constexpr (RE) {
        // per-lane exponents from the FP32 pass's maxima r_k (all lanes of
        // the wave take part in the ballot and shuffles)
        const int slot = active ? it.pad : -1;  // tile_items_task: the rescue-list slot
        const int cm = active && HL > 0 ? ilogbf((float)C0) : 0;
        const int res_t = res_lane(), gbase = lane - t;
        const unsigned long long gmask = group_mask();
        int r = 127;
        if (active && slot >= 0 && t < res_t)
            r = t < kReKnownLanes ? (int)(int8_t)p.rescue_exps[(int64_t)slot * kReKnownLanes + t]
                                  : -128;  // not recorded: extrapolated like an underflow
        re_bad = active && slot < 0;
        const unsigned long long inv = __ballot(active && t < res_t && r < kReValid) & gmask;
        const int lv = (inv ? __ffsll((long long)inv) - 1 - gbase : res_t) - 1;  // last known lane
        const int r_lv = __shfl(r, gbase + (lv > 0 ? lv : 0), 64);
        const int r_lv1 = __shfl(r, gbase + (lv > 1 ? lv - 1 : 0), 64);
        // lanes past the result lane (rows past the read) keep its scale
        const int te = t < res_t ? t : res_t;
        const int r_prev = __shfl(r, gbase + (te > 0 ? te - 1 : 0), 64);
        int E = 0;
        if (te >= 1) {
            if (te - 1 <= lv) {
                E = cm - r_prev;
            } else {
                int slope = lv >= 1 ? r_lv - r_lv1 : lv == 0 ? r_lv - cm : kReUnknownSlope;
                slope = slope < 0 ? slope : 0;
                const int e1 = lv >= 0 ? cm - r_lv : 0;  // E of lane lv + 1
                E = e1 - slope * (te - lv - 1);
            }
        }
        const int Ep = __shfl(E, lane > 0 ? lane - 1 : 0, 64);
        re_dE = t >= 1 ? E - Ep : 0;
        if (active && t <= res_t && (re_dE > 126 || re_dE < -126))
            re_bad = true;
        re_E = E;
        if (!first)
            nf = ldexpf(1.0f, re_dE < -126 ? -126 : re_dE > 126 ? 126 : re_dE);
        C0 = ldexpf((float)C0, -kReShift);
    }
    int ds_E = 0;  // DS: this lane's exponent (held = true * 2^ds_E)
    bool ds_any = false;  // DS: some lane of the wave is scaled (uniform)
    // DS: x capped at kDsSpan per lane above the group's lanes over it
    // (x_t <= x_k + kDsSpan * (t - k) for k < t: a prefix minimum of
    // x - kDsSpan * t)
    auto ds_span_cap = [&](int x) {
        int z = x - kDsSpan * t;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(z, o, 64);
            if (t >= o)
                z = y < z ? y : z;
        }
        return z + kDsSpan * t;
    };
    // DS: the lane above's value of an exponent.  The shifted value is
    // kept opaque to the compiler: fused into the subtraction that uses it
    // (v_subrev_u32_dpp), it measured as E_{t-1} - E_t on gfx950.
    auto ds_above = [&](int v) {
        int u = shr1i(v);
        asm volatile("" : "+v"(u));
        return u;
    };
    // DS: the transfer factor 2^(E_t - E_{t-1}), exact as long as the
    // exponents stay within kDsSpan
    auto ds_factor = [&]() -> T {
        const int d = ds_E - ds_above(ds_E);
        return first ? (T)0 : (T)ldexpf(1.0f, d);
    };
    if constexpr (DS) {
        // The insertion floor at the lane's last row R (the read's last
        // row for the lanes past it): the product of the rows' chain
        // coefficients from row 1 to R, which telescopes -- FAST: eps^R *
        // delta_0 / delta_R / prod(alpha_0 .. alpha_{R-1}); SAFE: eps per
        // row.  The alpha factors (each <= 1) are left out, which can only
        // put the estimate below the true floor.  No lane exchange needed.
        float fl = 0.0f;
        if (active && C0 > 0 && L > 0) {
            const int last = min((t + 1) * NR, L) - 1;
            fl = __log2f((float)C0) + (float)last * __log2f((float)R.eps);
            if (FAST && last > 0) {
                const uint32_t* rw = p.rows + bd.read_off[it.read];
                fl += __log2f(ph2pr_of(p.ph2pr, row_iq(rw[0]))) -
                      __log2f(ph2pr_of(p.ph2pr, row_iq(rw[last])));
            }
        }
        // the floor of the lane's last row at 2^kDsFloor
        const int e0 = (int)floorf((float)kDsFloor - fl);
        int x = first || e0 < 0 ? 0 : e0;
        if (__ballot(!first && x - ds_above(x) > kDsSpan))  // steep gap scores only
            x = ds_span_cap(x);
        ds_E = x;
        nf = ds_factor();
        ds_any = __ballot(ds_E > 0) != 0;
    }
    const T c0f = first ? C0 : (T)0;
    // M and D of the row above reach a lane two steps after they were
    // computed (diagonal dependency), I one step after: step u reads ring
    // entry u % 2 and refills it, so the unrolled steps need no copies
    T Mr[2] = {(T)0, (T)0}, Dr[2] = {c0f, c0f};
    T Iin = 0;
    T acc = 0;

    const int Hmax = wave_max(HL);
    const int nsteps = Hmax + G;
    // Hap letters: lane t consumes column c = s - t at step s.  Each lane
    // keeps a window of two dwords of its hap codes and shifts it by
    // (c mod 4) bytes; dword indices are clamped to the padded hap, so every
    // column up to the padding is exact, and columns outside it (before the
    // lane's first column or past the hap's end) only meet zero state or are
    // never accumulated.
    const uint32_t* hdw = reinterpret_cast<const uint32_t*>(p.hapc + (active ? hoff : 0));
    const int hdw_last = active ? (HL >> 2) : 0;  // padded length / 4 - 1
    const uint32_t shift = (uint32_t)(-t) & 3u;     // (s4 - t) mod 4, constant per lane
    auto hap_dword = [&](int q) { return hdw[min(max(q, 0), hdw_last)]; };
    const int q0 = (0 - t) >> 2;  // floor((s4 - t) / 4) of the first block
    uint32_t wlo = hap_dword(q0), whi = hap_dword(q0 + 1), wnx = hap_dword(q0 + 2);
    // Step s adds the chain of column s - t.  The chain is exactly 0 before
    // a lane's first column, so only the end of the lane's real columns,
    // s <= t + HL, needs a bound: steps below the wave's smallest HL add
    // unconditionally, the rest compare.
    const int acc_last = t + HL;
    const int n_free = wave_min(active ? HL : nsteps) + 1;  // steps s < n_free
    const int lane_off = lane;  // tab index of this lane's letter-0 entry

    // emission values of step 0; each step loads the next step's values
    // before it computes
    using Shape = TabShape<T, NR>;
    constexpr int NV = Shape::NV;
    const T* stab = reinterpret_cast<const T*>(tab + Shape::kVecs) + lane_off;
    V lam[2][NV];
    T lamS[2] = {(T)0, (T)0};  // the scalar-table row (Shape::kTail)
    uint32_t letters = __builtin_amdgcn_alignbyte(whi, wlo, shift);
    {
        const V* trow = tab + __umul24(letters & 3u, NV * 64) + lane_off;
#pragma unroll
        for (int qq = 0; qq < NV; ++qq)
            lam[0][qq] = trow[qq * 64];
        if constexpr (Shape::kTail)
            lamS[0] = stab[__umul24(letters & 3u, 64)];
    }

    // GE: every lane's eps is kEpsGpu, held in a scalar register (opaque to
    // the compiler, which would otherwise keep the constant in a vector one)
    T eps_ge = (T)0;
    if constexpr (GE) {
        int e = __builtin_bit_cast(int, kEpsGpu);
        asm volatile("" : "+s"(e));
        eps_ge = (T)__builtin_bit_cast(float, e);
    }
    auto sweep4 = [&](int s4, bool bounded) {
        // letters of steps s4+1 .. s4+4: the current window and the next one
        const uint32_t next = __builtin_amdgcn_alignbyte(wnx, whi, shift);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int cur = u & 1, nxt = cur ^ 1;
            // codes are 0..3 inside the hap; the padding's wildcard code
            // reads as base 0 there, and padding columns are never accumulated
            const uint32_t nletter = u < 3 ? (letters >> (8 * (u + 1))) & 3u : next & 3u;
            {
                const V* trow = tab + __umul24(nletter, NV * 64) + lane_off;
#pragma unroll
                for (int qq = 0; qq < NV; ++qq)
                    lam[nxt][qq] = trow[qq * 64];
                if constexpr (Shape::kTail)
                    lamS[nxt] = stab[__umul24(nletter, 64)];
            }
            T mp = Mr[cur], dp = Dr[cur], chain = Iin;
            // the cell of row j with emission value lvk
            auto cell = [&](int j, T lvk) {
                const T mo = M[j], dov = D[j];
                T t1;
                if (FAST) {
                    t1 = fma(R.c1[j], dp, mp);
                    t1 = fma(R.c0[j], chain, t1);
                    chain = fma(R.c2[j], chain, mp);
                } else {
                    t1 = fma(R.c1[j], dp, chain);
                    t1 = fma(R.c0[j], mp, t1);
                    chain = fma(R.c2[j], mp, R.c3[j] * chain);
                }
                M[j] = lvk * t1;
                D[j] = fma(GE ? eps_ge : d_eps<T, NR, FAST, ROWEPS>(R, j), dov, mo);
                mp = mo;
                dp = dov;
            };
#pragma unroll
            for (int qq = 0; qq < NV; ++qq) {
                T lv[K];
                TT::unpack(lam[cur][qq], lv);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int j = K * qq + k;
                    if (j >= (Shape::kTail ? NR - 1 : NR))
                        break;
                    cell(j, lv[k]);
                }
            }
            if constexpr (Shape::kTail)
                cell(NR - 1, lamS[cur]);
            if (bounded)
                acc += (s4 + u <= acc_last) ? chain : (T)0;
            else
                acc += chain;
            if constexpr (RE) {
                // a group's first lane must not see its neighbour's values:
                // scaled by that alignment's lane factors they may overflow,
                // and inf * 0 would poison this group.  The shifted value is
                // clamped to FLT_MAX first (v_min also maps NaN to it), then
                // masked as in the first pass.  (Not a select: the compiler
                // turns a select of a DPP result into an exec-masked DPP,
                // which reads 0 from the masked-off source lanes.)
                constexpr float kBig = 3.4028234663852886e38f;
                Mr[cur] = fminf(TT::shr(M[NR - 1]), kBig) * nf;
                Dr[cur] = fma(fminf(TT::shr(D[NR - 1]), kBig), nf, c0f);
                Iin = fminf(TT::shr(chain), kBig) * nf;
            } else {
                Mr[cur] = TT::shr(M[NR - 1]) * nf;
                Dr[cur] = fma(TT::shr(D[NR - 1]), nf, c0f);
                Iin = TT::shr(chain) * nf;
            }
        }
        letters = next;
        wlo = whi;
        whi = wnx;
        wnx = hap_dword(((s4 - t) >> 2) + 3);  // arithmetic shift: floor
    };
    // DS: every kDsEvery steps a scaled lane whose incoming values passed
    // 2^kDsGuard scales down to put them at 2^kDsGuardTo (never below
    // E = 0), rescaling what it holds (the values already taken from the
    // lane above included) and its transfer factor.  Called at wave-uniform
    // points; every lane of the wave takes part in the exchange.
    auto ds_check = [&]() {
        // values are never negative: no absolute values needed
        const float m = fmaxf(fmaxf(fmaxf((float)Mr[0], (float)Mr[1]), fmaxf((float)Dr[0], (float)Dr[1])),
                              (float)Iin);
        int x = ds_E;
        if (ds_E > 0 && m > kDsGuardF) {
            // v_frexp_exp reads inf as 0: an overflowed lane keeps its
            // exponent and fails acceptance
            const int g = ds_E + kDsGuardTo - (__builtin_amdgcn_frexp_expf(m) - 1);
            x = g < 0 ? 0 : g;
        }
        // any distance in one step: v_ldexp is exact wherever the result
        // stays normal
        if (__ballot(x != ds_E)) {
            // no lane more than kDsSpan above the lane over it, so the
            // transfer factor stays a normal float (a lane whose upper rows
            // meet a read's alignment scales down before its last row, which
            // the lane below sees, does); the scan only where a gap opened
            const int xu = ds_above(x);
            if (__ballot(!first && x - xu > kDsSpan))
                x = ds_span_cap(x);
            const int S = x - ds_E;
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                M[j] = ldexpf(M[j], S);
                D[j] = ldexpf(D[j], S);
            }
            Mr[0] = ldexpf(Mr[0], S);
            Mr[1] = ldexpf(Mr[1], S);
            Dr[0] = ldexpf(Dr[0], S);
            Dr[1] = ldexpf(Dr[1], S);
            Iin = ldexpf(Iin, S);
            acc = ldexpf(acc, S);
            ds_E = x;
            nf = ds_factor();
            ds_any = __ballot(ds_E > 0) != 0;
        }
    };
    int s4 = 0;
    if constexpr (DS) {
        // kDsEvery steps between checks with no branch in between: a
        // branch to the check after every 4 steps cost about 1.3% of the
        // scaled pass (measured on the mixed-length workload)
        static_assert(kDsEvery == 16, "four sweep4 blocks per check");
        for (; s4 + 16 <= n_free && s4 + 12 < nsteps; s4 += 16) {
            sweep4(s4, false);
            sweep4(s4 + 4, false);
            sweep4(s4 + 8, false);
            sweep4(s4 + 12, false);
            if (ds_any)
                ds_check();
        }
    }
    for (; s4 + 4 <= n_free && s4 < nsteps; s4 += 4) {
        sweep4(s4, false);
        if constexpr (DS)
            if (ds_any && (s4 & (kDsEvery - 4)) == kDsEvery - 4)
                ds_check();
    }
    for (; s4 < nsteps; s4 += 4) {
        sweep4(s4, true);
        if constexpr (DS)
            if (ds_any && (s4 & (kDsEvery - 4)) == kDsEvery - 4)
                ds_check();
    }

    if constexpr (RE) {
        // each lane's incoming magnitude, in its own scale: the lane above's
        // chain sum times 2^(E_t - E_{t-1}), within the window
        const float rmp = shr1f((float)acc);
        const int res_t = res_lane();
        bool bad = re_bad;
        if (active && t >= 1 && t <= res_t) {
            if (!(rmp > 0.0f && rmp <= 3.4028234663852886e38f)) {
                bad = true;
            } else {
                const int li = ilogbf(rmp) + re_dE;
                const int cre = ilogbf((float)C0);
                bad = bad || li < cre - kReLow || li > cre + kReHigh;
            }
        }
        const unsigned long long fb = __ballot(bad) & group_mask();
        if (active && t == res_t) {
            const float a = (float)acc;
            if (!fb && a >= 0x1p-100f && a <= 3.4028234663852886e38f) {
                // a * 2^(kReShift - E) = m * 2^K exactly: the score does not
                // depend on how the rows were split into lanes
                int e2 = 0;
                const float m = frexpf(a, &e2);
                const int K = e2 + kReShift - re_E;
                const double off = FAST ? (double)kLog10Scale - (double)kLog10FastKappa
                                        : (double)kLog10Scale + log10((double)(float)beta);
                bd.scores[it.out] =
                    (float)(log10((double)m) + (double)K * 0.30102999566398119521 - off);
            } else {
                const int n = atomicAdd(p.retry_counts + bin, 1);
                p.retry_items[rescue_base + n] = it;
            }
        }
        return 0;
    }
    // DS acceptance: every scaled lane up to the result lane holds a finite
    // running sum (an unscaled lane's sum can overflow harmlessly in the
    // FAST form's upper rows; an overflowed value reaches the result lane's
    // sum), and neighbouring exponents keep the transfer factor a normal
    // float.  Nothing that mattered was flushed: a lane starts with its
    // insertion floor in range and scales down only by what its largest
    // incoming value grew past 2^kDsGuard.
    unsigned long long ds_fb = 0;
    bool ds_done = false;  // finished by the scaling below the rescue threshold
    if constexpr (DS) {
        const int res_t = res_lane();
        const int dl = ds_E - ds_above(ds_E);
        bool bad = false;
        if (active && t <= res_t)
            bad = (ds_E > 0 && !((float)acc <= 3.4028234663852886e38f)) ||
                  (t >= 1 && (dl > 127 || dl < -126));
        ds_fb = __ballot(bad) & group_mask();
    }
    int slot = -1;  // REC: the rescued alignment's rescue-list slot
    if (active && t == L / NR) {
        if (DS && (ds_fb || ds_E > 0 || !((float)acc <= 3.4028234663852886e38f))) {
            const float a = (float)acc;
            if (!ds_fb && a >= 0x1p-100f && a <= 3.4028234663852886e38f) {
                // the unscaled sum, where it is a normal float: the score and
                // threshold test of the unscaled instance, bit for bit
                const float at = ldexpf(a, -ds_E);
                float raw;
                if (at >= 0x1p-126f) {
                    bd.scores[it.out] = FAST ? log10f(at) - (kLog10Scale - kLog10FastKappa)
                                             : log10f(at) - (kLog10Scale + log10f((float)beta));
                    raw = FAST ? at * nl : at / (float)beta;
                } else {
                    // a * 2^-E = m * 2^K exactly: the score does not depend
                    // on how the rows were split into lanes
                    int e2 = 0;
                    const float m = frexpf(a, &e2);
                    const int K = e2 - ds_E;
                    const double off = FAST ? (double)kLog10Scale - (double)kLog10FastKappa
                                            : (double)kLog10Scale + log10((double)(float)beta);
                    bd.scores[it.out] =
                        (float)(log10((double)m) + (double)K * 0.30102999566398119521 - off);
                    raw = ldexpf(FAST ? a * nl : a / (float)beta, -ds_E);
                }
                ds_done = !(raw >= p.rescue_min);
            } else {
                bd.scores[it.out] = -INFINITY;
                const int n = atomicAdd(p.rescue_counts + bin, 1);
                p.rescue_items[rescue_base + n] = it;
            }
        } else if (F32) {
            const float a = (float)acc;
            float score, raw;
            if (FAST) {
                score = log10f(a) - (kLog10Scale - kLog10FastKappa);
                raw = a * nl;
            } else {
                score = log10f(a) - (kLog10Scale + log10f((float)beta));
                raw = a / (float)beta;
            }
            bd.scores[it.out] = score;
            if (p.rescue_on && !(raw >= p.rescue_min)) {
                const int n = atomicAdd(p.rescue_counts + bin, 1);
                p.rescue_items[rescue_base + n] = it;
                if (REC && p.re_on)
                    slot = rescue_base + n;
            }
        } else {
            const double a = (double)acc;
            bd.scores[it.out] = FAST ? (float)(log10(a) - (kLog10StartF64 - log10((double)kFastKappa)))
                                     : (float)(log10(a) - log10((double)beta) - kLog10StartF64);
            // one scale for the whole alignment: accept sums that stay far
            // above double's floor (score above ~ -572), where no cell that
            // matters can have been flushed; the rest go to the
            // exponent-tracking wave kernel
            if (!(a >= 0x1p-900 && a < 1e300)) {
                const int n = atomicAdd(p.final_count, 1);
                p.final_items[n] = it;
            }
        }
    }
    if constexpr (REC) {
        // the group's lanes above the result lane store their magnitudes for
        // the range-extended pass
        const int res_t = res_lane();
        const int gs = __shfl(slot, lane - t + (res_t < G ? res_t : G - 1), 64);
        if (gs >= 0 && active && t < res_t && t < kReKnownLanes) {
            const int e = acc > 0.0f ? ilogbf((float)acc) : -128;
            p.rescue_exps[(int64_t)gs * kReKnownLanes + t] =
                (uint8_t)(int8_t)(e < -128 ? -128 : e > 127 ? 127 : e);
        }
    }
    if constexpr (DS)
        return __popcll(__ballot(ds_done));
    return 0;
}

// A task of 64/G consecutive alignments of an item list (the FP64 recompute
// of rescue lists): every group sets up its own read and sweeps its own hap.
template <typename T, int NR, bool FAST, bool ROWEPS, bool RE = false, bool DS = false,
          bool GE = false>
__device__ __forceinline__ int tile_items_task(const FwdParams& p, typename TileTraits<T>::V* tab,
                                                const WorkItem* items, int begin, int count, int G,
                                                int task, int rescue_base, int bin)
{
    const BatchDesc& bd = *p.bd;
    const int lane = threadIdx.x;
    const int W = 64 / G;
    const int grp = lane / G;
    const int t = lane - grp * G;
    const int local = task * W + grp;
    const bool active = grp < W && local < count;
    WorkItem it;
    it.read = 0;
    it.hap = 0;
    it.out = -1;
    it.pad = 0;
    if (active) {
        it = items[begin + local];
        if (RE)
            it.pad = begin + local;  // the rescue-list slot: its maxima in rescue_exps
    }
    TileRead<T, NR, ROWEPS> R;
    tile_setup<T, NR, FAST, ROWEPS>(p, tab, active, it.read, t, R);
    const int HL = active ? bd.hap_len[it.hap] : 0;
    const int hoff = active ? bd.hap_off[it.hap] : 0;
    return tile_sweep<T, NR, FAST, ROWEPS, RE, DS, GE>(p, tab, R, active, G, t, HL, hoff, it,
                                                       rescue_base, bin);
}

// A super-task: the groups set up their reads once and sweep the segment's
// haps in turn (all groups share each hap, so every sweep is uniform across
// the wave).  Empty haps were scored -inf by the scatter; haps holding a
// byte other than A/C/G/T belong to the exact bin.
template <typename T, int NR, bool FAST, bool ROWEPS, bool DS = false, bool GE = false>
__device__ __forceinline__ int tile_super_task(const FwdParams& p, typename TileTraits<T>::V* tab,
                                               const SuperTask& st, int G, int rescue_base, int bin)
{
    const BatchDesc& bd = *p.bd;
    const int lane = threadIdx.x;
    const int W = 64 / G;
    const int grp = lane / G;
    const int t = lane - grp * G;
    const bool active = grp < W && grp < st.nc;
    const int rd = active ? p.read_perm[st.perm + grp] : 0;
    TileRead<T, NR, ROWEPS> R;
    tile_setup<T, NR, FAST, ROWEPS>(p, tab, active, rd, t, R);
    int done = 0;
    for (int h = st.hs; h < st.he; ++h) {
        const int hap = st.h0 + h;
        const int HL = bd.hap_len[hap];
        if (HL <= 0 || p.hapflag[hap] != 0)
            continue;
        WorkItem it;
        it.read = rd;
        it.hap = hap;
        it.out = st.aoff + (rd - st.r0) * st.nh + h;
        it.pad = 0;
        done += tile_sweep<T, NR, FAST, ROWEPS, false, DS, GE>(p, tab, R, active, G, t, HL,
                                                              bd.hap_off[hap], it, rescue_base,
                                                              bin);
    }
    return done;
}

// Lane i < bs.n of a wave describes bin i of a BinSet: its id, lanes per
// alignment, begin, count and number of tasks (`per_task` units per task;
// 0: one super-task per task in super-task bins, 64 / lanes items per task
// in the others); incl is the inclusive prefix of tasks over the lanes.
// Returns the total number of tasks.
struct SetLane {
    int bin, G, begin, count, ntask, incl;
};
__device__ __forceinline__ int set_lanes(const BinSet& bs, int per_task, SetLane& s)
{
    const int lane = threadIdx.x;
    s.bin = 0;
    s.G = 1;
    s.begin = 0;
    s.count = 0;
    s.ntask = 0;
    if (lane < bs.n) {
        s.bin = bs.bin[lane];
        s.G = bs.lanes[lane] > 0 ? bs.lanes[lane] : 1;
        s.count = bs.counts[(int64_t)bs.count_stride * s.bin];
        s.begin = bs.range ? bs.range[2 * s.bin] : 0;
        const int per = per_task > 0                            ? per_task
                        : (bs.st && s.bin < bs.item_bin0) ? 1
                                                          : 64 / s.G;
        s.ntask = (s.count + per - 1) / per;
    }
    int x = s.ntask;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o)
            x += y;
    }
    s.incl = x;
    return __shfl(x, 63, 64);
}

// The lane (bin) whose task range holds `task` (task < total).
__device__ __forceinline__ int set_find(const SetLane& s, int task)
{
    const unsigned long long m = __ballot(s.incl > task);
    return __ffsll((long long)m) - 1;
}

// Next task of a wave from the launch's counter (one atomic per task).
__device__ __forceinline__ int next_task(int32_t* ctr)
{
    int t = 0;
    if (threadIdx.x == 0)
        t = atomicAdd(ctr, 1);
    return __shfl(t, 0, 64);
}

// Bins of one template in one persistent launch.  Waves take tasks in order
// from a counter: consecutive tasks of a length-sorted bin have similar
// lengths, and a wave that finishes early takes the next one.
// Waves per SIMD a tile kernel is built for: the 12-row tiles and the
// 13-row FAST tile fit three (at most 168 VGPRs, no spills) and issue more
// of their dependent FMA chains per cycle that way; the larger tiles need
// up to 256 VGPRs and run two.
// The range-extended instances (a smaller share of the time, and a few more
// live values in setup) are not held to three.
// The scaled instances are held to two where the register file allows no
// more (the 20-row SAFE tile would otherwise take AGPRs and run one).
template <typename T, int NR, bool FAST, bool RE, bool DS>
constexpr int tile_waves_per_simd()
{
    return (sizeof(T) == 4 && !RE && (NR == 12 || (NR == 13 && FAST))) ? 3 : DS ? 2 : 1;
}

// The tasks of one launch, taken in order from the launch's counter.
// Returns the alignments the scaling finished below the threshold (DS).
template <typename T, int NR, bool FAST, bool ROWEPS, bool RE, bool DS, bool GE>
__device__ __forceinline__ int tile_set_tasks(const FwdParams& p, const BinSet& bs,
                                              typename TileTraits<T>::V* tab, const SetLane& sl,
                                              int total)
{
    int done = 0;
    for (int task = next_task(bs.next_task); task < total; task = next_task(bs.next_task)) {
        // the task's bin, in scalar registers (uniform across the wave)
        const int k = set_find(sl, task);
        const int bin = __builtin_amdgcn_readfirstlane(__shfl(sl.bin, k, 64));
        const int G = __builtin_amdgcn_readfirstlane(__shfl(sl.G, k, 64));
        const int begin = __builtin_amdgcn_readfirstlane(__shfl(sl.begin, k, 64));
        const int count = __builtin_amdgcn_readfirstlane(__shfl(sl.count, k, 64));
        const int t0 = __builtin_amdgcn_readfirstlane(__shfl(sl.incl - sl.ntask, k, 64));
        // an item bin's results below the threshold join its tile bin's list
        const int rbin = bin >= bs.item_bin0 ? bin - bs.item_bin0 : bin;
        const int rescue_base =
            __builtin_amdgcn_readfirstlane(bs.seg ? bs.seg[2 * rbin] : 0);
        if constexpr (!RE) {
            if (bs.st && bin < bs.item_bin0) {
                done += tile_super_task<T, NR, FAST, ROWEPS, DS, GE>(
                    p, tab, bs.st[begin + (task - t0)], G, rescue_base, rbin);
                continue;
            }
        }
        done += tile_items_task<T, NR, FAST, ROWEPS, RE, DS, GE>(p, tab, bs.items, begin, count,
                                                                 G, task - t0, rescue_base, rbin);
    }
    return done;
}

template <typename T, int NR, bool FAST, bool ROWEPS = false, bool RE = false, bool DS = false>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(tile_waves_per_simd<T, NR, FAST, RE, DS>(), 8))) void k_fwd_tile_set(
    FwdParams p, BinSet bs)
{
    static_assert(NR >= 4 && NR <= 32, "rows per lane");
    __shared__ typename TileTraits<T>::V tab[TabShape<T, NR>::kWords];
    SetLane sl;
    const int total = set_lanes(bs, 0, sl);
    if (total == 0)  // empty bins (common in a persistent launch): no atomic traffic
        return;
    // DS: alignments finished by the scaling below the threshold.  The
    // scaled instances (the default FP32 pass) carry a copy of the sweep
    // for the reference GPU semantics, whose eps is one constant.
    int done;
    if (DS && !ROWEPS && p.semantics == 0)
        done = tile_set_tasks<T, NR, FAST, ROWEPS, RE, DS, DS && !ROWEPS>(p, bs, tab, sl, total);
    else
        done = tile_set_tasks<T, NR, FAST, ROWEPS, RE, DS, false>(p, bs, tab, sl, total);
    if constexpr (DS)
        if (done > 0 && threadIdx.x == 0)
            atomicAdd(p.scaled_total, (unsigned long long)done);
}

// ---------------------------------------------------------------------------
// k_fwd_wave
// ---------------------------------------------------------------------------
template <typename T>
struct WaveScalar;
template <>
struct WaveScalar<float> {
    __device__ static float shr(float v) { return shr1f(v); }
    __device__ static float ldexp(float v, int k) { return ldexpf(v, k); }
    __device__ static int ilogb(float v) { return ilogbf(v); }
};
template <>
struct WaveScalar<double> {
    __device__ static double shr(double v) { return shr1d(v); }
    __device__ static double ldexp(double v, int k) { return ::ldexp(v, k); }
    __device__ static int ilogb(double v) { return ::ilogb(v); }
};

// The items are those of the BinSet, one wave each (blocks take every
// gridDim-th item of the bins' concatenated item sequence).  `spill` holds
// spill_elems values of T; each resident block holds the border row's M, I,
// D per column and the power-of-two exponents they were stored with.
//
// Values carry a per-lane power-of-two exponent `sc` (true = stored * 2^sc).
// Before a lane takes in values from the lane above (or, for lane 0 of a
// lower tile, from the spilled border row) it compares their exponent, in
// its own scale, with its own largest value and rescales itself when the
// larger of the two leaves [2^-32, 2^96]; so nothing it receives can
// overflow or vanish, and the result is finite however low the score.
// T = float (multi-tile first pass) appends results below the rescue
// threshold (as the tile kernels define it) to the final list, which the
// T = double instance recomputes.
template <typename T, int NR>
__global__ __launch_bounds__(64) void k_fwd_wave(FwdParams p, BinSet bs, const int32_t* hmax_ptr,
                                                 T* spill, int64_t spill_elems,
                                                 int32_t* error_flag)
{
    using S = WaveScalar<T>;
    const BatchDesc& bd = *p.bd;
    constexpr int TR = 64 * NR;
    constexpr int kNone = -(1 << 28);
    const int lane = threadIdx.x;
    SetLane sl;
    const int n = set_lanes(bs, 1, sl);  // one item per wave
    const int hmax = *hmax_ptr;
    // per block: 3*(hmax+1) values of T, then 2*(hmax+1) int32 exponents,
    // in units of T (rounded up to 16 bytes)
    const int64_t bytes = 3 * (int64_t)(hmax + 1) * (int64_t)sizeof(T) + 8 * (int64_t)(hmax + 1);
    const int64_t stride = ((bytes + 15) / 16) * 16 / (int64_t)sizeof(T);
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
    // per column c: sp[3c+0] = M, sp[3c+1] = I, sp[3c+2] = D of the border
    // row; ex[2c] / ex[2c+1] the exponents M/D and I were stored with
    int32_t* ex = reinterpret_cast<int32_t*>(sp + 3 * (int64_t)(hmax + 1));

    for (int w = blockIdx.x; w < n; w += nblk) {
        const int kb = set_find(sl, w);
        const int begin = __shfl(sl.begin, kb, 64);
        const int w0 = __shfl(sl.incl - sl.ntask, kb, 64);
        const WorkItem it = bs.items[begin + (w - w0)];
        const int L = bd.read_len[it.read];
        const int HL = bd.hap_len[it.hap];
        const int roff = bd.read_off[it.read];
        const int hoff = bd.hap_off[it.hap];
        if (L <= 0 || HL <= 0) {
            if (lane == 0)
                bd.scores[it.out] = -INFINITY;
            continue;
        }
        if (HL > hmax) {
            if (lane == 0) {
                atomicOr(error_flag, 2);
                bd.scores[it.out] = __int_as_float(0x7fc00000);
            }
            continue;
        }
        // gap continuation: fixed (reference GPU semantics), the read's first
        // gcp quality (reference CPU semantics) or every row's own
        // (per-position gcp, set per row below)
        T eps = (T)kEpsGpu, beta = (T)kBetaGpu;
        if (p.semantics == 1) {
            const float e = ph2pr_of(p.ph2pr, bd.gcp[roff]);
            eps = (T)e;
            beta = (T)1 - (T)e;
        }
        const bool row_gcp = p.semantics == 2;
        const T C0 = S::ldexp((T)1, 64) / (T)HL;
        const int ntiles = (L + 1 + TR - 1) / TR;
        // Emission match: codes with 4 as the wildcard (reference GPU
        // semantics), or the raw bytes with a literal 'N' as the wildcard
        // (reference CPU semantics, gpuworker.cuh:1063) -- exact for any byte.
        const bool raw = p.semantics != 0;
        const int wild = raw ? (int)'N' : 4;
        T acc = 0;
        int sc = 0;
        for (int k = 0; k < ntiles; ++k) {
            const int r0 = k * TR + lane * NR;
            T ca[NR], cb[NR], cdl[NR], cs[NR], ce[NR], lm[NR], lx[NR];
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
                    if (row_gcp) {
                        const T e = (T)ph2pr_of(p.ph2pr, bd.gcp[roff + ri]);
                        ce[j] = e;
                        cb[j] = (T)1 - e;
                    } else {
                        ce[j] = eps;
                        cb[j] = beta;
                    }
                    lm[j] = (T)1 - pb;
                    code[j] = raw ? (int)bd.read_data[roff + ri] : row_code(w32);
                    lx[j] = code[j] == wild ? lm[j] : pb / (T)3;
                } else {
                    ca[j] = 0;
                    cb[j] = 0;
                    cdl[j] = ri == L ? (T)1 : (T)0;
                    cs[j] = 0;
                    ce[j] = 1;
                    lm[j] = 0;
                    lx[j] = 0;
                    code[j] = wild;
                }
            }
            T M[NR], D[NR];
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                M[j] = 0;
                D[j] = 0;
            }
            const bool top = (k == 0);
            sc = 0;
            acc = 0;
            T C0s = C0;
            T Ml = 0, Dl = (lane == 0 && top) ? C0 : (T)0, Mul = 0, Dul = Dl, Iin = 0;
            const bool last_tile = (k == ntiles - 1);
            const int tr = last_tile ? (L - k * TR) / NR : -1;
            const int nsteps = HL + 64;
            auto rescale = [&](int kk) {
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    M[j] = S::ldexp(M[j], kk);
                    D[j] = S::ldexp(D[j], kk);
                }
                Ml = S::ldexp(Ml, kk);
                Dl = S::ldexp(Dl, kk);
                Mul = S::ldexp(Mul, kk);
                Dul = S::ldexp(Dul, kk);
                Iin = S::ldexp(Iin, kk);
                acc = S::ldexp(acc, kk);
                C0s = S::ldexp(C0s, kk);
                sc -= kk;
            };
            // make room for incoming values of exponent e_in (own scale)
            auto fit = [&](int e_in) {
                T my = fmax(fmax(fmax(Ml, Dl), fmax(Mul, Dul)), fmax(Iin, acc));
#pragma unroll
                for (int j = 0; j < NR; ++j)
                    my = fmax(my, fmax(M[j], D[j]));
                const int e_my = my > (T)0 ? S::ilogb(my) : kNone;
                const int e = e_in > e_my ? e_in : e_my;
                if (e > kNone && (e > 96 || e < -32))
                    rescale(64 - e);
            };
            auto expo = [&](T v, int e) { return v > (T)0 ? S::ilogb(v) + e : kNone; };
            for (int s = 0; s < nsteps; ++s) {
                const int c = s - lane;  // this lane's 0-based column
                const int letter = (c >= 0 && c < HL)
                                       ? (raw ? (int)bd.hap_data[hoff + c] : (int)p.hapc[hoff + c])
                                       : wild;
                T mp = Mul, dp = Dul, chain = Iin;
                if (lane == 0 && !top) {
                    // row above the tile at column c-1, from the previous tile
                    // (only columns the previous tile wrote: c-1 < HL)
                    if (c >= 1 && c - 1 < HL) {
                        const int emd = ex[2 * (c - 1)] - sc, ei = ex[2 * (c - 1) + 1] - sc;
                        const T rm = sp[3 * (c - 1) + 0], ri = sp[3 * (c - 1) + 1];
                        const T rd = sp[3 * (c - 1) + 2];
                        int e_in = expo(rm, emd);
                        const int e2 = expo(rd, emd), e3 = expo(ri, ei);
                        e_in = e2 > e_in ? e2 : e_in;
                        e_in = e3 > e_in ? e3 : e_in;
                        fit(e_in);
                        const int emd2 = ex[2 * (c - 1)] - sc, ei2 = ex[2 * (c - 1) + 1] - sc;
                        mp = S::ldexp(rm, emd2);
                        chain = S::ldexp(ri, ei2);
                        dp = S::ldexp(rd, emd2);
                    } else {
                        mp = chain = dp = 0;
                    }
                }
#pragma unroll
                for (int j = 0; j < NR; ++j) {
                    const T mo = M[j], dov = D[j];
                    const T lam = (letter == code[j] || letter == wild) ? lm[j] : lx[j];
                    M[j] = lam * (ca[j] * mp + cb[j] * (chain + dp));
                    chain = cdl[j] * mp + ce[j] * chain;
                    // past the read (ce = 1, cs = 0) D stays at its start, 0
                    D[j] = cs[j] * mo + ce[j] * dov;
                    mp = mo;
                    dp = dov;
                }
                if (lane == tr && c >= 1 && c <= HL)
                    acc += chain;
                if (!last_tile && lane == 63) {
                    if (c >= 0 && c < HL) {
                        sp[3 * c + 0] = M[NR - 1];
                        sp[3 * c + 2] = D[NR - 1];
                        ex[2 * c] = sc;
                    }
                    if (c >= 1 && c <= HL) {
                        sp[3 * (c - 1) + 1] = chain;
                        ex[2 * (c - 1) + 1] = sc;
                    }
                }
                const T sM = S::shr(M[NR - 1]);
                const T sD = S::shr(D[NR - 1]);
                const T sI = S::shr(chain);
                const int ssc = shr1i(sc);
                if (lane != 0) {
                    int e_in = expo(sM, ssc - sc);
                    const int e2 = expo(sD, ssc - sc), e3 = expo(sI, ssc - sc);
                    e_in = e2 > e_in ? e2 : e_in;
                    e_in = e3 > e_in ? e3 : e_in;
                    fit(e_in);
                } else {
                    fit(kNone);
                }
                const int d = ssc - sc;
                Mul = Ml;
                Dul = Dl;
                Ml = lane == 0 ? (T)0 : S::ldexp(sM, d);
                Dl = lane == 0 ? (top ? C0s : (T)0) : S::ldexp(sD, d);
                Iin = lane == 0 ? (T)0 : S::ldexp(sI, d);
            }
            __syncthreads();  // spilled row visible to the next tile
        }
        // the result lane of the last tile owns acc
        const int tr = (L - (ntiles - 1) * TR) / NR;
        if (lane == tr) {
            const double lg = log10((double)acc) + (double)sc * 0.30102999566398119521 -
                              log10((double)C0 * (double)HL);
            bd.scores[it.out] = (float)lg;
            // same criterion as the tile kernels: the sum in the reference's
            // units (start FLT_MAX/16/H) below rescue_min, or not finite
            const double lmin = log10((double)p.rescue_min) - (double)kLog10Scale;
            if (sizeof(T) == 4 && p.rescue_on && !(acc > (T)0 && lg >= lmin && lg < 1e30)) {
                const int q = atomicAdd(p.final_count, 1);
                p.final_items[q] = it;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename K>
static int resident_blocks_per_cu(K kernel)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, 64, 0) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    return n > 0 ? n : 4;
}

template <typename T, int NR, bool FAST, bool ROWEPS = false, bool RE = false, bool DS = false>
static int set_blocks_per_cu()
{
    static const int n = resident_blocks_per_cu(k_fwd_tile_set<T, NR, FAST, ROWEPS, RE, DS>);
    return n;
}

static int grid_of(int grid, int per_cu)
{
    if (grid > 0)
        return grid;
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        else {
            (void)hipGetLastError();
            cus = 256;
        }
    }
    return cus * per_cu;
}

template <typename T, int NR, bool ROWEPS = false, bool RE = false, bool DS = false>
static hipError_t launch_set_nr(bool fast, const FwdParams& p, const BinSet& bs, int grid,
                                hipStream_t st)
{
    if (bs.n < 1 || bs.n > 64 || !bs.counts || !bs.next_task || (!bs.items && !bs.st))
        return hipErrorInvalidValue;
    if (RE && (bs.st || !bs.seg || !p.retry_items || !p.retry_counts || !p.rescue_exps))
        return hipErrorInvalidValue;
    if (DS && (!p.scaled_total || !p.rescue_on))
        return hipErrorInvalidValue;
    for (int i = 0; i < bs.n; ++i)
        if (bs.lanes[i] < 1 || bs.lanes[i] > 64)
            return hipErrorInvalidValue;
    if (fast) {
        hipLaunchKernelGGL((k_fwd_tile_set<T, NR, true, ROWEPS, RE, DS>),
                           dim3(grid_of(grid, set_blocks_per_cu<T, NR, true, ROWEPS, RE, DS>())),
                           dim3(64), 0, st, p, bs);
    } else {
        hipLaunchKernelGGL((k_fwd_tile_set<T, NR, false, ROWEPS, RE, DS>),
                           dim3(grid_of(grid, set_blocks_per_cu<T, NR, false, ROWEPS, RE, DS>())),
                           dim3(64), 0, st, p, bs);
    }
    return hipGetLastError();
}

bool tile_re_supported(int NR) { return re_rows(NR); }

hipError_t launch_fwd_tile_set_re(int NR, bool fast, const FwdParams& p, const BinSet& bs,
                                  int grid, hipStream_t st)
{
    switch (NR) {
    case 12: return launch_set_nr<float, 12, false, true>(fast, p, bs, grid, st);
    case 13: return launch_set_nr<float, 13, false, true>(fast, p, bs, grid, st);
    case 16: return launch_set_nr<float, 16, false, true>(fast, p, bs, grid, st);
    case 19: return launch_set_nr<float, 19, false, true>(fast, p, bs, grid, st);
    case 20: return launch_set_nr<float, 20, false, true>(fast, p, bs, grid, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fwd_tile_set_scaled(int NR, bool fast, const FwdParams& p, const BinSet& bs,
                                      int grid, hipStream_t st)
{
    switch (NR) {
    case 12: return launch_set_nr<float, 12, false, false, true>(fast, p, bs, grid, st);
    case 13: return launch_set_nr<float, 13, false, false, true>(fast, p, bs, grid, st);
    case 16: return launch_set_nr<float, 16, false, false, true>(fast, p, bs, grid, st);
    case 19: return launch_set_nr<float, 19, false, false, true>(fast, p, bs, grid, st);
    case 20: return launch_set_nr<float, 20, false, false, true>(fast, p, bs, grid, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fwd_tile_set(int NR, bool fast, const FwdParams& p, const BinSet& bs, int grid,
                               hipStream_t st, bool roweps)
{
    if (roweps)  // per-position gcp: the 16-row tiles only
        return NR == kRowEpsRows32 ? launch_set_nr<float, kRowEpsRows32, true>(fast, p, bs, grid, st)
                                   : hipErrorInvalidValue;
    switch (NR) {
    case 4: return launch_set_nr<float, 4>(fast, p, bs, grid, st);
    case 8: return launch_set_nr<float, 8>(fast, p, bs, grid, st);
    case 12: return launch_set_nr<float, 12>(fast, p, bs, grid, st);
    case 13: return launch_set_nr<float, 13>(fast, p, bs, grid, st);
    case 16: return launch_set_nr<float, 16>(fast, p, bs, grid, st);
    case 19: return launch_set_nr<float, 19>(fast, p, bs, grid, st);
    case 20: return launch_set_nr<float, 20>(fast, p, bs, grid, st);
    case 24: return launch_set_nr<float, 24>(fast, p, bs, grid, st);
    case 28: return launch_set_nr<float, 28>(fast, p, bs, grid, st);
    case 32: return launch_set_nr<float, 32>(fast, p, bs, grid, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fwd_tile_set_f64(int NR, bool fast, const FwdParams& p, const BinSet& bs,
                                   int grid, hipStream_t st, bool roweps)
{
    if (roweps)  // per-position gcp: the 8-row FP64 tiles only
        return NR == kRowEpsRows64 ? launch_set_nr<double, kRowEpsRows64, true>(fast, p, bs, grid, st)
                                   : hipErrorInvalidValue;
    switch (NR) {
    case 4: return launch_set_nr<double, 4>(fast, p, bs, grid, st);
    case 6: return launch_set_nr<double, 6>(fast, p, bs, grid, st);
    case 8: return launch_set_nr<double, 8>(fast, p, bs, grid, st);
    case 10: return launch_set_nr<double, 10>(fast, p, bs, grid, st);
    default: return hipErrorInvalidValue;
    }
}

int tile_set_blocks_per_cu(int NR, bool fast, bool f64, bool roweps, bool scaled)
{
    if (scaled && !f64 && !roweps) {
        switch (NR) {
        case 12: return fast ? set_blocks_per_cu<float, 12, true, false, false, true>()
                             : set_blocks_per_cu<float, 12, false, false, false, true>();
        case 13: return fast ? set_blocks_per_cu<float, 13, true, false, false, true>()
                             : set_blocks_per_cu<float, 13, false, false, false, true>();
        case 16: return fast ? set_blocks_per_cu<float, 16, true, false, false, true>()
                             : set_blocks_per_cu<float, 16, false, false, false, true>();
        case 19: return fast ? set_blocks_per_cu<float, 19, true, false, false, true>()
                             : set_blocks_per_cu<float, 19, false, false, false, true>();
        case 20: return fast ? set_blocks_per_cu<float, 20, true, false, false, true>()
                             : set_blocks_per_cu<float, 20, false, false, false, true>();
        default: return 0;
        }
    }
    if (roweps) {
        if (!f64 && NR == kRowEpsRows32)
            return fast ? set_blocks_per_cu<float, kRowEpsRows32, true, true>()
                        : set_blocks_per_cu<float, kRowEpsRows32, false, true>();
        if (f64 && NR == kRowEpsRows64)
            return fast ? set_blocks_per_cu<double, kRowEpsRows64, true, true>()
                        : set_blocks_per_cu<double, kRowEpsRows64, false, true>();
        return 0;
    }
    if (f64) {
        switch (NR) {
        case 4: return fast ? set_blocks_per_cu<double, 4, true>() : set_blocks_per_cu<double, 4, false>();
        case 6: return fast ? set_blocks_per_cu<double, 6, true>() : set_blocks_per_cu<double, 6, false>();
        case 8: return fast ? set_blocks_per_cu<double, 8, true>() : set_blocks_per_cu<double, 8, false>();
        case 10: return fast ? set_blocks_per_cu<double, 10, true>() : set_blocks_per_cu<double, 10, false>();
        default: return 0;
        }
    }
    switch (NR) {
    case 4: return fast ? set_blocks_per_cu<float, 4, true>() : set_blocks_per_cu<float, 4, false>();
    case 8: return fast ? set_blocks_per_cu<float, 8, true>() : set_blocks_per_cu<float, 8, false>();
    case 12: return fast ? set_blocks_per_cu<float, 12, true>() : set_blocks_per_cu<float, 12, false>();
    case 13: return fast ? set_blocks_per_cu<float, 13, true>() : set_blocks_per_cu<float, 13, false>();
    case 16: return fast ? set_blocks_per_cu<float, 16, true>() : set_blocks_per_cu<float, 16, false>();
    case 19: return fast ? set_blocks_per_cu<float, 19, true>() : set_blocks_per_cu<float, 19, false>();
    case 20: return fast ? set_blocks_per_cu<float, 20, true>() : set_blocks_per_cu<float, 20, false>();
    case 24: return fast ? set_blocks_per_cu<float, 24, true>() : set_blocks_per_cu<float, 24, false>();
    case 28: return fast ? set_blocks_per_cu<float, 28, true>() : set_blocks_per_cu<float, 28, false>();
    case 32: return fast ? set_blocks_per_cu<float, 32, true>() : set_blocks_per_cu<float, 32, false>();
    default: return 0;
    }
}

hipError_t launch_fwd_wave_f32(const FwdParams& p, const BinSet& bs, const int32_t* hmax_ptr,
                               float* spill, int64_t spill_elems, int32_t* error_flag, int blocks,
                               hipStream_t st)
{
    if (bs.n < 1 || bs.n > 64 || !bs.counts || !bs.items || blocks <= 0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_fwd_wave<float, 16>), dim3(blocks), dim3(64), 0, st, p, bs, hmax_ptr,
                       spill, spill_elems, error_flag);
    return hipGetLastError();
}

hipError_t launch_fwd_wave_f64(const FwdParams& p, const BinSet& bs, const int32_t* hmax_ptr,
                               double* spill, int64_t spill_elems, int32_t* error_flag, int blocks,
                               hipStream_t st)
{
    if (bs.n < 1 || bs.n > 64 || !bs.counts || !bs.items || blocks <= 0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_fwd_wave<double, 8>), dim3(blocks), dim3(64), 0, st, p, bs, hmax_ptr,
                       spill, spill_elems, error_flag);
    return hipGetLastError();
}

}  // namespace phmm
