/*
 * CPU oracle for the PairHMM forward algorithm -- TEST INFRASTRUCTURE ONLY
 * (see pairhmm_oracle.h).  Restates align_host (gpuworker.cuh:1022-1098)
 * and processBatchCPU (gpuworker.cuh:975-1020).
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  Contraction
 * must stay off: align_host's float cells are separate multiplies and adds.
 */
#include "pairhmm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <omp.h>

static int omp_default_threads(void) { return omp_get_max_threads(); }

void oracle_ph2pr(float out[128])
{
    /* gpuworker.cuh:479-483: tmp = pow(10.0, (-i) / 10.0) stored as float */
    for (int i = 0; i < 128; ++i)
        out[i] = (float)pow(10.0, (double)(-i) / 10.0);
}

static int base_code(uint8_t b)
{
    /* utility_kernels.cuh:16-20 */
    switch (b) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return 4;
    }
}

static int is_match(uint8_t x, uint8_t h, int semantics)
{
    if (semantics != ORACLE_SEM_GPU) /* gpuworker.cuh:1063 */
        return x == h || x == 'N' || h == 'N';
    /* GPU: both sides encoded, code 4 matches everything */
    int cx = base_code(x), ch = base_code(h);
    return cx == ch || cx == 4 || ch == 4;
}

static inline int qclamp(uint8_t q) { return q > 127 ? 127 : q; }

/* A double with its own exponent epoch: value = m * 2^(512*e), with m == 0
 * or 2^-512 <= m < 2^512.  The DP values of one row can span far more than
 * double's range (a decaying D carry-over next to a path that later
 * dominates, at scores below about -300), so every cell carries its epoch;
 * neighbouring cells almost always share it, which keeps the arithmetic
 * plain double. */
typedef struct {
    double m;
    int e;
} xd;

static inline xd xd_norm(xd x)
{
    /* magnitudes: qualities of 0 make alpha negative and cells signed (as
     * in the reference); non-finite values are left alone */
    if (!isfinite(x.m))
        return x;
    while (fabs(x.m) >= 0x1p512) {
        x.m *= 0x1p-512;
        x.e += 1;
    }
    while (x.m != 0.0 && fabs(x.m) < 0x1p-512) {
        x.m *= 0x1p512;
        x.e -= 1;
    }
    return x;
}

static inline xd xd_scale(xd x, double k)
{
    x.m *= k;
    return x;  /* k <= 1 and >= ~1e-13: renormalized by the caller */
}

static inline xd xd_add(xd a, xd b)
{
    if (a.m == 0.0)
        return b;
    if (b.m == 0.0)
        return a;
    if (a.e < b.e) {
        const xd t = a;
        a = b;
        b = t;
    }
    const int d = a.e - b.e;
    if (d >= 3)  /* b < 2^-512 of a */
        return a;
    a.m += ldexp(b.m, -512 * d);
    return a;
}

/* Gap continuation eps and gap-to-match beta of row r (1-based):
 *   SEM_GPU: fixed 0.1f / 0.9f (pairhmm_kernels.cuh:1195-1196);
 *   SEM_CPU: from the read's first gcp quality for every row
 *            (gpuworker.cuh:1009, 1048-1049);
 *   SEM_GCP: from the row's own gcp quality gq[r-1], indexed like the row's
 *            insertion and deletion qualities (gpuworker.cuh:1059-1061). */
static inline uint8_t gcp_of_row(const uint8_t* gq, int r, int semantics)
{
    return semantics == ORACLE_SEM_GCP ? gq[r - 1] : gq[0];
}

static double align_f64(const uint8_t* hap, int H, const uint8_t* read, int R,
                        const uint8_t* bq, const uint8_t* iq, const uint8_t* dq,
                        const uint8_t* gq, int semantics, const float* P)
{
    const double constant = (double)FLT_MAX / 16.0;
    double eps = (double)0.1f, beta = (double)0.9f;
    xd* buf = (xd*)malloc(sizeof(xd) * 6 * (size_t)(H + 1));
    xd *Mp = buf, *Ip = buf + (H + 1), *Dp = buf + 2 * (H + 1);
    xd *Mc = buf + 3 * (H + 1), *Ic = buf + 4 * (H + 1), *Dc = buf + 5 * (H + 1);
    const xd zero = {0.0, 0};
    for (int c = 0; c <= H; ++c) {
        Mp[c] = zero;
        Ip[c] = zero;
        Dp[c] = xd_norm((xd){constant / H, 0});
    }
    for (int r = 1; r <= R; ++r) {
        const double p = P[qclamp(bq[r - 1])];
        const double delta = P[qclamp(iq[r - 1])];
        const double sigma = P[qclamp(dq[r - 1])];
        const double alpha = 1.0 - (delta + sigma);
        if (semantics != ORACLE_SEM_GPU) {
            eps = (double)P[qclamp(gcp_of_row(gq, r, semantics))];
            beta = 1.0 - eps;
        }
        Mc[0] = Ic[0] = Dc[0] = zero;
        for (int c = 1; c <= H; ++c) {
            const double lambda = is_match(read[r - 1], hap[c - 1], semantics) ? 1.0 - p : p / 3.0;
            const xd gap = xd_add(Ip[c - 1], Dp[c - 1]);
            Mc[c] = xd_norm(xd_scale(xd_add(xd_scale(Mp[c - 1], alpha), xd_scale(gap, beta)), lambda));
            Ic[c] = xd_norm(xd_add(xd_scale(Mp[c], delta), xd_scale(Ip[c], eps)));
            Dc[c] = xd_norm(xd_add(xd_scale(Mc[c - 1], sigma), xd_scale(Dc[c - 1], eps)));
        }
        xd* t;
        t = Mp; Mp = Mc; Mc = t;
        t = Ip; Ip = Ic; Ic = t;
        t = Dp; Dp = Dc; Dc = t;
    }
    xd sum = zero;
    for (int c = 1; c <= H; ++c)
        sum = xd_norm(xd_add(sum, xd_norm(xd_add(Mp[c], Ip[c]))));
    free(buf);
    if (sum.m == 0.0)
        return -INFINITY;
    return log10(sum.m) + (double)sum.e * 512.0 * 0.30102999566398119521 - log10(constant);
}

/* Float cells, double result: the rounding of align_host, step for step. */
static double align_f32cells(const uint8_t* hap, int H, const uint8_t* read, int R,
                             const uint8_t* bq, const uint8_t* iq, const uint8_t* dq,
                             const uint8_t* gq, int semantics, const float* P)
{
    double result = 0;
    const double constant = (double)FLT_MAX / 16;
    float* buf = (float*)malloc(sizeof(float) * 6 * (size_t)(H + 1));
    float *M = buf, *I = buf + 2 * (H + 1), *D = buf + 4 * (H + 1);
    int target_row = 0, source_row = 0;
    const int W = H + 1;
    for (int col = 0; col <= H; col++) {
        M[col] = I[col] = 0;
        D[col] = (float)(constant / H);
    }
    float beta = (float)0.9, eps = (float)0.1;
    for (int row = 1; row <= R; row++) {
        if (semantics != ORACLE_SEM_GPU) {
            const uint8_t g = gcp_of_row(gq, row, semantics);
            beta = (float)(1.0 - P[qclamp(g)]);
            eps = P[qclamp(g)];
        }
        target_row = row & 1;
        source_row = !target_row;
        M[target_row * W] = I[target_row * W] = 0;
        D[target_row * W] = 0;
        const uint8_t read_base = read[row - 1];
        const float base_qual = P[qclamp(bq[row - 1])];
        const float alpha = (float)(1.0 - (double)(P[qclamp(iq[row - 1])] + P[qclamp(dq[row - 1])]));
        const float delta = P[qclamp(iq[row - 1])];
        const float sigma = P[qclamp(dq[row - 1])];
        for (int col = 1; col <= H; col++) {
            const float lambda = is_match(read_base, hap[col - 1], semantics)
                                     ? (float)(1.0 - (double)base_qual)
                                     : (float)((double)base_qual / 3.0);
            const float s1 = I[source_row * W + col - 1] + D[source_row * W + col - 1];
            const float s2 = beta * s1;
            const float s3 = alpha * M[source_row * W + col - 1];
            M[target_row * W + col] = lambda * (s3 + s2);
            const float i1 = delta * M[source_row * W + col];
            const float i2 = eps * I[source_row * W + col];
            I[target_row * W + col] = i1 + i2;
            const float d1 = sigma * M[target_row * W + col - 1];
            const float d2 = eps * D[target_row * W + col - 1];
            D[target_row * W + col] = d1 + d2;
        }
    }
    for (int col = 1; col <= H; col++) {
        const float mi = M[target_row * W + col] + I[target_row * W + col];
        result += mi;
    }
    free(buf);
    result = log10(result) - log10(constant);
    return result;
}

double oracle_align_gcp(const uint8_t* hap, int hap_len, const uint8_t* read, int read_len,
                        const uint8_t* base_q, const uint8_t* ins_q, const uint8_t* del_q,
                        const uint8_t* gcp_q, int precision, int semantics)
{
    float P[128];
    oracle_ph2pr(P);
    if (hap_len <= 0 || read_len <= 0)
        return -INFINITY;
    if (precision == ORACLE_F32CELLS)
        return align_f32cells(hap, hap_len, read, read_len, base_q, ins_q, del_q, gcp_q, semantics,
                              P);
    return align_f64(hap, hap_len, read, read_len, base_q, ins_q, del_q, gcp_q, semantics, P);
}

double oracle_align(const uint8_t* hap, int hap_len, const uint8_t* read, int read_len,
                    const uint8_t* base_q, const uint8_t* ins_q, const uint8_t* del_q,
                    uint8_t gcp0, int precision, int semantics)
{
    if (semantics == ORACLE_SEM_GCP)  /* needs every row's gcp */
        return NAN;
    return oracle_align_gcp(hap, hap_len, read, read_len, base_q, ins_q, del_q, &gcp0, precision,
                            semantics);
}

/* Map a global read-major output index to (group, read, hap) exactly as
 * processBatchCPU does (gpuworker.cuh:988-1011): alignments of group g occupy
 * [A_g, A_g + R_g*H_g) and index a = A_g + r*H_g + h. */
static double score_one(int64_t a, int G, const int* nH,
                        const int* rp, const int* hp, const int64_t* aprefix,
                        const uint8_t* read_data, const uint8_t* hap_data,
                        const uint8_t* bq, const uint8_t* iq, const uint8_t* dq,
                        const uint8_t* gq, const int* rl, const int* hl,
                        const int* ro, const int* ho, int precision, int semantics,
                        const float* P)
{
    int lo = 0, hi = G; /* first g with aprefix[g+1] > a */
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (aprefix[mid + 1] > a)
            hi = mid;
        else
            lo = mid + 1;
    }
    const int g = lo;
    const int64_t local = a - aprefix[g];
    const int h = (int)(local % nH[g]);
    const int r = (int)(local / nH[g]);
    const int read = rp[g] + r, hap = hp[g] + h;
    const int L = rl[read], HL = hl[hap];
    if (L <= 0 || HL <= 0)
        return -INFINITY;
    const int roff = ro[read], hoff = ho[hap];
    if (precision == ORACLE_F32CELLS)
        return align_f32cells(hap_data + hoff, HL, read_data + roff, L, bq + roff, iq + roff,
                              dq + roff, gq + roff, semantics, P);
    return align_f64(hap_data + hoff, HL, read_data + roff, L, bq + roff, iq + roff, dq + roff,
                     gq + roff, semantics, P);
}

static int64_t* alignment_prefix(int G, const int* nR, const int* nH)
{
    int64_t* ap = (int64_t*)malloc(sizeof(int64_t) * (size_t)(G + 1));
    ap[0] = 0;
    for (int g = 0; g < G; ++g)
        ap[g + 1] = ap[g] + (int64_t)nR[g] * nH[g];
    return ap;
}

int oracle_score_subset(int G, const int* nR, const int* nH, const int* rp, const int* hp,
                        const uint8_t* read_data, const uint8_t* hap_data, const uint8_t* bq,
                        const uint8_t* iq, const uint8_t* dq, const uint8_t* gq, const int* rl,
                        const int* hl, const int* ro, const int* ho, const int64_t* which,
                        int64_t n_which, int precision, int semantics, int nthreads, double* out)
{
    if (G < 0 || n_which < 0)
        return 1;
    float P[128];
    oracle_ph2pr(P);
    int64_t* ap = alignment_prefix(G, nR, nH);
    for (int64_t i = 0; i < n_which; ++i)
        if (which[i] < 0 || which[i] >= ap[G]) {
            free(ap);
            return 2;
        }
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : omp_default_threads())
    for (int64_t i = 0; i < n_which; ++i)
        out[i] = score_one(which[i], G, nH, rp, hp, ap, read_data, hap_data, bq, iq, dq, gq,
                           rl, hl, ro, ho, precision, semantics, P);
    free(ap);
    return 0;
}

int oracle_score_batch(int G, const int* nR, const int* nH, const int* rp, const int* hp,
                       const uint8_t* read_data, const uint8_t* hap_data, const uint8_t* bq,
                       const uint8_t* iq, const uint8_t* dq, const uint8_t* gq, const int* rl,
                       const int* hl, const int* ro, const int* ho, int precision, int semantics,
                       int nthreads, double* out)
{
    if (G < 0)
        return 1;
    float P[128];
    oracle_ph2pr(P);
    int64_t* ap = alignment_prefix(G, nR, nH);
    const int64_t total = ap[G];
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : omp_default_threads())
    for (int64_t a = 0; a < total; ++a)
        out[a] = score_one(a, G, nH, rp, hp, ap, read_data, hap_data, bq, iq, dq, gq, rl, hl,
                           ro, ho, precision, semantics, P);
    free(ap);
    return 0;
}
