/*
 * CPU oracle for the PairHMM forward algorithm -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, never the product: only tests/, the smoke
 * check and the CPU-baseline leg of bench.py may load it.  The engine
 * (libgpupairhmm.so) never links or calls it.
 *
 * It restates the reference's CPU forward algorithm `align_host`
 * (gpuworker.cuh:1022-1098) and its batch driver `processBatchCPU`
 * (gpuworker.cuh:975-1020) in plain C, in two arithmetic flavours:
 *   - ORACLE_F64      : every quantity in double, each cell with its own
 *                       power-of-two exponent epoch (the FP64 oracle; exact
 *                       at any score, whatever range a row spans);
 *   - ORACLE_F32CELLS : float M/I/D cells and coefficients, double result
 *                       sum -- the exact rounding of align_host.
 * and two input semantics:
 *   - ORACLE_SEM_GPU  : what the reference's GPU kernel computes: gap
 *                       continuation fixed at eps=0.1f, beta=0.9f
 *                       (pairhmm_kernels.cuh:1195-1196) and every byte other
 *                       than A/C/G/T is the N wildcard (utility_kernels.cuh:16-20);
 *   - ORACLE_SEM_CPU  : what align_host computes: eps/beta from the read's
 *                       first gcp quality (gpuworker.cuh:1009,1048-1049) and
 *                       only a literal 'N' is a wildcard (gpuworker.cuh:1063);
 *   - ORACLE_SEM_GCP  : align_host with every row's own gap continuation:
 *                       row r (1-based) takes eps_r = ph2pr[gcp[r-1]] and
 *                       beta_r = 1 - eps_r where align_host takes eps and
 *                       beta, indexed like the row's insertion and deletion
 *                       qualities (gpuworker.cuh:1059-1061); literal 'N'
 *                       wildcard.  Equal to ORACLE_SEM_CPU, bit for bit, when
 *                       a read's gcp qualities are all equal.
 * Both semantics agree for uppercase ACGTN input with gcp '+'.
 *
 * Pinning (see tests/test_oracle.py): the README known-answer record
 * (README.md:45-61), closed forms for R=1 and H=1, brute-force path
 * enumeration on tiny inputs, scale invariance, and -- where
 * /root/reference is present -- bit-exact agreement of ORACLE_F32CELLS with
 * the reference's own align_host compiled from its source (oracle/Makefile).
 *
 * Quality bytes index a 128-entry table; bytes above 127 (undefined in the
 * reference, which reads past its table) are clamped to 127 here and in the
 * engine.
 */
#ifndef PAIRHMM_ORACLE_H
#define PAIRHMM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_F64 = 0, ORACLE_F32CELLS = 1 };
enum { ORACLE_SEM_GPU = 0, ORACLE_SEM_CPU = 1, ORACLE_SEM_GCP = 2 };

/* ph2pr[q] = (float)10^(-q/10), q = 0..127  (gpuworker.cuh:473-486). */
void oracle_ph2pr(float out[128]);

/* One alignment.  Qualities already have 33 subtracted.  Returns the log10
 * score log10(sum_c M[R][c]+I[R][c]) - log10(FLT_MAX/16); -inf when either
 * length is 0. */
double oracle_align(const uint8_t* hap, int hap_len,
                    const uint8_t* read, int read_len,
                    const uint8_t* base_q, const uint8_t* ins_q,
                    const uint8_t* del_q, uint8_t gcp0,
                    int precision, int semantics);

/* The same with the read's gcp qualities (gcp_q[0..read_len-1]); the only
 * form that serves ORACLE_SEM_GCP (oracle_align returns NaN for it). */
double oracle_align_gcp(const uint8_t* hap, int hap_len,
                        const uint8_t* read, int read_len,
                        const uint8_t* base_q, const uint8_t* ins_q,
                        const uint8_t* del_q, const uint8_t* gcp_q,
                        int precision, int semantics);

/* A whole batch in the reference's SoA layout (common.cuh:109-201), scored
 * like processBatchCPU: out[off_g + r*H_g + h].  Returns 0 on success.
 * nthreads <= 0 uses every OpenMP thread. */
int oracle_score_batch(int num_groups,
                       const int* num_reads_per_group,
                       const int* num_haps_per_group,
                       const int* reads_prefix,   /* [G+1] */
                       const int* haps_prefix,    /* [G+1] */
                       const uint8_t* read_data,
                       const uint8_t* hap_data,
                       const uint8_t* base_q, const uint8_t* ins_q,
                       const uint8_t* del_q, const uint8_t* gcp_q,
                       const int* read_lengths, const int* hap_lengths,
                       const int* read_offsets, const int* hap_offsets,
                       int precision, int semantics, int nthreads,
                       double* out);

/* Same as oracle_score_batch but only for the alignment indices listed in
 * `which` (global read-major output indices); out[i] receives the score of
 * alignment which[i]. */
int oracle_score_subset(int num_groups,
                        const int* num_reads_per_group,
                        const int* num_haps_per_group,
                        const int* reads_prefix, const int* haps_prefix,
                        const uint8_t* read_data, const uint8_t* hap_data,
                        const uint8_t* base_q, const uint8_t* ins_q,
                        const uint8_t* del_q, const uint8_t* gcp_q,
                        const int* read_lengths, const int* hap_lengths,
                        const int* read_offsets, const int* hap_offsets,
                        const int64_t* which, int64_t n_which,
                        int precision, int semantics, int nthreads,
                        double* out);

#ifdef __cplusplus
}
#endif
#endif
