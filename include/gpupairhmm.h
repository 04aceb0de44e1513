/*
 * gpupairhmm -- C ABI of the MI355X (gfx950) PairHMM forward-algorithm engine.
 *
 * Every entry point takes plain pointers and sizes and returns an int status
 * (PHMM_OK = 0); no C++ exception crosses this boundary.  The text of the
 * last error of an engine is available from phmm_engine_last_error(), and of
 * calls that have no engine from phmm_last_error() (thread-local).
 *
 * Reference interfaces replaced (asbschmidt/gpuPairHMM):
 *   - the batch layout: BatchOfSequenceGroups (common.cuh:109-201)
 *   - the batch engine: GpuPairHMMWorker::processBatch_overlapped_float_
 *     coalesced_smem (gpuworker.cuh:559-973)
 *   - the device-pointer kernel dispatch: GpuPairHMMI /
 *     MultiConfigGpuPairHMM::launchPairHMMSingleTileKernel /
 *     launchPairHMMMultiTileKernel (gpuworker.cuh:101-125, 220-375)
 *   - the line-record parser and score writer: FileParserWorker
 *     (file_parser_worker.cu:31-218), OutputWriterWorker
 *     (output_writer_worker.cu:29-159)
 *   - the options: Options (common.cuh:16-24), main.cu:47-83
 */
#ifndef GPUPAIRHMM_H
#define GPUPAIRHMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHMM_ABI_VERSION 1

/* ---- status codes --------------------------------------------------- */
enum {
    PHMM_OK = 0,
    PHMM_EINVAL = 1,   /* malformed arguments or batch */
    PHMM_ENOMEM = 2,   /* device or host allocation failed */
    PHMM_EDEVICE = 3,  /* a HIP call failed */
    PHMM_EIO = 4,      /* file could not be opened / read / written */
    PHMM_EPARSE = 5,   /* malformed input record */
    PHMM_EBUSY = 6,    /* serving: request refused by the HBM budget */
    PHMM_EAGAIN = 7,   /* serving: no result ready within the wait */
    PHMM_ERANGE = 8    /* a size exceeds a fixed limit */
};

/* ---- options (common.cuh:16-24; main.cu:47-83) ---------------------- */
enum {
    PHMM_MODE_PLAIN = 0,   /* per-bin launches sized on the host (one host
                              round trip per transfer chunk), like the
                              reference's per-partition launches */
    PHMM_MODE_BATCHED = 1  /* binning, persistent forward kernels and FP64
                              rescue stay on the device: no host round trip */
};
enum {
    PHMM_SEM_GPU = 0,  /* reference GPU kernel semantics: gap continuation
                          fixed at eps=0.1, beta=0.9 (pairhmm_kernels.cuh:
                          1195-1196); every non-ACGT byte is the N wildcard
                          (utility_kernels.cuh:16-20) */
    PHMM_SEM_CPU = 1,  /* reference align_host semantics: eps/beta from the
                          read's first gcp quality (gpuworker.cuh:1009,
                          1048-1049); only a literal 'N' is a wildcard
                          (gpuworker.cuh:1063) */
    PHMM_SEM_GCP = 2   /* per-position gap continuation: align_host with
                          row r (1-based) taking eps_r = ph2pr[gcp[r-1]] and
                          beta_r = 1 - eps_r, indexed like the row's
                          insertion/deletion qualities (gpuworker.cuh:
                          1059-1061); literal-'N' wildcard.  Equals
                          PHMM_SEM_CPU when a read's gcp values are all
                          equal.  Device views need gcp_quals 4-byte
                          aligned. */
};

typedef struct phmm_options {
    /* Alignments per transfer chunk; a chunk closes at the first record
     * where the running count reaches this value.  <= 0: unlimited.
     * (--transferBatchsize, common.cuh:19, gpuworker.cuh:659-669) */
    int64_t transfer_batchsize;
    int32_t mode;              /* PHMM_MODE_* */
    int32_t semantics;         /* PHMM_SEM_* */
    int32_t fp64_rescue;       /* 1: recompute low/invalid FP32 results in FP64 */
    float rescue_min_sum;      /* FP32 likelihood sum (the reference's scaled
                                  units) below which a result is recomputed
                                  in FP64; default FLT_MIN, the floor of
                                  the reference's FP32 normal range */
    int32_t verbose;
    int32_t time_kernels;      /* 1: event-time the forward kernels (stats) */
    int32_t graphs;            /* PHMM_MODE_BATCHED: 1 (default) replays the
                                  call's kernel sequence as one captured HIP
                                  graph; 0 launches the kernels one by one */
    int32_t small_batch;       /* batches of at most this many alignments
                                  (either mode) use a one-template tile table
                                  (16 rows per lane for every class; under
                                  range_rescue 2 they rescue in FP64
                                  directly), so their latency is one launch
                                  per stage rather than one per tile
                                  template.  Scores are bit-identical across
                                  tables well above the rescue threshold;
                                  near and below it they agree to FP32
                                  accuracy (within 5e-5 log10).  0: always
                                  the tuned table.  Default 1 << 18. */
    int32_t f64_rows;          /* rows per lane of every FP64 rescue tile
                                  (4, 6, 8 or 10; SAFE tiles at most 8); 0
                                  (default): chosen per bin from the
                                  measured FP64 tile rates */
    int32_t range_rescue;      /* 1 (default): the FP32 tile kernels keep
                                  a power-of-two scale per tile lane, picked
                                  during the sweep from the running sum of
                                  the lane above, so sums below
                                  rescue_min_sum are finished in the same
                                  pass; what the scaling cannot vouch for
                                  goes to FP64.  2: a second FP32 pass over
                                  the results below rescue_min_sum, with
                                  per-lane scales chosen from the first
                                  pass's per-lane magnitudes (tuned table
                                  only), then FP64.  0: every such result is
                                  recomputed in FP64.  Tiles of 12, 13, 16,
                                  19 and 20 rows per lane only; not under
                                  PHMM_SEM_GCP. */
    int32_t reserved[2];
} phmm_options;

/* ---- batch view (BatchOfSequenceGroups, common.cuh:109-201) --------- *
 * All sequences are padded as the reference's parser pads them
 * (file_parser_worker.cu:81-86): every read/hap occupies len + (4 - len%4)
 * bytes (1..4 padding bytes) and offsets are exclusive prefix sums of the
 * padded lengths.  Qualities have 33 subtracted.  Scores are written
 * read-major per group: out[A_g + r*H_g + h] with A_g the exclusive prefix
 * of R_g*H_g (gpuworker.cuh:1011, output_writer_worker.cu:110-127).
 * The same struct describes host memory (phmm_score_batch) or device memory
 * (phmm_score_device); counts are always host values. */
typedef struct phmm_batch_view {
    int32_t num_groups;
    int32_t num_reads;
    int32_t num_haps;
    int32_t reserved0;
    int64_t read_bytes;        /* == read_offsets[num_reads] */
    int64_t hap_bytes;         /* == hap_offsets[num_haps] */
    int64_t num_alignments;    /* sum of R_g*H_g */
    const uint8_t* read_data;
    const uint8_t* hap_data;
    const uint8_t* base_quals;
    const uint8_t* ins_quals;
    const uint8_t* del_quals;
    const uint8_t* gcp_quals;
    const int32_t* read_lengths;        /* [num_reads] */
    const int32_t* hap_lengths;         /* [num_haps] */
    const int32_t* read_offsets;        /* [num_reads + 1] */
    const int32_t* hap_offsets;         /* [num_haps + 1] */
    const int32_t* num_reads_per_group; /* [num_groups] */
    const int32_t* num_haps_per_group;  /* [num_groups] */
    const int32_t* reads_prefix;        /* [num_groups + 1] */
    const int32_t* haps_prefix;         /* [num_groups + 1] */
} phmm_batch_view;

typedef struct phmm_engine phmm_engine;
typedef struct phmm_host_batch phmm_host_batch;
typedef struct phmm_device_batch phmm_device_batch;
typedef struct phmm_server phmm_server;

/* Cumulative counters of an engine. */
typedef struct phmm_stats {
    int64_t alignments;        /* alignments scored */
    int64_t cells;             /* sum of readLen*hapLen (gpuworker.cuh:495-520) */
    int64_t rescued;           /* alignments recomputed after the FP32 pass
                                  (range-extended FP32 or FP64) */
    int64_t fast_alignments;   /* alignments on the rescaled 5-op kernel */
    int64_t multitile_alignments;
    double engine_seconds;     /* wall time inside phmm_score_batch */
    double kernel_seconds;     /* event-timed forward kernels (if timed) */
    int64_t graph_launches;    /* HIP graph replays (batched mode) */
    int64_t graph_captures;    /* graphs captured (first call, new scratch
                                  sizes or a new tile table) */
    double rescue_seconds;     /* event-timed rescue stage (if timed) */
    int64_t f64_alignments;    /* of the rescued alignments, those recomputed
                                  in FP64 */
    int64_t scaled_alignments; /* alignments whose FP32 sum fell below
                                  rescue_min_sum and that the FP32 pass
                                  finished with per-lane scaling
                                  (range_rescue 1); not in `rescued` */
    int64_t reserved[3];
} phmm_stats;

/* One row of the tile table (the reference's per-architecture
 * tuningconfigs.hpp:33-239, here chosen on gfx950 by the autotuner): reads
 * of length <= max_read_len (and above the previous row's) use `lanes`
 * lanes x `rows` rows per lane; lanes*rows must be >= max_read_len + 1. */
typedef struct phmm_tile_config {
    int32_t lanes;         /* lanes per alignment, 1..64 */
    int32_t rows;          /* rows per lane: 4, 8, 12, 13, 16, 19, 20, 24, 28 or 32;
                              only 12, 13, 16, 19 and 20 scale in the FP32
                              pass (the others hand what falls below the
                              FP32 floor to FP64), and 24-32 spill registers
                              on gfx950: the tuned table uses 12-20 */
    int32_t max_read_len;  /* ascending across the table */
} phmm_tile_config;

/* ---- library --------------------------------------------------------- */
int phmm_abi_version(void);
const char* phmm_last_error(void);
void phmm_options_default(phmm_options* opt);

/* ---- engine (GpuPairHMMWorker, gpuworker.cuh:381-471) ---------------- */
int phmm_engine_create(int device, const phmm_options* opt, phmm_engine** out);
void phmm_engine_destroy(phmm_engine* eng);
const char* phmm_engine_last_error(const phmm_engine* eng);
int phmm_engine_stats(const phmm_engine* eng, phmm_stats* out);
int phmm_engine_reset_stats(phmm_engine* eng);
/* Replace the tile table (max_read_len ascending); n <= 32.  Reads longer
 * than the last row's max_read_len use the multi-tile kernel. */
int phmm_engine_set_tiles(phmm_engine* eng, const phmm_tile_config* tiles, int n);
int phmm_engine_get_tiles(const phmm_engine* eng, phmm_tile_config* tiles, int cap, int* n);

/* Host batch in, host scores out; synchronous.  Chunks the batch by
 * transfer_batchsize at record granularity and overlaps host<->device copies
 * with compute.  (processBatch_overlapped_float_coalesced_smem,
 * gpuworker.cuh:559-973) */
int phmm_score_batch(phmm_engine* eng, const phmm_batch_view* batch, float* scores_out);

/* ---- in-process callers: one region's likelihood matrix --------------- *
 * What a variant caller's per-region PairHMM call computes (GATK's
 * PairHMM.computeLikelihoods, GKL's computeLikelihoodsNative): R reads, each
 * with base, insertion, deletion and gap-continuation qualities as phred
 * values (not ASCII), against H haplotypes; out[r*H + h] receives the log10
 * likelihood of read r given hap h, as phmm_score_batch scores the same
 * region as one record.  No counterpart in the reference, whose callers go
 * through files; this is the adapter such a caller binds (over JNI, ctypes,
 * ...) instead of writing records.  Synchronous. */
typedef struct phmm_region_read {
    const uint8_t* bases;
    const uint8_t* base_quals;
    const uint8_t* ins_quals;
    const uint8_t* del_quals;
    const uint8_t* gcp_quals;
    int32_t length;
    int32_t reserved;
} phmm_region_read;
typedef struct phmm_region_hap {
    const uint8_t* bases;
    int32_t length;
    int32_t reserved;
} phmm_region_hap;
int phmm_compute_likelihoods(phmm_engine* eng, const phmm_region_read* reads, int32_t num_reads,
                             const phmm_region_hap* haps, int32_t num_haps, double* out);

/* Device batch in (array pointers are device memory), device scores out,
 * ordered on `hip_stream` (a hipStream_t; NULL = the engine's stream).  In
 * PHMM_MODE_BATCHED the call only enqueues work; in PHMM_MODE_PLAIN it waits
 * once for the per-bin counts.  (GpuPairHMMI::launchPairHMMSingleTileKernel /
 * launchPairHMMMultiTileKernel, gpuworker.cuh:101-125, 220-375) */
int phmm_score_device(phmm_engine* eng, const phmm_batch_view* device_batch,
                      float* d_scores, void* hip_stream);
/* Allocation-free variant of phmm_score_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The call allocates nothing -- no hipMalloc, hipFree or
 * allocating library routine: the batch-sized scratch comes from a device
 * array of PHMM_NOALLOC_SCRATCH_BYTES that every GPU the library runs on
 * holds from engine creation until the process ends.  A batch whose scratch
 * does not fit (about 4e5 alignments and up) returns PHMM_ERANGE.  Same
 * kernels as phmm_score_device, launched one by one and never
 * kernel-timed: the scores are bit-identical.  Calls of one engine are
 * ordered; calls of different engines on one GPU share that GPU's array and
 * race if they overlap. */
#define PHMM_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int phmm_score_device_noalloc(phmm_engine* eng, const phmm_batch_view* device_batch,
                              float* d_scores, void* hip_stream);
/* Wait for all work the engine enqueued on its own stream. */
int phmm_engine_synchronize(phmm_engine* eng);

/* Copy a host batch to engine-owned device memory; the returned view holds
 * device pointers and stays valid until phmm_device_batch_free. */
int phmm_device_batch_upload(phmm_engine* eng, const phmm_batch_view* host,
                             phmm_device_batch** out);
int phmm_device_batch_view(const phmm_device_batch* db, phmm_batch_view* out);
void phmm_device_batch_free(phmm_device_batch* db);
/* Device scratch helpers for callers without their own allocator. */
int phmm_device_alloc(phmm_engine* eng, size_t bytes, void** out);
int phmm_device_free(phmm_engine* eng, void* ptr);
/* Page-locked host memory (hipHostMalloc) for callers that keep their
 * batches where the copy engines read them directly, as the reference keeps
 * its batches in pinned memory (PinnedAllocator, common.cuh:115-116):
 * phmm_score_batch copies such arrays to the device without staging them. */
int phmm_host_alloc_pinned(phmm_engine* eng, size_t bytes, void** out);
int phmm_host_free_pinned(phmm_engine* eng, void* ptr);
int phmm_memcpy_d2h(phmm_engine* eng, void* dst, const void* src, size_t bytes);
int phmm_memcpy_h2d(phmm_engine* eng, void* dst, const void* src, size_t bytes);

/* ---- host batches: parser, writer, synthetic workloads --------------- */
/* Parse line records (file_parser_worker.cu:88-218) from memory. */
int phmm_host_batch_parse(const char* text, size_t len, phmm_host_batch** out);
/* Parse a file in batches: records are appended until the running count of
 * alignments reaches file_batchsize (<= 0: whole file), then `cb` is called
 * with the batch; a non-zero return from cb stops parsing and is returned. */
typedef int (*phmm_batch_callback)(const phmm_batch_view* batch, void* user);
int phmm_parse_file(const char* path, int64_t file_batchsize, phmm_batch_callback cb, void* user);
int phmm_host_batch_view(const phmm_host_batch* b, phmm_batch_view* out);
void phmm_host_batch_free(phmm_host_batch* b);
/* Synthetic workload `config` (0..4) with `seed`; num_records < 0 keeps the
 * configuration's record count. */
int phmm_synth_generate(int config, uint64_t seed, int32_t num_records, phmm_host_batch** out);
/* Line-record text of a batch (the input format), caller frees with
 * phmm_free_text. */
int phmm_host_batch_to_text(const phmm_batch_view* batch, char** text, size_t* len);
/* Scores as the reference's writer prints them: per group "R H\n" then R
 * lines of H "%.5f" values (output_writer_worker.cu:110-127). */
int phmm_format_scores(const phmm_batch_view* batch, const float* scores, char** text, size_t* len);
void phmm_free_text(char* text);
/* Sum of readLen*hapLen over all alignments (gpuworker.cuh:495-520). */
int64_t phmm_batch_cells(const phmm_batch_view* batch);

/* ---- serving: continuous batching of active-region requests --------- *
 * Admitted requests stay in a device arena until their scores are returned;
 * a worker merges waiting requests on the device (up to the per-launch caps)
 * whenever one of `streams` launch contexts is free, so launches overlap on
 * the GPU.  No counterpart in the reference. */
typedef struct phmm_server_options {
    int64_t hbm_budget_bytes;  /* device bytes the server may hold: its launch
                                  contexts (engine scratch, merge buffers,
                                  allocated at creation at the caps below)
                                  plus the arena of admitted requests */
    int32_t max_inflight;      /* requests admitted but not yet returned */
    int32_t max_batch_requests;/* requests merged into one device launch */
    int64_t max_batch_alignments; /* alignments per launch; a larger request
                                  is refused with PHMM_ERANGE */
    int64_t max_batch_bytes;   /* batch-image bytes per launch (5 x read bytes
                                  + hap bytes + index arrays); a larger
                                  request is refused with PHMM_ERANGE */
    int32_t streams;           /* launch contexts (streams), 1..3; default 2 */
    int32_t reserved0;
    int64_t latency_batch_alignments; /* alignments per merged launch while
                                  the queue holds less than 4 x
                                  max_batch_alignments (a launch above the
                                  engine's small_batch runs the tuned table's
                                  longer sequence, and requests arriving
                                  behind it wait for it); beyond that backlog
                                  launches grow to max_batch_alignments for
                                  throughput.  <= 0: max_batch_alignments.
                                  Default 1 << 19. */
} phmm_server_options;

void phmm_server_options_default(phmm_server_options* opt);
int phmm_server_create(phmm_engine* eng, const phmm_server_options* opt, phmm_server** out);
void phmm_server_destroy(phmm_server* srv);
/* Submit one request (a host batch, usually one region = one group).  The
 * data is copied (the caller may reuse its buffers on return); PHMM_EBUSY
 * when the arena (the budget) or the in-flight limit refuses it -- it is
 * admitted again once earlier requests complete. */
int phmm_server_submit(phmm_server* srv, const phmm_batch_view* request, uint64_t tag);
/* Wait up to timeout_ms (< 0: forever) for a finished request; copies its
 * scores into `scores` (capacity floats) and its submit-to-completion time
 * into *latency_seconds (may be NULL).  PHMM_EAGAIN when none finished in
 * time; PHMM_ERANGE (result kept, *n_scores set) when capacity is short.  A
 * request whose launch failed returns that launch's status (its scores are
 * NaN); other requests are unaffected. */
int phmm_server_poll(phmm_server* srv, int timeout_ms, uint64_t* tag, float* scores,
                     int64_t capacity, int64_t* n_scores, double* latency_seconds);
typedef struct phmm_server_stats {
    int64_t submitted, refused, completed, launches;
    int64_t inflight, queued, max_queue_depth;
    int64_t bytes_in_use;      /* scratch_bytes + arena bytes of admitted requests */
    int64_t cells;
    int64_t scratch_bytes;     /* device bytes the launch contexts hold */
    int64_t arena_bytes;       /* the request arena: budget - scratch_bytes */
    int64_t failed;            /* requests whose launch failed */
    int64_t max_concurrent;    /* most launches on the device at once */
    int64_t reserved[4];
} phmm_server_stats;
int phmm_server_stats_get(const phmm_server* srv, phmm_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* GPUPAIRHMM_H */
