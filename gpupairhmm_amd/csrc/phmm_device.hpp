// Shared device-side definitions of the gfx950 PairHMM engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phmm {

// One alignment of the sorted work list: global read index, global hap
// index and the read-major output slot (gpuworker.cuh:1011).
struct __attribute__((aligned(16))) WorkItem {
    int32_t read;
    int32_t hap;
    int32_t out;
    int32_t pad;
};

// Work of a tile bin: up to 64/lanes reads of one record (consecutive in
// read_perm, all in the same bin) against a segment [hs, he) of the record's
// haps.  One wave sets the reads up once and sweeps the segment's haps one
// after another; segments hold at most a column budget (at least one hap),
// so tasks stay numerous enough to balance.
struct __attribute__((aligned(16))) SuperTask {
    int32_t perm;    // index in read_perm of the first read
    int32_t nc;      // reads in the task (groups of the wave in use)
    int32_t h0;      // first hap of the record (global index)
    int32_t nh;      // haps of the record
    int32_t hs, he;  // the segment: record haps hs .. he-1
    int32_t r0;      // first read of the record (global index)
    int32_t aoff;    // output index of the record's first alignment
};

// Scale of the forward variables: the reference starts row 0's deletion
// state at FLT_MAX/16 / H and returns log10(sum) - log10(FLT_MAX/16)
// (pairhmm_kernels.cuh:1203, 1673, 1760-1763).
constexpr float kScale = 2.1267646664908054e37f;  // FLT_MAX / 16 (exact)
constexpr float kLog10Scale = 37.3277194f;      // log10(FLT_MAX / 16)
constexpr double kScaleD = 2.1267646664908054e37;

// Headroom divisor of the rescaled kernel's start constant (see
// k_fwd_tile_set): every rescaled cell stays below FLT_MAX/16.
constexpr float kFastKappa = 2.0f;
constexpr float kLog10FastKappa = 0.30103000f;

// The FP64 recompute starts at 2^1000/H: scores down to about -600 stay in
// double's normal range.
constexpr double kStartF64 = 0x1p1000;
constexpr double kLog10StartF64 = 301.02999566398120;  // 1000 * log10(2)

constexpr int kMaxTiles = 32;
// Bins: 2 per tile class (FAST, SAFE), the multi-tile bin, and the exact
// bin (haps with a byte other than A/C/G/T; under reference CPU semantics
// also reads with a byte other than A/C/G/T/N).
constexpr int kMaxBins = 2 * kMaxTiles + 2;
// Work bins: the bins above, then one item bin per tile bin (the alignments
// of a record's reads left over after full chunks of 64/lanes reads).
constexpr int kMaxWorkBins = kMaxBins + 2 * kMaxTiles;
// Rows per lane with a compiled tile kernel: multiples of 4, plus 13 and 19
// (8 lanes of them hold reads of 103 and 151 bases, common sequencer read
// lengths, with at most one spare row).
__host__ __device__ inline bool tile_rows_supported(int r)
{
    return (r >= 4 && r <= 32 && r % 4 == 0) || r == 13 || r == 19;
}
constexpr int kTaskCounters = 128;  // set launches per call (plain mode: one per bin)
constexpr int kF64Sets = 16;        // of them, the last ones: FP64 tile sets
constexpr int kReSets = 24;         // before those: range-extended FP32 sets
constexpr int kRetrySlots = 128;    // retry counts per bin
constexpr int kHapBuckets = 3072;   // hap-length buckets of the sort key
constexpr int kKeyParts = 8;        // counters per (bin, bucket): groups spread over them
constexpr int kKeysPerBin = kHapBuckets * kKeyParts;
constexpr int kAlignParts = 256;    // partial counters per bin (atomic spreading)
constexpr int kHmaxParts = 512;     // partial maxima of the hap pass (atomic spreading)

// Hap length -> sort bucket: exact below 2048, then width 16.
__host__ __device__ inline int hap_bucket(int h)
{
    if (h < 2048)
        return h < 0 ? 0 : h;
    int b = 2048 + ((h - 2048) >> 4);
    return b < kHapBuckets - 1 ? b : kHapBuckets - 1;
}

// Gap-continuation constants of the reference GPU kernel
// (pairhmm_kernels.cuh:1195-1196).
constexpr float kEpsGpu = 0.1f;
constexpr float kBetaGpu = 0.9f;

// Packed read row: base code | base q << 8 | ins q << 16 | del q << 24,
// qualities clamped to 127.
__host__ __device__ inline int row_code(uint32_t w) { return (int)(w & 0xffu); }
__host__ __device__ inline int row_bq(uint32_t w) { return (int)((w >> 8) & 0xffu); }
__host__ __device__ inline int row_iq(uint32_t w) { return (int)((w >> 16) & 0xffu); }
__host__ __device__ inline int row_dq(uint32_t w) { return (int)(w >> 24); }

// A/C/G/T -> 0..3, every other byte -> 4 (utility_kernels.cuh:16-20).
__host__ __device__ inline int base_code(uint8_t b)
{
    return b == 'A' ? 0 : b == 'C' ? 1 : b == 'G' ? 2 : b == 'T' ? 3 : 4;
}

// Reference CPU semantics (align_host, gpuworker.cuh:1063), and per-position
// gcp semantics, which extend them: only a literal 'N' is the wildcard; any
// byte other than A/C/G/T/N gets code 5 and sends its read or hap to the
// exact bin, whose kernel compares raw bytes.
constexpr int kExoticCode = 5;
__host__ __device__ inline int base_code_sem(uint8_t b, int semantics)
{
    const int c = base_code(b);
    return (semantics != 0 && c == 4 && b != 'N') ? kExoticCode : c;
}

// Per-read metadata from the read pass.
struct __attribute__((aligned(8))) ReadMeta {
    int32_t bin;     // work-list bin; -1: empty read
    float nl;        // product of alpha over the read (rescaled kernel)
};

// The per-call part of a batch: device pointers of a self-contained batch
// view, the output array and the sizes.  It lives in device memory (one
// small copy per call), so every kernel launch of a call -- and a captured
// graph of them -- is independent of the batch.
struct BatchDesc {
    const uint8_t* read_data;
    const uint8_t* hap_data;
    const uint8_t* bq;
    const uint8_t* iq;
    const uint8_t* dq;
    const uint8_t* gcp;
    const int32_t* read_len;
    const int32_t* hap_len;
    const int32_t* read_off;
    const int32_t* hap_off;
    const int32_t* nreads;     // per group
    const int32_t* nhaps;      // per group
    const int32_t* rprefix;    // [G+1]
    const int32_t* hprefix;    // [G+1]
    float* scores;
    int32_t G, nR, nH;
    int32_t col_budget;  // hap columns per super-task segment
    int64_t read_bytes, hap_bytes, num_alignments;
};

// Everything the forward kernels read besides the batch descriptor.
struct FwdParams {
    const BatchDesc* bd;
    const uint32_t* rows;        // packed read rows, indexed like read bytes
    const uint8_t* hapc;         // hap base codes, indexed like hap bytes
    const ReadMeta* meta;
    const int32_t* read_perm;    // reads of each record ordered by bin
    const uint8_t* hapflag;      // 1: the hap holds a byte other than A/C/G/T (exact bin)
    const float* ph2pr;          // 128 entries
    WorkItem* rescue_items;      // per tile bin: alignments to recompute in FP64
    int32_t* rescue_counts;      // per bin
    WorkItem* final_items;       // alignments the FP64 tile pass could not finish
    int32_t* final_count;
    // Range-extended FP32 pass (re_on): the FP32 pass records, for the first
    // kReKnownLanes lanes of a rescued alignment, ilogb of each lane's last
    // row's magnitude (rescue_exps: 16 bytes per rescue-list slot, the
    // slot's index as in rescue_items; lanes past them count as unknown); the pass recomputes the rescue
    // lists with per-lane power-of-two scales and sends what it cannot
    // accept to the retry lists (per tile bin, like the rescue lists), which
    // the FP64 tile pass recomputes.
    WorkItem* retry_items;
    int32_t* retry_counts;
    uint8_t* rescue_exps;
    int32_t re_on;
    int32_t pad0;
    // In-pass scaling (the scaled instances of the FP32 tiles): alignments
    // whose sum fell below rescue_min and that the scaling finished are
    // counted here (one atomic per wave and launch).
    unsigned long long* scaled_total;
    float rescue_min;            // FP32 sum (reference units) threshold
    int32_t rescue_on;
    int32_t semantics;           // 0: reference GPU, 1: reference CPU, 2: per-position gcp
    int32_t pad;
};

// Range-extended FP32 pass.  Lane t of a group scales its values by 2^E_t
// (E_0 = 0) and multiplies what it takes from lane t-1 by 2^(E_t - E_{t-1});
// the start constant is the FP32 pass's times 2^-kReShift.  E_t is chosen
// from the FP32 pass's per-lane magnitudes r_k (ilogb): E_t = ilogb(start) -
// r_{t-1} while r_0 .. r_{t-1} are all known (>= kReValid), beyond that the
// last known per-lane decay continues.  A result is accepted when every
// lane's incoming magnitude lies within [start/2^kReLow, start*2^kReHigh]
// (in its own scale), the lane-to-lane factors fit a float and the sum is
// finite and at least 2^kReMinSum.  The window trades two risks: a lane
// scaled too high overflows (detected: the sum or a magnitude is not
// finite), a lane scaled too low flushes values that a path with a better
// suffix would have needed (not detectable locally).  kReShift leaves 30
// bits above the first pass's own headroom for extrapolation errors;
// kReLow keeps every lane's values within 2^70 of the first pass's start,
// i.e. at least about 2^150 above the FP32 floor.
constexpr int kReShift = 30;
constexpr int kReValid = -120;
constexpr int kReLow = 40;
constexpr int kReHigh = 50;
constexpr int kReMinSum = -100;
constexpr int kReUnknownSlope = -150;  // per-lane decay when lane 0's maximum is unknown
constexpr int kReKnownLanes = 16;      // lanes per alignment with a recorded magnitude
// Rows per lane with a range-extended instance (the tuned tables' rows).
__host__ __device__ constexpr bool re_rows(int NR)
{
    return NR == 12 || NR == 13 || NR == 16 || NR == 19 || NR == 20;
}

// In-pass scaling (the scaled FP32 tile instances; the same rows as the
// range-extended pass).  Lane t of a group holds its values times 2^E_t
// (E_t >= 0; E_0 = 0) and multiplies what it takes from lane t-1 by
// 2^(E_t - E_{t-1}).  E_t starts from the insertion floor: every cell of a
// row r >= 1 holds at least the path that enters row 1 anywhere and
// inserts down to r, start * prod of the rows' chain coefficients, which
// telescopes (FAST: eps^r * delta_0 / delta_r / prod alpha; SAFE: eps^r),
// so lane t starts with its last row's floor at 2^kDsFloor (E_t =
// kDsFloor - log2 of that floor, never below 0: lanes whose floor stays
// above it are not scaled at all) and nothing it will hold can be below its
// range; no lane starts or stays more than kDsSpan above the lane over it.  Every kDsEvery steps a scaled lane whose incoming
// values passed 2^kDsGuard scales down (never below E = 0) to put them at
// 2^kDsGuardTo: a cell grows by the gap-open factor (2^13 at insertion
// quality 40) in the step where a read's alignment reaches it and by the
// gap-extension factor (2^3.3 at gcp 10) in the steps before, so over
// kDsEvery steps about 2^63, which with the FAST form's D (about 2^13 above
// its M) just fits above 2^kDsGuard (a growth past it overflows, which
// fails acceptance); what the scale-down flushes lies at least 2^149 below
// the lane's largest value.  Scaling by
// powers of two is exact, so the sums are those of FP32 with an unbounded
// exponent wherever no value leaves the normal range; lanes that never
// leave E = 0 compute exactly what the unscaled instance computes.  A
// result is finished in the pass when every lane up to the result lane
// holds a running sum of at least 2^kReMinSum (finite where the lane is
// scaled; an unscaled lane's sum may overflow harmlessly, as in the
// unscaled instance), the result lane's sum is finite and neighbouring
// exponents keep the transfer factor a normal float; the rest join the
// rescue list.
constexpr int kDsEvery = 16;
constexpr int kDsFloor = -100;
constexpr int kDsSpan = 127;  // largest E_t - E_{t-1}
constexpr int kDsGuard = 55;
constexpr float kDsGuardF = 0x1p55f;  // 2^kDsGuard
constexpr int kDsGuardTo = 0;

// Everything the binning kernels read and write (binning.hip).
struct BinParams {
    const BatchDesc* bd;
    int32_t NB;                // bins of reads: 2*ntiles + 2
    int32_t NW;                // work bins: NB + 2*ntiles (item bin of tile bin k: NB + k)
    // scratch
    uint32_t* rows;
    uint8_t* hapc;
    ReadMeta* meta;
    int32_t* gb_count;         // [G*(NB+1)]  slot NB: empty reads
    int32_t* gb_start;         // [G*(NB+1)]
    int32_t* read_perm;        // [nR]
    int64_t* align_off;        // [G+1]
    int32_t* key_count;        // [NW*kKeysPerBin]
    int32_t* key_cursor;       // [NW*kKeysPerBin]
    int32_t* bin_range;        // [2*NW]: begin, count (tile bins: super-tasks; others: items)
    int32_t* hmax;             // [1]: the longest hap (k_scan_bins folds hmax_part)
    int32_t* hmax_part;        // [kHmaxParts]: per-block maxima of the hap pass
    int32_t* rescue_counts;    // [NB]
    int32_t* retry_counts;     // [kRetrySlots]: retry lists' counts per bin
    int32_t* final_count;      // [1]
    int32_t* error_flag;       // [1]
    int32_t* task_ctr;         // [kTaskCounters]: task counters of the set launches
    SuperTask* st;             // tile bins' work, in bin_range units
    int32_t* align_range;      // [2*NB]: per bin, begin / count in alignments
    int32_t* align_part;       // [NB*kAlignParts]: partial alignment counts per bin (k_scan_keys sums each bin's into its first)
    uint8_t* hapflag;          // [nH]: 1 = hap holds a byte other than A/C/G/T
    WorkItem* items;
    const float* ph2pr;
    // tiles
    int32_t ntiles;
    int32_t tiles[kMaxTiles];  // ascending max read length per tile class
    int32_t lanes[kMaxTiles];  // lanes per alignment of each tile class
    int32_t semantics;
    int32_t allow_fast;
};

// Binning launches.  `hd` is the host copy of the descriptor bp.bd points to;
// with fixed_grids every grid is independent of the batch sizes (graph
// capture; `small`: the smaller grids of the small-batch sequence), else
// grids are sized to them.
hipError_t run_binning(const BinParams& bp, const BatchDesc& hd, bool fixed_grids, bool small,
                       hipStream_t st);
// *total += sum of counts[0..n)
hipError_t launch_add_count(int64_t* total, const int32_t* counts, int n, hipStream_t st);
// Several such sums in one launch.
struct CountSums {
    static constexpr int kMax = 4;
    int32_t n = 0;
    int64_t* total[kMax] = {};
    const int32_t* counts[kMax] = {};
    int32_t len[kMax] = {};
    void add(int64_t* t, const int32_t* c, int l)
    {
        total[n] = t;
        counts[n] = c;
        len[n] = l;
        ++n;
    }
};
hipError_t launch_add_counts(const CountSums& c, hipStream_t st);
// Several bins in one persistent launch.  Bin bin[i] holds count_i =
// counts[count_stride*bin[i]] units starting at begin_i = range ?
// range[2*bin[i]] : 0, both read on the device.  A unit is a SuperTask
// (st != null and bin < item_bin0: tile kernels, one task each) or a
// WorkItem of `items` (tile kernels: 64/lanes per task; wave kernels: one
// each).  The bins' tasks form one sequence in the given order; tile-kernel
// waves take the next task from *next_task (zeroed before the launch).
// Rescue lists: results of bin b (an item bin: of its tile bin b -
// item_bin0) below the threshold go to tile bin's list, which starts at
// seg[2*bin] in FwdParams::rescue_items.
struct BinSet {
    int32_t n = 0;
    int32_t item_bin0 = 1 << 30;   // first item bin id (tile kernels with st)
    int32_t bin[kMaxBins] = {};
    int32_t lanes[kMaxBins] = {};  // tile kernels: lanes per alignment
    const WorkItem* items = nullptr;
    const SuperTask* st = nullptr;
    const int32_t* range = nullptr;
    const int32_t* counts = nullptr;
    int32_t count_stride = 0;
    const int32_t* seg = nullptr;
    int32_t* next_task = nullptr;
};
// Tile kernels with a gap continuation per row (per-position gcp
// semantics) exist for these rows per lane only.
constexpr int kRowEpsRows32 = 16;
constexpr int kRowEpsRows64 = 8;
// grid: 0 = as many blocks as fit on the device at once; roweps: the
// per-position gcp instances (kRowEpsRows32 / kRowEpsRows64 rows only)
hipError_t launch_fwd_tile_set(int NR, bool fast, const FwdParams& p, const BinSet& bs, int grid,
                               hipStream_t st, bool roweps = false);
hipError_t launch_fwd_tile_set_f64(int NR, bool fast, const FwdParams& p, const BinSet& bs,
                                   int grid, hipStream_t st, bool roweps = false);
// The range-extended FP32 pass over rescue lists (item bins of a BinSet
// whose seg gives each bin's retry-list base); rows per lane with such an
// instance: tile_re_supported.
bool tile_re_supported(int NR);
hipError_t launch_fwd_tile_set_re(int NR, bool fast, const FwdParams& p, const BinSet& bs,
                                  int grid, hipStream_t st);
// The FP32 tile sets with in-pass scaling (rows per lane as for
// tile_re_supported; not under per-position gcp).
hipError_t launch_fwd_tile_set_scaled(int NR, bool fast, const FwdParams& p, const BinSet& bs,
                                      int grid, hipStream_t st);
hipError_t launch_fwd_wave_f32(const FwdParams& p, const BinSet& bs, const int32_t* hmax_ptr,
                               float* spill, int64_t spill_elems, int32_t* error_flag, int blocks,
                               hipStream_t st);
hipError_t launch_fwd_wave_f64(const FwdParams& p, const BinSet& bs, const int32_t* hmax_ptr,
                               double* spill, int64_t spill_elems, int32_t* error_flag, int blocks,
                               hipStream_t st);
// Resident blocks of 64 threads per CU for a tile-set instance.
int tile_set_blocks_per_cu(int NR, bool fast, bool f64, bool roweps = false, bool scaled = false);

}  // namespace phmm
