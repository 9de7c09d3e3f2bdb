// Internal interface between the C-ABI host layer (wfa_api.cpp) and the HIP
// kernels (wfa_kernels.hip).  Not installed; no torch types anywhere.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace wfa {

constexpr int32_t kOffsetNull = INT32_MIN / 2;  // WAVEFRONT_OFFSET_NULL

// Component indices (M, I1, D1, I2, D2), as affine2p_matrix_type order-free ids.
enum Comp : int { kM = 0, kI1 = 1, kD1 = 2, kI2 = 3, kD2 = 4 };

// Per-pair device status codes.  0 / -100 mirror WF_STATUS_ALG_COMPLETED /
// WF_STATUS_MAX_STEPS_REACHED; the others are internal and make the host
// re-run the pair with a larger configuration.
enum DevStatus : int {
  kDevCompleted = 0,
  kDevMaxSteps = -100,
  kDevRingOverflow = -2001,   // a wavefront wider than the launch's ring capacity
  kDevArenaFull = -2002,      // the backtrace-record arena ran out
  kDevBadTrace = -2003,       // backtrace found an inconsistent record (bug guard)
  // Final statuses of the heuristic cut-off instances:
  kDevUnreachable = -2004,    // more null steps in a row than the score scope
                              // (WF_STATUS_END_UNREACHABLE, wavefront_extend.c:94-101)
  kDevStaleTrace = -2005,     // the reference's backtrace would read a cell of a
                              // wavefront the cut-off emptied (not replayed)
  kDevZdropped = -2006,       // z-drop: partial alignment ending at the best cell seen
};

// Distance metrics, numbered as distance_metric_t (wavefront_penalties.h:41-47).
enum Metric : int { kIndel = 0, kEdit = 1, kLinear = 2, kAffine = 3, kAffine2p = 4 };

// Internal (Eizenga-transformed) penalties plus derived scalars.  The
// single-component metrics (indel, edit, gap-linear) keep only M: o1 is the
// cost of one gap symbol and e1 is 0, so an M cell's gap sources sit at
// s - o1 (indel: no mismatch source; edit: x = o1 = 1).
struct Penalties {
  int x, o1, e1, o2, e2;   // o2/e2 unused for gap-affine
  int scope;               // max_score_scope
  int g;                   // gcd of all score increments: only multiples of g can be non-null
  int nslots;              // ring slots = (scope-1)/g + 1
  int metric;              // Metric
  int ncomp;               // wavefront components: 1 (M), 3 (M, I1, D1) or 5
  // A gap wavefront is read only e/g steps after it is computed, so the
  // generic forward kernel keeps I1/D1 in rings of r1 = e1/g + 1 slots and
  // I2/D2 in rings of r2 = e2/g + 1 (0 without them); M keeps nslots.
  int r1, r2;
};

// Slots of every component's ring in the generic forward kernel.
__host__ __device__ inline int ring_slots(const Penalties& p) { return p.nslots + 2 * p.r1 + 2 * p.r2; }

// Edit distance prunes a wavefront of at least this many diagonals
// (wavefront_compute_edit.c:219-278).
constexpr int kEditPruneWidth = 1000;

// Record byte (one per diagonal of every non-null step):
//   bits 0..2  winning M source: 0 X, 1 D2, 2 D1, 3 I2, 4 I1
//   bits 3..6  gap-extension flags for I1, D1, I2, D2 at this diagonal
//   bit  7     (cut-off instances) a source the reference's backtrace probes
//              here lies in a wavefront the cut-off emptied, and changes a
//              choice or a value of this diagonal: a backtrace through it is
//              not replayed
// A step's block is a 16-byte header {s, lo, hi, previous block} followed by
// hi - lo + 1 record bytes, padded to 16 bytes.  The offsets themselves are
// not stored: the backtrace walks (score, diagonal, component) with these
// bits and rebuilds the match runs by extending forward, which reproduces the
// forward pass's extensions exactly.
constexpr uint32_t kRecSrcMask = 7u;
constexpr uint32_t kRecI1Ext = 1u << 3;
constexpr uint32_t kRecD1Ext = 1u << 4;
constexpr uint32_t kRecI2Ext = 1u << 5;
constexpr uint32_t kRecD2Ext = 1u << 6;
constexpr uint32_t kRecStale = 1u << 7;
__host__ __device__ inline unsigned long long rec_units_for(int width) {
  return 1 + (unsigned long long)((unsigned)(width + 15) / 16u);  // width >= 0
}
enum RecSrc : uint32_t { kSrcX = 0, kSrcD2 = 1, kSrcD1 = 2, kSrcI2 = 3, kSrcI1 = 4 };
constexpr uint32_t kNoBlock = 0xFFFFFFFFu;
constexpr uint32_t kChunkUnits = 4096;   // record chunk grabbed per workgroup, in 16-byte units

// Heuristic cut-off parameters (wavefront_heuristic.c:506-561), as
// wavefront_aligner_init_heuristic selects the strategies: the kernel
// instances with the cut-off run when any is set.
struct HeurArgs {
  int wfadaptive;                // 0 none, 1 wf-adaptive, 2 wfmash (weighted distances, :134-145)
  int min_len, thr, steps;       // and the countdown shared by every strategy
  int xdrop, xdrop_thr;          // x-drop (:294-370) on Smith-Waterman scores
  int zdrop, zdrop_thr;          // z-drop (:371-431), when x-drop is not set
  int swg_match;                 // (swg_match * (v + h) - s) / 2: -match, or -1 when match is 0
  int banded, band_lo, band_hi;  // banded: 0 none, 1 static, 2 adaptive (:451-500)
  // kernel instance: 0 exact, 1 wf-adaptive / wfmash / banded, 2 with an x-drop or z-drop
  __host__ __device__ int cut() const { return (xdrop || zdrop) ? 2 : ((wfadaptive || banded) ? 1 : 0); }
};

struct ForwardArgs {
  int n_pairs;
  const int* order;            // processing order (pair ids), longest first
  const int* plen;
  const int* tlen;
  const uint64_t* words;       // packed sequences
  const int64_t* pat_pos;      // first symbol of each pattern (index into the packed words)
  const int64_t* txt_pos;      // first symbol of each text
  const signed char* comp_begin_p;  // per-pair begin / end components (null: comp_begin / comp_end)
  const signed char* comp_end_p;
  int* out_status;
  int* out_score;              // wavefront score at the end (s_end)
  long long* out_dsteps;       // diagonal-steps of the pair (sum of hi-lo+1)
  uint32_t* out_last_block;    // last record block of the pair, 16-byte units
  uint4* rec_arena;            // record arena (16-byte units)
  unsigned long long rec_units;
  unsigned long long* rec_top; // bump pointer into the arena, 16-byte units
  int* ring_global;            // per-workgroup ring scratch (global-ring kernels)
  long long ring_stride;       // ints per workgroup
  unsigned int* queue;         // work-queue counter
  unsigned long long* dsteps_total;  // sum of diagonal-steps over the launch
  unsigned int* retry_count;   // pairs that must be re-run (ring overflow / arena full)
  unsigned int* fail_count;    // pairs that finished with another non-completed status
  int wcap;                    // ring capacity in diagonals per slot
  int lmax_words;              // LDS capacity per sequence, in 64-bit words
  int max_steps;               // INT_MAX if unlimited
  int records;                 // 0 = score only
  int comp_begin, comp_end;
  int defer_shift;             // HBM-ring family: defer a pair when the arena has < L^2 >> shift units left (0: never)
  // Ends-free form (alignment_endsfree): the score-0 wavefront spans
  // [-pbf, tbf] and the alignment ends at the first diagonal whose extended
  // cell reaches the text end with at most pef pattern symbols left, or the
  // pattern end with at most tef text symbols left.  0: end-to-end.
  int endsfree;
  int pbf, pef, tbf, tef;
  int* out_end_k;              // end cell of each completed pair: diagonal and offset
  int* out_end_off;
  int* out_end_rec;            // z-drop: the end cell's record byte (0xFF: no source), or null
  HeurArgs heur;                // heuristic cut-off (none: the exact instances)
  Penalties pen;
};

// Free-end value that covers any sequence (the extension form frees both ends entirely).
constexpr int kFreeAll = 0x3FFFFFFF;

struct BacktraceArgs {
  int n_pairs;
  const int* order;            // pairs to trace (the forward pass's subset)
  const int* plen;
  const int* tlen;
  int* status;                 // in/out (kDevBadTrace on a corrupt walk)
  const int* score;
  const uint32_t* last_block;
  const uint4* rec_arena;
  unsigned long long rec_units;
  char* ops;                   // operations arena
  const int64_t* ops_off;      // slot of pair i: ops[ops_off[i] .. +plen+tlen)
  int* ops_begin;              // CIGAR = ops[ops_off + ops_begin .. ops_off + plen + tlen)
  unsigned int* fail_count;    // incremented for a pair whose trace is inconsistent
  const signed char* comp_end_p;  // per-pair end component (null: comp_end)
  const signed char* comp_begin_p; // per-pair begin component (null: comp_begin)
  int comp_begin, comp_end;
  const uint64_t* words;       // packed sequences (match runs are re-extended)
  const int64_t* pat_pos;
  const int64_t* txt_pos;
  int bits;                    // 2 or 8 bits per symbol
  int endsfree;                // trace from (end_k, end_off), free leading / trailing gaps
  const int* end_k;
  const int* end_off;
  const int* end_rec;          // z-drop: the end cell's record byte (0xFF: no source)
  Penalties pen;
};

// Kernel families.  kSmall: one 64-lane wavefront per pair, wavefront ring
// and both sequences in LDS.  kLarge: one workgroup per pair, ring in HBM.
// kLean: like kSmall, gap-affine with 2-bit sequences, specialised on the
// penalties' slot structure (wfa_lean.hip).
enum Family : int { kSmall = 0, kLarge = 1, kLean = 2 };

struct LaunchShape {
  int family = 0;
  int block = 0;         // threads per workgroup
  int grid = 0;          // persistent workgroups
  size_t lds_bytes = 0;  // dynamic LDS
};

// Returns hipSuccess or the launch error.  bits: 2 (ACGT packed) or 8 (bytes).
hipError_t launch_forward(const ForwardArgs& a, int ncomp, int bits,
                          const LaunchShape& shape, hipStream_t stream);
hipError_t launch_backtrace(const BacktraceArgs& a, hipStream_t stream);
// Device counters of one run (one 64-byte block).
struct Counters {
  unsigned int queue;            // work queue of the current launch
  unsigned int retry;            // pairs to re-run (ring overflow / arena full)
  unsigned long long rec_top;    // record arena bump pointer (16-byte units)
  unsigned long long dsteps;     // diagonal-steps of the pairs that finished
  unsigned int fail;             // pairs that finished with a status other than completed
  unsigned int pad0;
  unsigned long long pad[4];
};
static_assert(sizeof(Counters) == 64, "Counters layout");
// Zeroes the queue and the arena pointer; with all=1 also retry; with all=2
// also the run totals (diagonal-steps, failures).
hipError_t launch_reset_counters(Counters* c, int all, hipStream_t stream);

// ------------------------------------------------------------------ BiWFA
// One bidirectional sub-problem: align pattern[pb,pe) with text[tb,te) of a
// top-level pair, starting in comp_begin and ending in comp_end.
struct BiProblem {
  int pair;
  int pb, pe, tb, te;
  int comp_begin, comp_end;
  int pad;
};
// kBiPairAbort: a paired-workgroup search whose partner did not answer in
// time (the host searches the sub-problem again on one workgroup)
enum BiStatus : int { kBiOk = 0, kBiEndReached = -2, kBiPairAbort = -3 };
struct BiOutcome {             // wf_bialign_breakpoint_t plus the search status
  int status;                  // kBiOk, kBiEndReached, kDevMaxSteps, kDevRingOverflow, kDevBadTrace
  int score, score_forward, score_reverse;
  int k_forward, k_reverse, offset_forward, offset_reverse;
  int component;
  int fwd_end_score;           // score at which the forward side reached its end, -1 if not
  int rev_end_score;
  int pad;
  long long dsteps;
};

struct BiArgs {
  int n;
  const BiProblem* probs;
  const uint64_t* words;       // forward sequences
  const uint64_t* rev_words;   // reversed sequences
  const int64_t* fwd_ppos;     // per top-level pair: first symbol of pattern / text ...
  const int64_t* fwd_tpos;
  const int64_t* rev_ppos;     // ... and of their reversed copies
  const int64_t* rev_tpos;
  const int* plen_total;
  const int* tlen_total;
  BiOutcome* out;
  int* ring;                   // per workgroup: forward ring, then reverse ring
  long long ring_stride;       // ints per workgroup
  int wcap;
  unsigned int* queue;
  unsigned long long* dsteps_total;
  int max_steps;
  int lmax_words;              // LDS words per sequence (four: forward and reversed pattern and text); 0: global
  HeurArgs heur;
  Penalties pen;
  // paired workgroups (launch_breakpoint_pair): per team, 4 mailbox words
  // (zeroed before the launch) and the reverse side's handover record
  unsigned long long* mail;
  int* handover;
  unsigned long long pair_wait_ticks;  // bound of every wait, 100 MHz device clock ticks
};
hipError_t launch_breakpoint(const BiArgs& a, int ncomp, int bits, int block, int grid, hipStream_t stream);
// Paired workgroups, one sub-problem per team of two (2 * a.n workgroups of
// `block` threads; 1,024 only), no heuristic cut-off; a.pen.nslots must be
// one more than the penalties' score scope needs, and every workgroup must
// fit on the device at once (the caller checks against the occupancy).
hipError_t launch_breakpoint_pair(const BiArgs& a, int ncomp, int bits, int block, hipStream_t stream);
int breakpoint_pair_max_blocks_per_cu(int ncomp, int bits, int block, size_t lds_bytes);
size_t breakpoint_pair_handover_ints(int ncomp);
// Dynamic LDS of a breakpoint launch: the four sequences of a problem.
size_t breakpoint_lds_bytes(int lmax_words);
// A piece of a final CIGAR: len operations copied from src (or, when src < 0,
// len copies of op) to dst.
struct OpsPiece {
  long long dst;
  long long src;
  int len;
  char op;
  char pad[3];
};
hipError_t launch_gather_ops(const OpsPiece* pieces, int n, const char* src, char* dst, hipStream_t stream);
hipError_t launch_reverse_pack(const uint64_t* words, const int64_t* src_pos, const int64_t* dst_pos, const int* len,
                               int n, int bits, uint64_t* out, hipStream_t stream);
int breakpoint_max_blocks_per_cu(int ncomp, int bits, int block, bool cut, size_t lds_bytes = 0);

// Run-length CIGARs (wfa_cigar.hip): per pair the runs of its operations
// ops[ops_off + ops_begin .. ops_off + plen + tlen) as (length << 4) | SAM
// code (7 '=', 8 X, 1 I, 2 D), pairs back to back.  A pair whose status is
// not 0 has none unless keep_failed (BiWFA keeps the CIGAR appended before a
// failure).
struct CigarRunsArgs {
  int n;
  int keep_failed;
  const char* ops;
  const int64_t* ops_off;
  const int* ops_begin;
  const int* plen;
  const int* tlen;
  const int* status;           // null: every pair has its CIGAR
  long long* counts;           // runs per pair
  long long* offsets;          // n + 1 entries: offsets[i] = runs before pair i
  uint32_t* runs;
};
// Counting pass and scan (temp == null: only sets *temp_bytes, the scan's storage).
hipError_t launch_cigar_runs_count(const CigarRunsArgs& a, void* temp, size_t* temp_bytes, hipStream_t stream);
hipError_t launch_cigar_runs_write(const CigarRunsArgs& a, hipStream_t stream);

// Lean family: whether the penalties' slot structure has an instance, its
// dynamic LDS, occupancy and launch (ForwardArgs as for launch_forward; the
// per-pair component arrays must be null and comp_begin = comp_end = M).
bool lean_supported(const Penalties& pen);
int lean_width(int wcap);  // smallest ring width with an instance >= wcap, 0 if none
size_t lean_lds_bytes(const Penalties& pen, int wcap, int lmax_words);
int lean_max_blocks_per_cu(const Penalties& pen, int wcap, size_t lds_bytes);
hipError_t launch_forward_lean(const ForwardArgs& a, const LaunchShape& shape, hipStream_t stream);

// Dynamic-LDS bytes and occupancy for a configuration.
size_t forward_lds_bytes(int family, int ring_slots, int wcap, int lmax_words);
int forward_max_blocks_per_cu(int family, int block, int ncomp, int bits, int cut, size_t lds_bytes);
int pick_block(int family, int max_width);

}  // namespace wfa
