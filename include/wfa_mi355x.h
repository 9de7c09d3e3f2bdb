/*
 * wfa_mi355x.h -- C ABI of the MI355X-native exact WFA library.
 *
 * Drop-in for the WFA2-lib v2.3 alignment API (wavefront/wavefront_align.h,
 * wavefront/wfa.h, wavefront/wavefront_attributes.h, alignment/cigar.h): the
 * same type names, field names, enum values, defaults, entry points, status
 * codes and result members, so that callers written against the reference
 * (README.md:88-112, examples/wfa_basic.c, examples/wfa_repeated.c,
 * tools/align_benchmark/benchmark/benchmark_gap_affine.c:128-154) compile
 * unchanged.  Alignment runs on one AMD Instinct MI355X (gfx950) through HIP.
 *
 * Exact path (bit-identical scores and CIGARs to the reference):
 *   distance_metric indel / edit / gap_linear / gap_affine / gap_affine_2p,
 *   heuristic wf_heuristic_none, wf_heuristic_wfadaptive (the default
 *   attributes), wf_heuristic_wfmash, wf_heuristic_xdrop / _zdrop,
 *   wf_heuristic_banded_static / _banded_adaptive, alone or combined,
 *   compute_score or compute_alignment;
 *   alignment_end2end in memory modes high / med / low / ultralow;
 *   alignment_endsfree and the extension form in memory modes high / med /
 *   low (the reference rejects them in ultralow with exit(1), and so does
 *   this library), ending on the first diagonal in ascending order as the
 *   reference's scalar extend kernel does.
 * Configurations outside it are rejected with WF_STATUS_UNSUPPORTED and a
 * "[WFA]" message on stderr; they never silently produce different results:
 * z-drop alignments with a CIGAR in memory modes med / low, lambda sequences,
 * free beginnings with a match bonus or with
 * indel/edit distances, indel/edit extension outside memory mode high, and
 * (per pair) a banded alignment whose reference backtrace would take
 * operations back through a wavefront the cut-off emptied.
 * Configuration errors (invalid penalties, ends-free lengths larger than the
 * sequences, x-drop / z-drop with indel or edit) print the reference's
 * message and exit(1) as it does.
 *
 * Not thread-safe per aligner (as the reference, wavefront_aligner.c); one
 * aligner per host thread, any number of aligners per device.
 */
#ifndef WFA_MI355X_H_
#define WFA_MI355X_H_

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status
 * wavefront/wfa.h:45-55 */
#define WF_STATUS_ALG_COMPLETED            0
#define WF_STATUS_ALG_PARTIAL              1
#define WF_STATUS_MAX_STEPS_REACHED     -100
#define WF_STATUS_OOM                   -200
#define WF_STATUS_UNATTAINABLE          -300
#define WF_STATUS_OK                      -1
#define WF_STATUS_END_REACHED             -2
#define WF_STATUS_END_UNREACHABLE         -3
/* No counterpart in the reference: configuration outside the exact device path. */
#define WF_STATUS_UNSUPPORTED           -400

/* ------------------------------------------------------------- penalties
 * alignment/{linear,affine,affine2p}_penalties.h, wavefront_penalties.h:41-67 */
typedef enum {
  indel         = 0,
  edit          = 1,
  gap_linear    = 2,
  gap_affine    = 3,
  gap_affine_2p = 4
} distance_metric_t;

typedef struct { int match; int mismatch; int indel; } linear_penalties_t;
typedef struct { int match; int mismatch; int gap_opening; int gap_extension; } affine_penalties_t;
typedef struct {
  int match;
  int mismatch;
  int gap_opening1;
  int gap_extension1;
  int gap_opening2;
  int gap_extension2;
} affine2p_penalties_t;

typedef enum {
  affine2p_matrix_M,
  affine2p_matrix_I1,
  affine2p_matrix_I2,
  affine2p_matrix_D1,
  affine2p_matrix_D2
} affine2p_matrix_type;

typedef struct {
  distance_metric_t distance_metric;
  int match;
  int mismatch;
  int gap_opening1;
  int gap_extension1;
  int gap_opening2;
  int gap_extension2;
  linear_penalties_t linear_penalties;
  affine_penalties_t affine_penalties;
  affine2p_penalties_t affine2p_penalties;
  int internal_gap_e;
} wavefront_penalties_t;

/* ------------------------------------------------------------------ CIGAR
 * alignment/cigar.h:45-59 */
typedef struct {
  char* operations;
  int max_operations;
  int begin_offset;
  int end_offset;
  int score;
  int end_v;
  int end_h;
  bool has_misms;
  uint32_t* cigar_buffer;
  int cigar_length;
} cigar_t;

/* ------------------------------------------------------------ attributes
 * wavefront/wavefront_attributes.h:51-133, wavefront_heuristic.h:42-69 */
typedef enum { compute_score, compute_alignment } alignment_scope_t;
typedef enum { alignment_end2end, alignment_endsfree } alignment_span_t;
typedef struct {
  alignment_span_t span;
  bool extension;
  int pattern_begin_free;
  int pattern_end_free;
  int text_begin_free;
  int text_end_free;
} alignment_form_t;

typedef struct {
  uint64_t total;
  uint64_t samples;
  uint64_t min;
  uint64_t max;
  double m_oldM;
  double m_newM;
  double m_oldS;
  double m_newS;
} profiler_counter_t;
typedef struct {
  struct timespec begin_timer;
  profiler_counter_t time_ns;
  uint64_t accumulated;
} profiler_timer_t;

typedef struct {
  int max_alignment_steps;
  int probe_interval_global;
  int probe_interval_compact;
  uint64_t max_partial_compacts;
  uint64_t max_memory_compact;
  uint64_t max_memory_resident;
  uint64_t max_memory_abort;
  int verbose;
  bool check_alignment_correct;
  profiler_timer_t timer;
  int max_num_threads;
  int min_offsets_per_thread;
} alignment_system_t;

typedef enum {
  wavefront_memory_high     = 0,
  wavefront_memory_med      = 1,
  wavefront_memory_low      = 2,
  wavefront_memory_ultralow = 3,
} wavefront_memory_t;

typedef enum {
  wf_heuristic_none            = 0x0000000000000000ul,
  wf_heuristic_banded_static   = 0x0000000000000001ul,
  wf_heuristic_banded_adaptive = 0x0000000000000002ul,
  wf_heuristic_wfadaptive      = 0x0000000000000004ul,
  wf_heuristic_xdrop           = 0x0000000000000010ul,
  wf_heuristic_zdrop           = 0x0000000000000020ul,
  wf_heuristic_wfmash          = 0x0000000000000040ul,
} wf_heuristic_strategy;
typedef struct {
  wf_heuristic_strategy strategy;
  int steps_between_cutoffs;
  int min_k;
  int max_k;
  int min_wavefront_length;
  int max_distance_threshold;
  int xdrop;
  int zdrop;
  int steps_wait;
  int max_sw_score;
  int max_wf_score;
  int max_sw_score_offset;
  int max_sw_score_k;
} wavefront_heuristic_t;

typedef struct {
  bool enabled;
  int resolution_points;
  int align_level;
} wavefront_plot_attr_t;

/* External allocators are accepted for source compatibility and ignored:
   device memory comes from the library's own arena. */
typedef struct _mm_allocator_t mm_allocator_t;

typedef struct {
  distance_metric_t distance_metric;
  alignment_scope_t alignment_scope;
  alignment_form_t alignment_form;
  linear_penalties_t linear_penalties;
  affine_penalties_t affine_penalties;
  affine2p_penalties_t affine2p_penalties;
  wavefront_heuristic_t heuristic;
  wavefront_memory_t memory_mode;
  mm_allocator_t* mm_allocator;
  wavefront_plot_attr_t plot;
  alignment_system_t system;
} wavefront_aligner_attr_t;

/* Identical defaults to wavefront_attributes.c:38-100, including the
   heuristic wf-adaptive(10,50,1), which the device runs with the reference's
   results; set heuristic.strategy = wf_heuristic_none (or call
   wavefront_aligner_set_heuristic_none) for exact alignment. */
extern wavefront_aligner_attr_t wavefront_aligner_attr_default;

/* ---------------------------------------------------------------- aligner
 * wavefront/wfa.h:60-120 */
typedef struct _wavefront_aligner_t wavefront_aligner_t;

typedef struct {
  int status;
  int score;
  bool dropped;
  int num_null_steps;
  uint64_t memory_used;
  void (*wf_align_compute)(wavefront_aligner_t* const, const int);
  int (*wf_align_extend)(wavefront_aligner_t* const, const int);
} wavefront_align_status_t;

typedef enum {
  wf_align_regular = 0,
  wf_align_biwfa = 1,
  wf_align_biwfa_breakpoint_forward = 2,
  wf_align_biwfa_breakpoint_reverse = 3,
  wf_align_biwfa_subsidiary = 4
} wavefront_align_mode_t;

typedef struct {
  int score;
  int k;
  int32_t offset;
} wavefront_pos_t;

/* The pair last aligned (wavefront_sequences.h members callers read). */
typedef struct {
  char* pattern;
  int pattern_length;
  char* text;
  int text_length;
} wavefront_sequences_t;

struct _wavefront_aligner_t {
  wavefront_align_mode_t align_mode;
  char* align_mode_tag;
  wavefront_align_status_t align_status;
  wavefront_sequences_t sequences;
  alignment_scope_t alignment_scope;
  alignment_form_t alignment_form;
  wavefront_penalties_t penalties;
  wavefront_heuristic_t heuristic;
  wavefront_memory_t memory_mode;
  affine2p_matrix_type component_begin;
  affine2p_matrix_type component_end;
  wavefront_pos_t alignment_end_pos;
  cigar_t* cigar;
  alignment_system_t system;
  void* device;                   /* opaque device state (stream, arena) */
};

/* Lifecycle, wavefront_aligner.c:422-504 */
wavefront_aligner_t* wavefront_aligner_new(wavefront_aligner_attr_t* attributes);
void wavefront_aligner_reap(wavefront_aligner_t* const wf_aligner);
void wavefront_aligner_delete(wavefront_aligner_t* const wf_aligner);

/* Configuration, wavefront_aligner.c:508-643 */
void wavefront_aligner_set_alignment_end_to_end(wavefront_aligner_t* const wf_aligner);
void wavefront_aligner_set_alignment_free_ends(wavefront_aligner_t* const wf_aligner,
    const int pattern_begin_free, const int pattern_end_free,
    const int text_begin_free, const int text_end_free);
void wavefront_aligner_set_alignment_extension(wavefront_aligner_t* const wf_aligner);
void wavefront_aligner_set_heuristic_none(wavefront_aligner_t* const wf_aligner);
void wavefront_aligner_set_heuristic_wfadaptive(wavefront_aligner_t* const wf_aligner,
    const int min_wavefront_length, const int max_distance_threshold, const int score_steps);
void wavefront_aligner_set_heuristic_wfmash(wavefront_aligner_t* const wf_aligner,
    const int min_wavefront_length, const int max_distance_threshold, const int score_steps);
void wavefront_aligner_set_heuristic_xdrop(wavefront_aligner_t* const wf_aligner,
    const int xdrop, const int score_steps);
void wavefront_aligner_set_heuristic_zdrop(wavefront_aligner_t* const wf_aligner,
    const int ydrop, const int score_steps);
void wavefront_aligner_set_heuristic_banded_static(wavefront_aligner_t* const wf_aligner,
    const int band_min_k, const int band_max_k);
void wavefront_aligner_set_heuristic_banded_adaptive(wavefront_aligner_t* const wf_aligner,
    const int band_min_k, const int band_max_k, const int score_steps);
void wavefront_aligner_set_max_alignment_steps(wavefront_aligner_t* const wf_aligner,
    const int max_alignment_steps);
void wavefront_aligner_set_max_memory(wavefront_aligner_t* const wf_aligner,
    const uint64_t max_memory_resident, const uint64_t max_memory_abort);
void wavefront_aligner_set_max_num_threads(wavefront_aligner_t* const wf_aligner,
    const int max_num_threads);
void wavefront_aligner_set_min_offsets_per_thread(wavefront_aligner_t* const wf_aligner,
    const int min_offsets_per_thread);

/* Messages, wavefront_aligner.c:71-98 */
char* wavefront_align_strerror(const int error_code);
char* wavefront_align_strerror_short(const int error_code);

/* Align, wavefront_align.c:150-266.  Synchronous; inputs are copied. */
int wavefront_align(wavefront_aligner_t* const wf_aligner,
    const char* const pattern, const int pattern_length,
    const char* const text, const int text_length);
typedef int (*alignment_match_funct_t)(int, int, void*);  /* wavefront_sequences.h:55 */
int wavefront_align_lambda(wavefront_aligner_t* const wf_aligner,
    alignment_match_funct_t const match_funct, void* match_funct_arguments,
    const int pattern_length, const int text_length);
/* 2 bits per base, 4 bases per byte, base i in bits 2*(i%4) of byte i/4
   (A=0, C=1, G=2, T=3), the layout wavefront_sequences.c:102-140 decodes. */
int wavefront_align_packed2bits(wavefront_aligner_t* const wf_aligner,
    const uint8_t* const pattern, const int pattern_length,
    const uint8_t* const text, const int text_length);
int wavefront_align_resume(wavefront_aligner_t* const wf_aligner);

/* ------------------------------------------------------------ batch entry
 * No counterpart in the reference.  Per pair, identical to wavefront_align
 * with the same aligner attributes, for num_pairs pairs in one device pass. */
typedef struct {
  const char* pattern;
  int pattern_length;
  const char* text;
  int text_length;
} wfa_pair_t;

typedef struct {
  int status;            /* as wavefront_align's return value */
  int score;             /* as cigar->score */
  int end_v;             /* as cigar->end_v */
  int end_h;             /* as cigar->end_h */
  int64_t cigar_offset;  /* operations at cigar_arena[cigar_offset .. +cigar_length) */
  int cigar_length;      /* 0 for score-only or failed pairs */
} wfa_result_t;

/* Bytes of cigar_arena that wavefront_align_batch needs for these pairs:
   sum of (pattern_length + text_length). */
size_t wavefront_align_batch_arena_bytes(const wfa_pair_t* pairs, int num_pairs);

/* Returns WF_STATUS_ALG_COMPLETED, or the first failing per-pair status.
   The caller owns pairs, results and cigar_arena (may be NULL when the
   aligner is score-only).  Operations are M/X/I/D as in cigar_t. */
int wavefront_align_batch(wavefront_aligner_t* const wf_aligner,
    const wfa_pair_t* pairs, int num_pairs, wfa_result_t* results,
    char* cigar_arena, size_t cigar_arena_bytes);

/* ------------------------------------------------ device-resident batches
 * No counterpart in the reference.  For callers that align the same input
 * repeatedly, or that want the inputs resident in HBM before timing:
 * upload once, run any number of times, download results when needed. */
typedef struct wfa_device_batch_t wfa_device_batch_t;
wfa_device_batch_t* wfa_device_batch_upload(wavefront_aligner_t* const wf_aligner,
    const wfa_pair_t* pairs, int num_pairs);
/* Align every pair of the batch on the device; results stay on the device. */
int wfa_device_batch_run(wavefront_aligner_t* const wf_aligner, wfa_device_batch_t* batch);
/* Copy results (and CIGARs, when cigar_arena is not NULL) to the host. */
int wfa_device_batch_download(wavefront_aligner_t* const wf_aligner, wfa_device_batch_t* batch,
    wfa_result_t* results, char* cigar_arena, size_t cigar_arena_bytes);
void wfa_device_batch_free(wfa_device_batch_t* batch);
/* Run-length CIGARs, made on the device from the last run: each pair's
 * CIGAR as the runs cigar_get_CIGAR returns with show_mismatches
 * (alignment/cigar.c:180-239): (length << 4) | code, code 7 ('=') for
 * matches, 8 (X), 1 (I), 2 (D), pairs back to back.  They are a few times
 * smaller than the operation bytes, so fewer bytes cross PCIe.
 * wfa_device_batch_cigar_runs returns the number of runs (to size `runs`),
 * or a negative status; wfa_device_batch_download_runs fills results as
 * wfa_device_batch_download does, with cigar_offset / cigar_length counting
 * runs of `runs`.  Not for the extension form or z-drop (their CIGARs are
 * trimmed on the host: WF_STATUS_UNSUPPORTED). */
int64_t wfa_device_batch_cigar_runs(wavefront_aligner_t* const wf_aligner, wfa_device_batch_t* batch);
int wfa_device_batch_download_runs(wavefront_aligner_t* const wf_aligner, wfa_device_batch_t* batch,
    wfa_result_t* results, uint32_t* runs, size_t max_runs);
/* Statistics of the last run: total diagonal-steps (sum over non-null score
   steps of hi-lo+1) and forward / backtrace kernel milliseconds (HIP events
   on the aligner's stream). */
typedef struct {
  int64_t diagonal_steps;
  double forward_ms;
  double backtrace_ms;
  int forward_launches;
  int64_t record_arena_bytes;  /* the aligner's backtrace-record arena after the run */
} wfa_run_stats_t;
int wfa_device_batch_stats(wfa_device_batch_t* batch, wfa_run_stats_t* stats);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP stream capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same run as wfa_device_batch_run, with every device
 * buffer beyond the batch's inputs and outputs (run counters, launch order
 * arrays, HBM wavefront rings, backtrace-record arena) taken from one static
 * 256 MB device array per GPU, held until the process ends.  No HIP
 * allocation or free is made.  The batch must have been uploaded with this
 * aligner (wfa_device_batch_upload sizes its CIGAR output for aligners that
 * compute alignments).  Limits, each refused with a [WFA] message before any
 * device work: at most about 20 M pairs per call; each pair's wavefront ring
 * (4 x ring slots x (n + m + 8) bytes) within 64 MB; memory mode ultralow is
 * refused (WF_STATUS_UNSUPPORTED).  Batches whose records exceed the arena
 * run in waves; a pair whose own records do not fit ends WF_STATUS_OOM.
 * Calls that overlap on one GPU share the array and race: serialise them.
 * No kernel timings are kept (wfa_run_stats_t forward_ms / backtrace_ms 0). */
int wfa_device_batch_run_noalloc(wavefront_aligner_t* const wf_aligner, wfa_device_batch_t* batch);

/* ------------------------------------------------------------ serving
 * No counterpart in the reference (its wavefront_align is synchronous per
 * pair).  A request loop for single-pair requests that arrive over time:
 * requests are queued by length class (short / medium with the memory-high
 * algorithm, long with BiWFA), and worker lanes -- each an aligner with its
 * own HIP stream and resident device buffers -- admit the queued requests of
 * their class into device batches, up to max_batch_pairs and up to the lane's
 * memory budget on each request's estimated footprint: O(s^2) bytes of
 * backtrace records for the memory-high classes and O(s) for BiWFA, with the
 * score s estimated from the length, the penalties and error_prior.  Batches
 * of different lanes run concurrently on the GPU; with two lanes per class,
 * one admits new requests while the other's batch is on the device.  Per
 * request, results are identical to wavefront_align with the same
 * attributes. */
typedef struct wfa_server_t wfa_server_t;
typedef struct {
  int short_lanes;          /* lanes for requests with max(n, m) <= short_max_length */
  int medium_lanes;         /* lanes for requests between the two limits */
  int long_lanes;           /* lanes for requests with max(n, m) >= long_min_length (BiWFA) */
  int short_max_length;
  int long_min_length;
  int max_batch_pairs;      /* requests admitted into one device batch */
  size_t lane_memory_bytes; /* per-lane budget: record arena and admission limit */
  double error_prior;       /* expected error rate of a request, for its footprint estimate (0.1) */
  /* Shares of the device's CUs per class (0.125, 0.375, 0.5): the CUs are
     split short | medium | long from the top (short and medium must not
     overlap: short + medium <= 1; long takes the lowest CUs and may overlap
     them), so one class's long-running kernels never hold another class's
     batches off the CUs.  Each share is at least 0.125 (one of the eight
     accelerator dies) and at most 1; wfa_server_create refuses others. */
  double short_cu_share, medium_cu_share, long_cu_share;
} wfa_server_config_t;

typedef struct {
  int64_t id;               /* as returned by wfa_server_submit */
  int status;               /* as wavefront_align's return value */
  int score;
  int64_t cigar_offset;     /* operations at cigar_arena[cigar_offset .. +cigar_length) */
  int cigar_length;
  double submit_time;       /* seconds on wfa_server_clock(): submitted, */
  double start_time;        /* its device batch started, */
  double finish_time;       /* and finished */
} wfa_server_result_t;

void wfa_server_config_default(wfa_server_config_t* config);
/* attributes: as for wavefront_aligner_new (memory_mode is chosen per class). */
wfa_server_t* wfa_server_create(const wavefront_aligner_attr_t* attributes, const wfa_server_config_t* config);
/* Copies the sequences.  Thread-safe.  Returns the request id (>= 0). */
int64_t wfa_server_submit(wfa_server_t* server,
    const char* pattern, int pattern_length, const char* text, int text_length);
/* Submits pairs[i] at arrival_s[i] seconds after the call, from a thread of
   the server (the pairs are copied first); returns at once. */
int wfa_server_replay(wfa_server_t* server, const wfa_pair_t* pairs, const double* arrival_s, int num_pairs);
/* Waits up to timeout_ms for finished requests, then copies up to
   max_results of them (in completion order) and their CIGARs into
   cigar_arena.  Returns the number copied. */
int wfa_server_collect(wfa_server_t* server, wfa_server_result_t* results, int max_results,
    char* cigar_arena, size_t cigar_arena_bytes, int timeout_ms);
/* Requests submitted and not yet collected. */
int64_t wfa_server_pending(wfa_server_t* server);
double wfa_server_clock(void);
/* Finishes the replays and every queued request, then frees the server. */
void wfa_server_destroy(wfa_server_t* server);

/* ------------------------------------------------------------ file I/O (new)
 * The input and output formats of the reference's align_benchmark tool,
 * streamed through the device: a .seq file of line pairs (">pattern",
 * "<text"; tools/align_benchmark/align_benchmark.c:313-339) is aligned in
 * batches of batch_pairs pairs (0: 65536) with the aligner's configuration,
 * and one .alg line per pair is written (benchmark_utils.c:103-179):
 * "score\tCIGAR", or with full_output "plen\ttlen\tscore\tpattern\ttext\tCIGAR".
 * Parsing, device work and formatting of consecutive batches overlap.
 * Returns 0 or the first failing per-pair status. */
typedef struct {
  int64_t pairs;
  int64_t bytes_in;
  int64_t bytes_out;
  double seconds;         /* whole call */
  double read_seconds;    /* reader thread: parsing */
  double align_seconds;   /* calling thread: wavefront_align_batch */
  double write_seconds;   /* writer thread: formatting and writing */
} wfa_io_stats_t;
int wfa_align_seq_file(wavefront_aligner_t* const wf_aligner, const char* input_path, const char* output_path,
    int batch_pairs, int full_output, wfa_io_stats_t* stats);

/* ------------------------------------------------------------ CIGAR utils
 * alignment/cigar.c:244-470 */
bool cigar_is_null(const cigar_t* const cigar);
int cigar_score_edit(const cigar_t* const cigar);                                          /* cigar.c:244-259 */
int cigar_score_gap_linear(const cigar_t* const cigar, const linear_penalties_t* const penalties); /* :260-276 */
int cigar_score_gap_affine(const cigar_t* const cigar, const affine_penalties_t* const penalties);
int cigar_score_gap_affine2p(const cigar_t* const cigar, const affine2p_penalties_t* const penalties);
bool cigar_check_alignment(FILE* const stream,
    const char* const pattern, const int pattern_length,
    const char* const text, const int text_length,
    const cigar_t* const cigar, const bool verbose);
int cigar_sprint(char* const buffer, const cigar_t* const cigar, const bool print_matches);
/* SAM CIGAR runs (length << 4 | BAM code) and their text (alignment/cigar.c:180-239, :470-502) */
void cigar_get_CIGAR(cigar_t* const cigar, const bool show_mismatches, uint32_t** const cigar_buffer,
    int* const cigar_length);
int cigar_sprint_SAM_CIGAR(char* const buffer, cigar_t* const cigar, const bool show_mismatches);
void cigar_print_SAM_CIGAR(FILE* const stream, cigar_t* const cigar, const bool show_mismatches);
void cigar_print(FILE* const stream, const cigar_t* const cigar, const bool print_matches);
void cigar_print_pretty(FILE* const stream, const cigar_t* const cigar,
    const char* const pattern, const int pattern_length,
    const char* const text, const int text_length);

#ifdef __cplusplus
}
#endif

#endif /* WFA_MI355X_H_ */
