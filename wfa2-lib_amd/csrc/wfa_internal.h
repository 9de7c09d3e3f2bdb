// Host-side internals shared by the C-ABI layer (wfa_api.cpp) and the BiWFA
// orchestration (wfa_bialign.cpp).  Not installed.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>
#include <vector>

#include "wfa_device.h"
#include "wfa_mi355x.h"

namespace wfa_host {

bool hip_ok(hipError_t e, const char* what);

// A fixed range of device memory handed out front to back, 256-byte aligned
// (no HIP allocation; the no-allocation entry point carves its buffers from
// the library's static device array this way).
struct ScratchRegion {
  char* base = nullptr;
  size_t bytes = 0, used = 0;
  void* take(size_t n) {
    const size_t at = (used + 255) & ~(size_t)255;
    if (at > bytes || n > bytes - at) return nullptr;
    used = at + n;
    return base + at;
  }
};

// Set on threads whose aligners run beside other aligners' kernels (the
// serving lanes).  hipFree and hipHostFree wait for every stream of the
// device: a short lane growing a buffer would wait for a long lane's
// seconds-long BiWFA kernels.  On such threads a buffer that grows keeps its
// old allocation until its owner releases it (at most about its current
// size more, as buffers grow geometrically).
extern thread_local bool t_keep_retired;

// Grow-only device allocation, freed by its owner's destructor.  hipFree
// waits for the whole device (other aligners' kernels included), so buffers
// that are used again and again grow geometrically up to max_bytes and are
// never freed between runs.  A buffer bound to a ScratchRegion takes its
// memory from the region instead and never allocates or frees.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  ScratchRegion* region = nullptr;
  std::vector<void*> retired;  // outgrown allocations kept under t_keep_retired
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), cap(o.cap), region(o.region), retired(std::move(o.retired)) {
    o.p = nullptr; o.cap = 0; o.region = nullptr;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; cap = o.cap; region = o.region; retired = std::move(o.retired);
      o.p = nullptr; o.cap = 0; o.region = nullptr;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  bool ensure(size_t bytes, size_t max_bytes = SIZE_MAX);
  void release();
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// A typed view into a DevBuf owned elsewhere (never freed through the view).
struct DevView {
  void* p = nullptr;
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Grow-only pinned host staging buffer (hipHostMalloc), freed by its owner.
struct HostBuf {
  void* p = nullptr;
  size_t cap = 0;
  std::vector<void*> retired;  // outgrown allocations kept under t_keep_retired
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() { release(); }
  bool ensure(size_t bytes);
  void release();
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// One forward launch (+ backtrace) over a subset of jobs with one kernel family.
struct PlanGroup {
  int family = 0;
  int block = 64;
  int wcap = 0;
  int lmax_words = 0;
  int max_grid = 0;            // cap on the persistent grid (0: none), set when jobs must run one at a time
  std::vector<int> ids;
  DevBuf d_ids;
  bool ids_on_device = false;  // d_ids holds ids (uploaded once per plan)
  double est_rec_bytes = -1;   // rough record footprint of the group (computed once)
  int lclass = -1;             // lean family: length class (0 short, 1 long)
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  PlanGroup() = default;
  PlanGroup(const PlanGroup&) = delete;
  PlanGroup& operator=(const PlanGroup&) = delete;
  PlanGroup(PlanGroup&& o) noexcept { *this = std::move(o); }
  PlanGroup& operator=(PlanGroup&& o) noexcept;
  ~PlanGroup() { release(); }
  void release();
};

// Per-aligner device state: one stream, the backtrace-record arena, the HBM
// wavefront rings, the run counters.
struct Engine {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  DevBuf rec;
  DevBuf ring;
  DevBuf counters;
  size_t rec_bytes = 0;     // record arena size for the next run (grows on demand)
  size_t rec_cap = 0;       // largest record arena this engine may allocate
  bool ok = false;
  // Stream on a hardware queue of its own: HIP spreads ordinary streams over a
  // few shared queues, on which kernels of different streams run in order;
  // a stream created with a CU mask (here: every CU) gets a new queue.
  bool own_queue = false;
  // With own_queue, the stream's CU mask covers the CUs [cu_lo, cu_hi) x CU
  // count (fractions); num_cus is then the count of those CUs, and the
  // persistent launches size their grids to them.
  double cu_lo = 0.0, cu_hi = 1.0;
  ::wfa_device_batch_t* scratch = nullptr;  // reused by wavefront_align / wavefront_align_batch
  // Plan groups no longer in use (their order buffers and events), handed
  // out again by take_group: retries and re-planning allocate nothing new.
  std::vector<PlanGroup> group_pool;
  // BiWFA: grow-only per-aligner buffers of the breakpoint launches, the base
  // jobs and the CIGAR gather (wfa_bialign.cpp).
  struct BiBuffers {
    DevBuf pos, len, probs, out, pieces;
    DevBuf pl, tl, pp, tp, cb, ce, st, sc, last, ds, off, beg, ops;
    DevBuf mail, hand;  // paired-workgroup searches: mailboxes and handover records
    hipEvent_t ev[2] = {nullptr, nullptr};
  } bi;
  // Set for the duration of a no-allocation run: plan groups take their
  // order arrays from `order`, and the HBM-ring family's grid is limited so
  // that its rings fit in ring_bytes (the engine's rec/ring/counters buffers
  // are then views into the same static array).  No events are recorded.
  struct NoAlloc {
    ScratchRegion order;
    size_t ring_bytes = 0;
  }* na = nullptr;
  // Single-pair calls (wavefront_align) replay one captured graph per launch
  // configuration: the upload, the counter reset, the forward and backtrace
  // launches and the downloads of the results, counters and CIGAR, under one
  // launch and one wait.  A configuration is captured the second time in a
  // row it is seen; `key` holds the bytes of the arguments and copies the
  // graph was captured with.
  struct PlainGraph {
    std::vector<unsigned char> key, seen;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool off = false;      // a capture failed: single pairs run uncaptured from then on
    long long replays = 0;
  } plain;
  HostBuf plain_out;       // pinned: results (0), counters (64) and CIGAR (256) of the single pair
  ~Engine();
};

// A group with buffers and events from the engine's pool (or fresh ones),
// its ids and plan fields cleared.
PlanGroup take_group(Engine* e);
// Returns groups' buffers and events to the engine's pool and clears the vector.
void give_back_groups(Engine* e, std::vector<PlanGroup>& groups);
// Moves the device buffers and events of old groups into new ones (same
// index); what is left goes to the engine's pool, so re-planning allocates
// nothing new when the plan does not grow.
void recycle_plan(Engine* e, std::vector<PlanGroup>& old_plan, std::vector<PlanGroup>& new_plan);

Engine* engine_of(wavefront_aligner_t* wf);
bool engine_init(Engine* e);
wfa::Counters* counters_of(Engine* e);

wfa::Penalties make_penalties(const wavefront_penalties_t& p);
int classic_score(const wavefront_penalties_t& p, int plen, int tlen, int s);

// Ends-free form of a run (alignment_endsfree); the extension form frees
// both ends entirely (pef = tef = wfa::kFreeAll).
struct EndsFree {
  int on = 0;
  int pbf = 0, pef = 0, tbf = 0, tef = 0;
};

// A set of alignment jobs on the device: pairs of a batch, or the base-case
// sub-problems of BiWFA.  Job i aligns plen[i] pattern symbols starting at
// symbol ppos[i] of `words` with tlen[i] text symbols starting at tpos[i],
// from component cbeg[i] to cend[i] (M when the arrays are null); results go
// to status/score/last/dsteps[i] and, with records, the CIGAR into
// ops[ops_off[i] + ops_begin[i] .. ops_off[i] + plen[i] + tlen[i]).
struct JobSet {
  int n = 0;
  int bits = 2;
  const int* h_plen = nullptr;
  const int* h_tlen = nullptr;
  const uint64_t* words = nullptr;
  const int* plen = nullptr;
  const int* tlen = nullptr;
  const int64_t* ppos = nullptr;
  const int64_t* tpos = nullptr;
  const signed char* cbeg = nullptr;
  const signed char* cend = nullptr;
  int* status = nullptr;
  int* score = nullptr;
  uint32_t* last = nullptr;
  long long* dsteps = nullptr;
  char* ops = nullptr;
  const int64_t* ops_off = nullptr;
  int* ops_begin = nullptr;
  int* end_k = nullptr;        // end cell (diagonal, offset) of each completed job
  int* end_off = nullptr;
  int* end_rec = nullptr;      // z-drop: record byte of the end cell (0xFF: no source)
  EndsFree ef;
  wfa::HeurArgs heur{};        // wf-adaptive (none for BiWFA base jobs)
};

// The aligner's wf-adaptive parameters (none unless the strategy is wf-adaptive).
wfa::HeurArgs heuristic_of(const wavefront_aligner_t* wf);

struct RunTotals {
  double forward_ms = 0;
  double backtrace_ms = 0;
  int launches = 0;
  unsigned long long dsteps = 0;
  unsigned int fail = 0;
};

std::vector<PlanGroup> make_plan(const JobSet& js, const std::vector<int>& ids, const wfa::Penalties& pen);

// Aligns every job of the plan (retrying the ones that overflow an LDS ring
// or the record arena); results stay on the device.  h_status is scratch
// for the host copy of the statuses.  Returns WF_STATUS_OK or WF_STATUS_OOM.
int run_jobs(wavefront_aligner_t* wf, Engine* e, const JobSet& js, std::vector<PlanGroup>& plan,
             int records, RunTotals* tot, std::vector<int>& h_status);

}  // namespace wfa_host

// A batch of pairs resident in HBM (the wfa_device_batch_t of the C ABI).
struct wfa_device_batch_t {
  int n = 0;
  int bits = 2;
  int64_t total_ops = 0;
  std::vector<int> plen, tlen;
  std::vector<int64_t> ops_off;
  std::vector<int64_t> pat_pos, txt_pos;   // symbol positions (host copy)
  int64_t words_used = 0;                  // words taken by the forward sequences
  std::vector<int> sorted;                 // pair ids, longest first
  // `words` holds the packed sequences and, after them, the per-pair inputs
  // (one host-to-device copy); `d_res` holds the per-pair results (one copy
  // back).  The d_* views point into them.
  wfa_host::DevBuf words, d_res, d_last, d_dsteps, d_ops, d_end_rec;
  wfa_host::DevView d_patw, d_txtw, d_ops_off, d_plen, d_tlen;
  wfa_host::DevView d_status, d_score, d_ops_begin, d_end_k, d_end_off;
  // reversed copies of the sequences (BiWFA), built on first use
  wfa_host::DevBuf rev_words, d_rev_ppos, d_rev_tpos;
  bool have_reverse = false;
  // launch plan, cached for the penalties it was made for
  std::vector<wfa_host::PlanGroup> plan;
  std::vector<int> plan_key;
  // results: on the device after a run, copied to the host on demand
  std::vector<int> status, score, ops_begin, end_k, end_off;
  wfa_host::EndsFree ef;                   // form of the last run
  wfa::HeurArgs heur{};                    // heuristic of the last run
  bool have_results = false;
  bool host_results = false;
  bool ops_on_device = false;
  int first_fail = 0;
  bool bi_mode = false;                    // last run was BiWFA: scores in bi_score
  std::vector<int> bi_score;
  wfa_run_stats_t stats{};
  int records = 1;
  wfa_host::HostBuf h_words;               // pinned staging of the packed sequences and per-pair inputs
  wfa_host::HostBuf h_res;                 // pinned staging of the results
  wfa_host::JobSet jobs() const;
  // (Re)fills the batch with new pairs, reusing every buffer that is large
  // enough; with cigars, the CIGAR output buffer is sized as well.  Returns
  // false on an allocation or copy failure.
  // With bucketed, the per-pair inputs start at a power-of-two word offset
  // (so that pairs of similar lengths share one layout) and nothing is
  // copied: *upload_bytes tells how much of the staging buffer to upload.
  bool load(wfa_host::Engine* e, const wfa_pair_t* pairs, int n, bool cigars, bool bucketed = false,
            size_t* upload_bytes = nullptr);
  const char* h_ops = nullptr;             // a host copy of the operations (single-pair graph), or null
  // run-length CIGARs of the last run (wfa_device_batch_download_runs), made on first request
  wfa_host::DevBuf d_run_counts, d_run_off, d_runs, d_scan_temp;
  bool runs_ready = false;
  long long total_runs = 0;
  std::vector<long long> h_run_off;
  ~wfa_device_batch_t();
};

namespace wfa_host {
// BiWFA (memory mode ultralow) over a whole batch; fills the batch's device
// results like the unidirectional run.  Returns the first failing status.
int bialign_run(wavefront_aligner_t* wf, Engine* e, wfa_device_batch_t* b, RunTotals* tot);
// Copies the results to the host (once per run) and, when cigar_dst is
// given, the operations arena with them, under one wait.
bool download_results(Engine* e, wfa_device_batch_t* b, char* cigar_dst = nullptr);
// wfa_device_batch_run with every device buffer beyond the batch's inputs and
// outputs carved from [scratch, scratch + scratch_bytes): no HIP allocation.
int device_batch_run_in(wavefront_aligner_t* wf, wfa_device_batch_t* b, char* scratch, size_t scratch_bytes);
}  // namespace wfa_host
