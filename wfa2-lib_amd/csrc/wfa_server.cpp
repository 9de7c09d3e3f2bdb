// Request loop for alignments that arrive one at a time (no counterpart in
// the reference library, whose wavefront_align is synchronous per pair).
//
// Requests are classified by length: short and medium ones are aligned with
// the memory-high algorithm, long ones with BiWFA (O(s) memory).  Each class
// has its own queue and its own worker lanes; a lane owns an aligner (one HIP
// stream, a resident batch and a record arena) and runs back-to-back device
// batches.  A free lane admits the queued requests of its class in arrival
// order, up to a pair count and up to its device-memory budget on the
// estimated footprint of each request (O(s^2) bytes of backtrace records for
// memory-high, O(s) for BiWFA, s estimated from the length and an error
// prior), so requests that arrive while
// other batches are on the GPU join the next batch of any free lane and the
// batches of different lanes overlap on the device.  Results and their CIGARs
// are queued for wfa_server_collect with submit / start / finish times.
#include <algorithm>
#include <climits>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "wfa_internal.h"

namespace {

double now_s() {
  using clock = std::chrono::steady_clock;
  return std::chrono::duration<double>(clock::now().time_since_epoch()).count();
}

struct Request {
  int64_t id;
  std::string pattern, text;
  double submit;
  double footprint;  // estimated device bytes
};

struct Done {
  int64_t id;
  int status, score;
  std::string cigar;
  double submit, start, finish;
};

enum Class { kShort = 0, kMedium = 1, kLong = 2, kClasses = 3 };

struct Queue {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::unique_ptr<Request>> q;
};

}  // namespace

struct wfa_server_t {
  wfa_server_config_t cfg{};
  wavefront_aligner_attr_t attr{};
  Queue queues[kClasses];
  std::mutex done_mu;
  std::condition_variable done_cv;
  std::deque<Done> done;
  std::vector<std::thread> lanes;
  std::vector<std::thread> replayers;
  std::atomic<int64_t> next_id{0};
  std::atomic<int64_t> outstanding{0};
  std::atomic<bool> stop{false};
  std::atomic<int> lane_errors{0};

  int classify(int plen, int tlen) const {
    const int L = std::max(plen, tlen);
    if (L >= cfg.long_min_length) return kLong;
    if (L <= cfg.short_max_length) return kShort;
    return kMedium;
  }
  // Estimated alignment score of a request: error_prior * L errors, equally
  // substitutions, 1-symbol insertions and 1-symbol deletions, at the
  // aligner's penalties.
  double score_estimate(double L) const {
    double sub = 1, gap = 1;
    switch (attr.distance_metric) {
      case indel: sub = 2; gap = 1; break;
      case edit: sub = 1; gap = 1; break;
      case gap_linear: sub = attr.linear_penalties.mismatch; gap = attr.linear_penalties.indel; break;
      case gap_affine:
        sub = attr.affine_penalties.mismatch;
        gap = attr.affine_penalties.gap_opening + attr.affine_penalties.gap_extension;
        break;
      default:
        sub = attr.affine2p_penalties.mismatch;
        gap = std::min(attr.affine2p_penalties.gap_opening1 + attr.affine2p_penalties.gap_extension1,
                       attr.affine2p_penalties.gap_opening2 + attr.affine2p_penalties.gap_extension2);
    }
    return std::max(1.0, cfg.error_prior * L * (sub + 2 * gap) / 3);
  }
  // Device bytes a request is expected to take.  Memory-high: its backtrace
  // records, one byte per diagonal-step plus 16 per score step, with about
  // 0.2 s^2 diagonal-steps (9 L^2 e^2 measured on gap-affine (0,4,6,2) reads,
  // where s is about 6.7 e L), plus its packed sequences and CIGAR: O(s^2).
  // BiWFA: the breakpoint rings of both directions (5 slots x 3 components x
  // (n + m) int32 each, the device's allocation) and the records of its
  // base cases (scores under 250): O(s) beyond the sequences.
  double footprint(int cls, int plen, int tlen) const {
    const double L = std::max(plen, tlen);
    const double s = score_estimate(L);
    if (cls == kLong) return 2.0 * 5 * 3 * 4 * (plen + tlen + 8.0) + 64.0 * s + 2 * L + (1 << 20);
    return 0.2 * s * s + 16 * s + 2 * L + 4096;
  }
  void lane_main(int cls);
};

void wfa_server_t::lane_main(int cls) {
  wfa_host::t_keep_retired = true;  // no hipFree while other lanes' kernels run
  wavefront_aligner_attr_t a = attr;
  a.memory_mode = cls == kLong ? wavefront_memory_ultralow : wavefront_memory_high;
  wavefront_aligner_t* wf = wavefront_aligner_new(&a);
  if (!wf) {
    ++lane_errors;
    return;
  }
  // The lane's record arena is its memory budget.  Each lane has a hardware
  // queue of its own: on a shared queue, one lane's kernels (seconds for a
  // long batch, hundreds of milliseconds for a medium one) would hold up the
  // batches of the others (streams sharing a queue run in order).
  if (wfa_host::Engine* e = wfa_host::engine_of(wf)) {
    e->rec_cap = cfg.lane_memory_bytes;
    e->own_queue = true;
    // CUs from the top: short | medium | long
    const double s = cfg.short_cu_share, m = cfg.medium_cu_share, l = cfg.long_cu_share;
    if (cls == kShort) { e->cu_lo = 1.0 - s; e->cu_hi = 1.0; }
    else if (cls == kMedium) { e->cu_lo = 1.0 - s - m; e->cu_hi = 1.0 - s; }
    else { e->cu_lo = 0.0; e->cu_hi = l; }
  }
  Queue& Q = queues[cls];
  std::vector<std::unique_ptr<Request>> batch;
  std::vector<wfa_pair_t> pairs;
  std::vector<wfa_result_t> results;
  std::vector<char> arena;
  for (;;) {
    batch.clear();
    {
      std::unique_lock<std::mutex> lk(Q.mu);
      Q.cv.wait(lk, [&] { return stop.load() || !Q.q.empty(); });
      if (Q.q.empty()) break;  // stopping and drained
      double used = 0;
      while (!Q.q.empty() && (int)batch.size() < cfg.max_batch_pairs) {
        const double f = Q.q.front()->footprint;
        if (!batch.empty() && used + f > (double)cfg.lane_memory_bytes) break;
        used += f;
        batch.push_back(std::move(Q.q.front()));
        Q.q.pop_front();
      }
    }
    const double start = now_s();
    const int n = (int)batch.size();
    pairs.resize(n);
    results.resize(n);
    size_t bytes = 0;
    for (int i = 0; i < n; ++i) {
      Request& r = *batch[i];
      pairs[i] = wfa_pair_t{r.pattern.data(), (int)r.pattern.size(), r.text.data(), (int)r.text.size()};
      bytes += r.pattern.size() + r.text.size();
    }
    if (arena.size() < std::max<size_t>(bytes, 1)) arena.resize(std::max<size_t>(bytes, 1));
    for (auto& r : results) r = wfa_result_t{WF_STATUS_OOM, INT32_MIN, -1, -1, 0, 0};  // if the batch fails early
    (void)wavefront_align_batch(wf, pairs.data(), n, results.data(), arena.data(), arena.size());
    const double finish = now_s();
    {
      std::lock_guard<std::mutex> lk(done_mu);
      for (int i = 0; i < n; ++i) {
        Done d;
        d.id = batch[i]->id;
        d.status = results[i].status;
        d.score = results[i].score;
        if (d.status == WF_STATUS_ALG_COMPLETED && results[i].cigar_length > 0)
          d.cigar.assign(arena.data() + results[i].cigar_offset, (size_t)results[i].cigar_length);
        d.submit = batch[i]->submit;
        d.start = start;
        d.finish = finish;
        done.push_back(std::move(d));
      }
    }
    done_cv.notify_all();
  }
  wavefront_aligner_delete(wf);
}

extern "C" void wfa_server_config_default(wfa_server_config_t* c) {
  // Two lanes for the short and medium classes: while one lane's batch is on
  // the device, the other admits the requests that arrive meanwhile.  Every
  // lane's stream has a hardware queue of its own (CU-masked stream), so the
  // batches of different lanes overlap on the device.
  c->short_lanes = 2;
  c->medium_lanes = 2;
  c->long_lanes = 1;
  c->error_prior = 0.1;
  // The persistent kernels of a batch hold their CUs until the batch ends
  // (seconds for 100 kbp BiWFA): without a split, the p99 of short requests
  // was 1.5-2 s of waiting for CUs.  Short requests need few CUs; long ones
  // the most (one CU per 100 kbp problem for seconds).
  c->short_cu_share = 0.125;
  c->medium_cu_share = 0.375;
  c->long_cu_share = 0.5;
  c->short_max_length = 2048;
  c->long_min_length = 32768;
  c->max_batch_pairs = 8192;
  c->lane_memory_bytes = (size_t)8 << 30;
}

extern "C" wfa_server_t* wfa_server_create(const wavefront_aligner_attr_t* attributes,
                                           const wfa_server_config_t* config) {
  wfa_server_t* s = new (std::nothrow) wfa_server_t();
  if (!s) return nullptr;
  if (config) s->cfg = *config;
  else wfa_server_config_default(&s->cfg);
  s->attr = attributes ? *attributes : wavefront_aligner_attr_default;
  if (s->cfg.max_batch_pairs < 1) s->cfg.max_batch_pairs = 1;
  if (!(s->cfg.error_prior > 0 && s->cfg.error_prior <= 1)) s->cfg.error_prior = 0.1;
  const double sh_s = s->cfg.short_cu_share, sh_m = s->cfg.medium_cu_share, sh_l = s->cfg.long_cu_share;
  bool shares_ok = sh_s + sh_m <= 1.0 + 1e-9;
  for (double sh : {sh_s, sh_m, sh_l}) shares_ok = shares_ok && sh >= 0.125 && sh <= 1.0;
  if (!shares_ok) {
    fprintf(stderr, "[WFA] wfa_server_create: CU shares (%g, %g, %g) must each lie in [0.125, 1] with short + "
            "medium <= 1\n", sh_s, sh_m, sh_l);
    delete s;
    return nullptr;
  }
  const int counts[kClasses] = {std::max(1, s->cfg.short_lanes), std::max(1, s->cfg.medium_lanes),
                                std::max(1, s->cfg.long_lanes)};
  for (int c = 0; c < kClasses; ++c)
    for (int i = 0; i < counts[c]; ++i) s->lanes.emplace_back(&wfa_server_t::lane_main, s, c);
  return s;
}

extern "C" int64_t wfa_server_submit(wfa_server_t* s, const char* pattern, int plen, const char* text, int tlen) {
  if (!s || s->stop.load() || plen < 0 || tlen < 0) return WF_STATUS_UNATTAINABLE;
  auto r = std::make_unique<Request>();
  r->id = s->next_id.fetch_add(1);
  r->pattern.assign(pattern, (size_t)plen);
  r->text.assign(text, (size_t)tlen);
  r->submit = now_s();
  const int cls = s->classify(plen, tlen);
  r->footprint = s->footprint(cls, plen, tlen);
  const int64_t id = r->id;
  s->outstanding.fetch_add(1);
  {
    std::lock_guard<std::mutex> lk(s->queues[cls].mu);
    s->queues[cls].q.push_back(std::move(r));
  }
  s->queues[cls].cv.notify_one();
  return id;
}

extern "C" int wfa_server_collect(wfa_server_t* s, wfa_server_result_t* out, int max_results, char* cigar_arena,
                                  size_t cigar_arena_bytes, int timeout_ms) {
  if (!s || max_results <= 0) return 0;
  std::unique_lock<std::mutex> lk(s->done_mu);
  if (s->done.empty() && timeout_ms != 0) {
    s->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms),
                        [&] { return !s->done.empty(); });
  }
  int n = 0;
  size_t used = 0;
  while (n < max_results && !s->done.empty()) {
    Done& d = s->done.front();
    if (!d.cigar.empty() && (cigar_arena == nullptr || used + d.cigar.size() > cigar_arena_bytes)) {
      if (n == 0 && cigar_arena != nullptr) {
        fprintf(stderr, "[WFA] wfa_server_collect: a CIGAR of %zu bytes does not fit the arena\n", d.cigar.size());
      }
      if (cigar_arena != nullptr) break;
    }
    wfa_server_result_t& r = out[n];
    r.id = d.id;
    r.status = d.status;
    r.score = d.score;
    r.cigar_offset = (int64_t)used;
    r.cigar_length = 0;
    if (cigar_arena && !d.cigar.empty()) {
      memcpy(cigar_arena + used, d.cigar.data(), d.cigar.size());
      r.cigar_length = (int)d.cigar.size();
      used += d.cigar.size();
    }
    r.submit_time = d.submit;
    r.start_time = d.start;
    r.finish_time = d.finish;
    s->done.pop_front();
    ++n;
  }
  s->outstanding.fetch_sub(n);
  return n;
}

extern "C" int64_t wfa_server_pending(wfa_server_t* s) { return s ? s->outstanding.load() : 0; }

extern "C" double wfa_server_clock(void) { return now_s(); }

extern "C" int wfa_server_replay(wfa_server_t* s, const wfa_pair_t* pairs, const double* arrival_s, int n) {
  if (!s || n < 0) return WF_STATUS_UNATTAINABLE;
  // Own copies: the caller may free its arrays after this returns.
  auto req = std::make_shared<std::vector<std::pair<std::string, std::string>>>();
  auto at = std::make_shared<std::vector<double>>(arrival_s, arrival_s + n);
  req->reserve(n);
  for (int i = 0; i < n; ++i)
    req->emplace_back(std::string(pairs[i].pattern, (size_t)pairs[i].pattern_length),
                      std::string(pairs[i].text, (size_t)pairs[i].text_length));
  const double t0 = now_s();
  s->replayers.emplace_back([s, req, at, t0]() {
    for (size_t i = 0; i < req->size() && !s->stop.load(); ++i) {
      const double due = t0 + (*at)[i];
      double t = now_s();
      if (due > t + 2e-3) std::this_thread::sleep_for(std::chrono::duration<double>(due - t - 1e-3));
      while (now_s() < due) std::this_thread::yield();
      const auto& pr = (*req)[i];
      wfa_server_submit(s, pr.first.data(), (int)pr.first.size(), pr.second.data(), (int)pr.second.size());
    }
  });
  return WF_STATUS_ALG_COMPLETED;
}

extern "C" void wfa_server_destroy(wfa_server_t* s) {
  if (!s) return;
  for (auto& t : s->replayers) t.join();
  s->stop.store(true);
  for (auto& q : s->queues) {
    std::lock_guard<std::mutex> lk(q.mu);
    q.cv.notify_all();
  }
  for (auto& t : s->lanes) t.join();
  delete s;
}
