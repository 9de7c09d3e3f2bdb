// C-ABI host layer: the drop-in wavefront_aligner_new / wavefront_align /
// wavefront_aligner_delete entry points (wavefront/wavefront_align.c:212-241,
// wavefront/wavefront_aligner.c:422-504) plus wavefront_align_batch and the
// device-resident batch API, all driving the HIP kernels.
//
// One aligner owns one HIP stream and a grow-only device arena.  A batch is
// packed on the host (2 bits per base when every byte is A/C/G/T, else one
// byte per symbol), copied to HBM once, and aligned by:
//   1. forward kernels, one launch per kernel family (pairs sorted by length;
//      short pairs on the LDS-resident one-wavefront family, long pairs on the
//      workgroup family with the wavefront ring in HBM);
//   2. one backtrace kernel per family launch;
//   3. a retry pass, only for pairs that overflowed the LDS ring (moved to the
//      HBM-ring family) or the record arena (arena reset, smaller pass).
// Memory mode ultralow (BiWFA) is driven by wfa_bialign.cpp on the same jobs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>
#include <tuple>
#include <vector>

#include "wfa_internal.h"

namespace wfa_host {

using wfa::BacktraceArgs;
using wfa::ForwardArgs;
using wfa::LaunchShape;
using wfa::Penalties;

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  fprintf(stderr, "[WFA] HIP error in %s: %s\n", what, hipGetErrorString(e));
  return false;
}

thread_local bool t_keep_retired = false;

bool DevBuf::ensure(size_t bytes, size_t max_bytes) {
  if (bytes <= cap && p) return true;
  if (region) {
    p = region->take(bytes);
    cap = p ? bytes : 0;
    return p != nullptr;
  }
  // Geometric growth, clamped to max_bytes: a buffer used again and again
  // (serving lanes, repeated batches) reallocates rarely.  If the grown size
  // does not fit, the exact request is tried.
  size_t want = std::max<size_t>(std::max<size_t>(bytes, 256), p ? std::min(2 * cap, cap + (size_t(4) << 30)) : 0);
  want = std::max(bytes, std::min(want, max_bytes));
  if (p && t_keep_retired) retired.push_back(p);
  else if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  if (hipMalloc(&p, want) != hipSuccess) {
    p = nullptr;
    if (want == bytes || hipMalloc(&p, bytes) != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return false;
    }
    want = bytes;
  }
  cap = want;
  return true;
}

void DevBuf::release() {
  if (p && !region) (void)hipFree(p);
  for (void* q : retired) (void)hipFree(q);
  retired.clear();
  p = nullptr;
  cap = 0;
}

void PlanGroup::release() {
  d_ids.release();
  ids_on_device = false;
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  ev[0] = ev[1] = ev[2] = nullptr;
}

PlanGroup& PlanGroup::operator=(PlanGroup&& o) noexcept {
  if (this == &o) return *this;
  release();
  family = o.family;
  block = o.block;
  wcap = o.wcap;
  lmax_words = o.lmax_words;
  max_grid = o.max_grid;
  lclass = o.lclass;
  ids = std::move(o.ids);
  d_ids = std::move(o.d_ids);
  ids_on_device = o.ids_on_device;
  est_rec_bytes = o.est_rec_bytes;
  for (int j = 0; j < 3; ++j) {
    ev[j] = o.ev[j];
    o.ev[j] = nullptr;
  }
  o.ids_on_device = false;
  return *this;
}

bool HostBuf::ensure(size_t bytes) {
  if (bytes <= cap && p) return true;
  const size_t want = std::max<size_t>(std::max<size_t>(bytes, 4096), p ? 2 * cap : 0);
  if (p && t_keep_retired) retired.push_back(p);
  else if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
  if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
    p = nullptr;
    return false;
  }
  cap = want;
  return true;
}

void HostBuf::release() {
  if (p) (void)hipHostFree(p);
  for (void* q : retired) (void)hipHostFree(q);
  retired.clear();
  p = nullptr;
  cap = 0;
}

PlanGroup take_group(Engine* e) {
  PlanGroup g;
  if (e && e->na) {
    g.d_ids.region = &e->na->order;
  } else if (e && !e->group_pool.empty()) {
    g = std::move(e->group_pool.back());
    e->group_pool.pop_back();
  }
  g.family = 0;
  g.block = 64;
  g.wcap = 0;
  g.lmax_words = 0;
  g.max_grid = 0;
  g.ids.clear();
  g.ids_on_device = false;
  g.est_rec_bytes = -1;
  g.lclass = -1;
  return g;
}

void give_back_groups(Engine* e, std::vector<PlanGroup>& groups) {
  for (auto& g : groups) {
    g.ids.clear();
    g.ids_on_device = false;
    if (e && !g.d_ids.region) e->group_pool.push_back(std::move(g));
  }
  groups.clear();
}

void recycle_plan(Engine* e, std::vector<PlanGroup>& old_plan, std::vector<PlanGroup>& new_plan) {
  size_t i = 0;
  for (; i < old_plan.size() && i < new_plan.size(); ++i) {
    PlanGroup& o = old_plan[i];
    PlanGroup& n = new_plan[i];
    std::swap(o.d_ids, n.d_ids);
    for (int j = 0; j < 3; ++j) std::swap(o.ev[j], n.ev[j]);
    n.ids_on_device = false;
  }
  // groups beyond the old plan take pooled buffers and events
  for (; i < new_plan.size(); ++i) {
    if (!e || e->group_pool.empty()) break;
    PlanGroup& n = new_plan[i];
    PlanGroup& o = e->group_pool.back();
    std::swap(o.d_ids, n.d_ids);
    for (int j = 0; j < 3; ++j) std::swap(o.ev[j], n.ev[j]);
    n.ids_on_device = false;
    e->group_pool.pop_back();
  }
  give_back_groups(e, old_plan);
}

Engine::~Engine() {
  for (auto& ev : bi.ev)
    if (ev) (void)hipEventDestroy(ev);
  if (plain.exec) (void)hipGraphExecDestroy(plain.exec);
  if (plain.graph) (void)hipGraphDestroy(plain.graph);
}

Engine* engine_of(wavefront_aligner_t* wf) { return static_cast<Engine*>(wf->device); }
wfa::Counters* counters_of(Engine* e) { return e->counters.as<wfa::Counters>(); }

// Record blocks are addressed by 32-bit indices of 16-byte units.
constexpr size_t kMaxArenaBytes = 0xFFFFFFF0ull * 16;
// The arena for long pairs (HBM-ring family) is sized at L^2 bytes of
// records per pair (1-20 kbp pairs of 1-15% error average about L^2 / 4 and
// reach 0.7 L^2 at 15%; a sample that runs out of arena re-runs its last
// pairs almost alone, a launch at a few percent occupancy), and such a pair
// starts only while the arena has L^2 / 2 bytes left (L^2 >> 5 units of 16
// bytes).
constexpr double kLargeRecordShareOfL2 = 1.0;
// LDS for the two sequences of a pair of the HBM-ring family.
constexpr size_t kLargeSeqLdsBytes = 64 * 1024;
constexpr int kLargeDeferShift = 5;

bool engine_init(Engine* e) {
  if (e->ok) return true;
  if (!hip_ok(hipGetDevice(&e->device), "hipGetDevice")) return false;
  hipDeviceProp_t prop;
  if (!hip_ok(hipGetDeviceProperties(&prop, e->device), "hipGetDeviceProperties")) return false;
  e->num_cus = prop.multiProcessorCount;
  if (e->own_queue) {
    const int n = e->num_cus;
    int lo = std::max(0, std::min(n - 1, (int)(e->cu_lo * n)));
    int hi = std::max(lo + 1, std::min(n, (int)(e->cu_hi * n + 0.999)));
    std::vector<uint32_t> mask((n + 31) / 32, 0u);
    for (int c = lo; c < hi; ++c) mask[c / 32] |= 1u << (c % 32);
    if (!hip_ok(hipExtStreamCreateWithCUMask(&e->stream, (uint32_t)mask.size(), mask.data()), "hipStreamCreate"))
      return false;
    e->num_cus = hi - lo;
  } else if (!hip_ok(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "hipStreamCreate")) {
    return false;
  }
  if (!e->counters.ensure(64)) return false;
  // Record arena: up to half of the free HBM (block indices address 64 GB),
  // or WFA_RECORD_ARENA_MB; each run sizes it from its pairs' lengths.
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  size_t cap = std::min<size_t>(free_b / 2, kMaxArenaBytes);
  const char* env = getenv("WFA_RECORD_ARENA_MB");
  if (env) cap = std::min<size_t>(strtoull(env, nullptr, 10) << 20, kMaxArenaBytes);
  if (e->rec_cap) cap = std::min(cap, e->rec_cap);  // a cap set before the first run
  e->rec_cap = std::max<size_t>(cap, 64ull << 20);
  // The arena starts at 64 MB; each run grows it to its pairs' estimate.
  e->rec_bytes = std::min<size_t>(64ull << 20, e->rec_cap);
  e->ok = true;
  return true;
}

namespace {

void engine_free(Engine* e) {
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  delete e->scratch;
  e->scratch = nullptr;
  e->rec.release();
  e->ring.release();
  e->counters.release();
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

}  // namespace

namespace {
int gcd(int a, int b) {
  while (b) {
    const int t = a % b;
    a = b;
    b = t;
  }
  return a;
}
}  // namespace

// Device penalties and ring geometry; scopes as wavefront_components.c:45-124.
wfa::Penalties make_penalties(const wavefront_penalties_t& p) {
  wfa::Penalties r{};
  r.metric = (int)p.distance_metric;
  if (p.distance_metric <= gap_linear) {
    // One component (M): gaps are read from M at s - o1 (e1 = 0).
    r.ncomp = 1;
    r.x = p.distance_metric == indel ? 1 : p.mismatch;  // indel has no mismatch source
    r.o1 = p.distance_metric == gap_linear ? p.gap_opening1 : 1;
    r.e1 = 0;
    r.o2 = 0;
    r.e2 = 1;
    r.scope = p.distance_metric == gap_linear ? std::max(r.x, r.o1) + 1 : 2;
    r.g = std::max(gcd(r.x, r.o1), 1);
    r.nslots = (r.scope - 1) / r.g + 1;
    return r;
  }
  const bool two = p.distance_metric == gap_affine_2p;
  r.ncomp = two ? 5 : 3;
  r.x = p.mismatch;
  r.o1 = p.gap_opening1;
  r.e1 = p.gap_extension1;
  r.o2 = two ? p.gap_opening2 : 0;
  r.e2 = two ? p.gap_extension2 : 1;
  int indel_scope = r.o1 + r.e1;
  if (two) indel_scope = std::max(indel_scope, r.o2 + r.e2);
  r.scope = std::max(indel_scope, r.x) + 1;
  int g = gcd(r.x, gcd(r.o1 + r.e1, r.e1));
  if (two) g = gcd(g, gcd(r.o2 + r.e2, r.e2));
  r.g = std::max(g, 1);
  r.nslots = (r.scope - 1) / r.g + 1;
  r.r1 = r.e1 / r.g + 1;
  r.r2 = two ? r.e2 / r.g + 1 : 0;
  return r;
}

// wavefront_penalties_set_{indel,edit,linear,affine,affine2p}
// (wavefront_penalties.c:39-173).
void set_penalties(wavefront_penalties_t* wp, const wavefront_aligner_attr_t* attr) {
  memset(wp, 0, sizeof(*wp));
  wp->distance_metric = attr->distance_metric;
  if (attr->distance_metric == indel || attr->distance_metric == edit) {
    wp->match = 0;
    wp->mismatch = attr->distance_metric == indel ? -1 : 1;
    wp->gap_opening1 = 1;
    wp->gap_extension1 = -1;
    wp->gap_opening2 = -1;
    wp->gap_extension2 = -1;
  } else if (attr->distance_metric == gap_linear) {
    const linear_penalties_t& a = attr->linear_penalties;
    if (a.match > 0) {
      fprintf(stderr, "[WFA::Penalties] Match score must be negative or zero (M=%d)\n", a.match);
      exit(1);
    }
    if (a.mismatch <= 0 || a.indel <= 0) {
      fprintf(stderr, "[WFA::Penalties] Penalties (X=%d,D=%d,I=%d) must be (X>0,D>0,I>0)\n", a.mismatch, a.indel,
              a.indel);
      exit(1);
    }
    if (a.match < 0) {
      wp->match = a.match;
      wp->mismatch = 2 * a.mismatch - 2 * a.match;
      wp->gap_opening1 = 2 * a.indel - a.match;
    } else {
      wp->match = 0;
      wp->mismatch = a.mismatch;
      wp->gap_opening1 = a.indel;
    }
    wp->gap_extension1 = -1;
    wp->gap_opening2 = -1;
    wp->gap_extension2 = -1;
    wp->linear_penalties = a;
    wp->internal_gap_e = a.indel;
  } else if (attr->distance_metric == gap_affine_2p) {
    const affine2p_penalties_t& a = attr->affine2p_penalties;
    if (a.match > 0) {
      fprintf(stderr, "[WFA::Penalties] Match score must be negative or zero (M=%d)\n", a.match);
      exit(1);
    }
    if (a.mismatch <= 0 || a.gap_opening1 < 0 || a.gap_extension1 <= 0 || a.gap_opening2 < 0 ||
        a.gap_extension2 <= 0) {
      fprintf(stderr, "[WFA::Penalties] Penalties (X=%d,O1=%d,E1=%d,O2=%d,E2=%d) must be (X>0,O1>=0,E1>0,O2>=0,E2>0)\n",
              a.mismatch, a.gap_opening1, a.gap_extension1, a.gap_opening2, a.gap_extension2);
      exit(1);
    }
    if (a.match < 0) {
      wp->match = a.match;
      wp->mismatch = 2 * a.mismatch - 2 * a.match;
      wp->gap_opening1 = 2 * a.gap_opening1;
      wp->gap_extension1 = 2 * a.gap_extension1 - a.match;
      wp->gap_opening2 = 2 * a.gap_opening2;
      wp->gap_extension2 = 2 * a.gap_extension2 - a.match;
    } else {
      wp->match = 0;
      wp->mismatch = a.mismatch;
      wp->gap_opening1 = a.gap_opening1;
      wp->gap_extension1 = a.gap_extension1;
      wp->gap_opening2 = a.gap_opening2;
      wp->gap_extension2 = a.gap_extension2;
    }
    wp->affine2p_penalties = a;
    wp->internal_gap_e = a.gap_extension1;
  } else if (attr->distance_metric == gap_affine) {
    const affine_penalties_t& a = attr->affine_penalties;
    if (a.match > 0) {
      fprintf(stderr, "[WFA::Penalties] Match score must be negative or zero (M=%d)\n", a.match);
      exit(1);
    }
    if (a.mismatch <= 0 || a.gap_opening < 0 || a.gap_extension <= 0) {
      fprintf(stderr, "[WFA::Penalties] Penalties (X=%d,O=%d,E=%d) must be (X>0,O>=0,E>0)\n",
              a.mismatch, a.gap_opening, a.gap_extension);
      exit(1);
    }
    if (a.match < 0) {
      wp->match = a.match;
      wp->mismatch = 2 * a.mismatch - 2 * a.match;
      wp->gap_opening1 = 2 * a.gap_opening;
      wp->gap_extension1 = 2 * a.gap_extension - a.match;
    } else {
      wp->match = 0;
      wp->mismatch = a.mismatch;
      wp->gap_opening1 = a.gap_opening;
      wp->gap_extension1 = a.gap_extension;
    }
    wp->gap_opening2 = -1;
    wp->gap_extension2 = -1;
    wp->affine_penalties = a;
    wp->internal_gap_e = a.gap_extension;
  }
}

// wavefront_compute_classic_score (wavefront_compute.c:108-120).
int classic_score(const wavefront_penalties_t& p, int plen, int tlen, int s) {
  if (p.distance_metric <= edit) return s;
  const int swg_match = -p.match;
  if (swg_match == 0) return -s;
  return (swg_match * (plen + tlen) - s) / 2;
}


namespace {

// Configurations the exact device path covers; everything else is refused.
int check_supported(wavefront_aligner_t* wf) {
  const distance_metric_t m = wf->penalties.distance_metric;
  const alignment_form_t& f = wf->alignment_form;
  const bool endsfree = f.span == alignment_endsfree;
  const bool begin_free = endsfree && !f.extension && (f.pattern_begin_free > 0 || f.text_begin_free > 0);
  const char* why = nullptr;
  if (m < indel || m > gap_affine_2p) why = "an unknown distance metric";
  else if (endsfree && !f.extension &&
           (f.pattern_begin_free < 0 || f.pattern_end_free < 0 || f.text_begin_free < 0 || f.text_end_free < 0))
    why = "negative ends-free lengths";
  else if (endsfree && wf->memory_mode == wavefront_memory_ultralow)
    why = "the ends-free form in memory mode ultralow";
  // A free beginning with a match bonus starts wavefronts at later scores
  // (wavefront_compute.c:124-262), which the device does not model.
  else if (begin_free && wf->penalties.match != 0) why = "free beginnings with a match bonus (match < 0)";
  // Indel / edit with a free beginning: the reference's backtrace sees only
  // k = 0 of the score-0 wavefront (its initialised range stays [0,0],
  // wavefront.c:89-109) and writes 'M' over mismatching bases.
  else if (begin_free && m <= edit) why = "free beginnings with indel/edit distances";
  // Indel / edit extension in memory med/low leaves the score unset in the
  // reference (piggyback backtrace); only memory high is defined.
  else if (f.extension && m <= edit && wf->memory_mode != wavefront_memory_high)
    why = "indel/edit extension outside memory mode high";
  // The strategies take effect as wavefront_aligner_init_heuristic selects
  // them (wavefront_aligner.c:189-228): wf-adaptive before wfmash, x-drop
  // before z-drop, banded static before banded adaptive.
  // A z-dropped alignment in memory mode med / low traces the piggybacked
  // CIGAR of another cell (wavefront_unialign.c:177-189).
  else if ((wf->heuristic.strategy & wf_heuristic_zdrop) && !(wf->heuristic.strategy & wf_heuristic_xdrop) &&
           wf->alignment_scope == compute_alignment &&
           (wf->memory_mode == wavefront_memory_med || wf->memory_mode == wavefront_memory_low))
    why = "the z-drop heuristic with alignment in memory mode med or low";
  else if (make_penalties(wf->penalties).nslots > 64)
    why = "penalties whose score scope needs more than 64 wavefront slots";
  if (!why) return WF_STATUS_OK;
  if (wf->system.verbose >= 0)
    fprintf(stderr, "[WFA] Unsupported configuration on the device path: %s\n", why);
  return WF_STATUS_UNSUPPORTED;
}

// wavefront_align_presets__checks (wavefront_align.c:46-106): configuration
// errors the reference reports with exit(1), checked for every pair.
void presets_check(wavefront_aligner_t* wf, const int* plens, const int* tlens, int n) {
  const alignment_form_t& f = wf->alignment_form;
  if ((wf->heuristic.strategy & (wf_heuristic_xdrop | wf_heuristic_zdrop)) &&
      (wf->penalties.distance_metric == edit || wf->penalties.distance_metric == indel)) {
    fprintf(stderr, "[WFA] Heuristics drops are not compatible with 'edit'/'indel' distance metrics\n");
    exit(1);
  }
  for (int i = 0; i < n; ++i) {
    const int plen = plens[i], tlen = tlens[i];
    int pbf = f.pattern_begin_free, pef = f.pattern_end_free, tbf = f.text_begin_free, tef = f.text_end_free;
    if (f.span == alignment_endsfree && f.extension) { pbf = 0; pef = plen; tbf = 0; tef = tlen; }
    if (wf->memory_mode == wavefront_memory_ultralow) {
      if (pbf > 0 || pef > 0 || tbf > 0 || tef > 0) {
        fprintf(stderr, "[WFA] BiWFA ends-free has not been tested properly yet (let me know and I'll do it)\n");
        exit(1);
      }
      if (f.extension) {
        fprintf(stderr, "[WFA] BiWFA extension is not implemented yet (let me know and I'll add it)\n");
        exit(1);
      }
    }
    if (f.span == alignment_endsfree && (pbf > plen || pef > plen || tbf > tlen || tef > tlen)) {
      fprintf(stderr, "[WFA] Ends-free parameters must be not larger than the sequences "
              "(P0=%d,Pf=%d,T0=%d,Tf=%d). Must be (P0<=|P|,Pf<=|P|,T0<=|T|,Tf<=|T|) where (|P|,|T|)=(%d,%d)\n",
              pbf, pef, tbf, tef, plen, tlen);
      exit(1);
    }
  }
}

void presets_check(wavefront_aligner_t* wf, const wfa_pair_t* pairs, int n) {
  std::vector<int> pl(n), tl(n);
  for (int i = 0; i < n; ++i) { pl[i] = pairs[i].pattern_length; tl[i] = pairs[i].text_length; }
  presets_check(wf, pl.data(), tl.data(), n);
}

}  // namespace

wfa::HeurArgs heuristic_of(const wavefront_aligner_t* wf) {
  wfa::HeurArgs h{};
  const int st = (int)wf->heuristic.strategy;
  h.wfadaptive = (st & wf_heuristic_wfadaptive) ? 1 : (st & wf_heuristic_wfmash) ? 2 : 0;
  h.xdrop = (st & wf_heuristic_xdrop) ? 1 : 0;
  h.zdrop = (!h.xdrop && (st & wf_heuristic_zdrop)) ? 1 : 0;
  h.banded = (st & wf_heuristic_banded_static) ? 1 : (st & wf_heuristic_banded_adaptive) ? 2 : 0;
  if (!h.cut()) return h;
  h.min_len = wf->heuristic.min_wavefront_length;
  h.thr = wf->heuristic.max_distance_threshold;
  h.steps = wf->heuristic.steps_between_cutoffs;
  h.xdrop_thr = wf->heuristic.xdrop;
  h.zdrop_thr = wf->heuristic.zdrop;
  h.swg_match = wf->penalties.match != 0 ? -wf->penalties.match : -1;  // wavefront_heuristic.c:303-304
  h.band_lo = wf->heuristic.min_k;
  h.band_hi = wf->heuristic.max_k;
  return h;
}

namespace {

// Ends-free form of a run; extension frees both ends entirely (wavefront_align.c:57-63).
EndsFree ends_free_of(const wavefront_aligner_t* wf) {
  EndsFree e;
  const alignment_form_t& f = wf->alignment_form;
  if (f.span != alignment_endsfree) return e;
  e.on = 1;
  if (f.extension) {
    e.pef = e.tef = wfa::kFreeAll;
  } else {
    e.pbf = f.pattern_begin_free;
    e.pef = f.pattern_end_free;
    e.tbf = f.text_begin_free;
    e.tef = f.text_end_free;
  }
  return e;
}

// cigar_maxtrim_gap_{linear,affine,affine2p} (alignment/cigar_utils.c:111-278):
// the extension form keeps the prefix of the CIGAR with the best score, a
// match scoring -match (1 when match is 0).  `examined` is false where the
// reference leaves the CIGAR as it is (indel / edit, an empty affine2p CIGAR).
struct Trim {
  bool examined = false, trimmed = false;
  int keep = 0, score = INT32_MIN, end_v = -1, end_h = -1;
};
Trim maxtrim(const wavefront_penalties_t& p, const char* ops, int n) {
  Trim t;
  t.keep = n;
  const distance_metric_t m = p.distance_metric;
  if (m <= edit) return t;
  if (m == gap_affine_2p && n <= 0) return t;
  t.examined = true;
  int match = 0, mismatch = 0, o1 = 0, e1 = 0, o2 = 0, e2 = 0;
  if (m == gap_linear) {
    match = p.linear_penalties.match; mismatch = p.linear_penalties.mismatch; o1 = p.linear_penalties.indel;
  } else if (m == gap_affine) {
    match = p.affine_penalties.match; mismatch = p.affine_penalties.mismatch;
    o1 = p.affine_penalties.gap_opening; e1 = p.affine_penalties.gap_extension;
  } else {
    const affine2p_penalties_t& a = p.affine2p_penalties;
    match = a.match; mismatch = a.mismatch; o1 = a.gap_opening1; e1 = a.gap_extension1;
    o2 = a.gap_opening2; e2 = a.gap_extension2;
  }
  const int mscore = match != 0 ? match : -1;
  int best = 0, best_i = 0, best_v = 0, best_h = 0, score = 0, v = 0, h = 0;
  if (m == gap_affine_2p) {
    // scored per run of equal operations, compared at run ends
    int i = 0;
    while (i < n) {
      int j = i;
      while (j < n && ops[j] == ops[i]) ++j;
      const int L = j - i;
      switch (ops[i]) {
        case 'M': score -= mscore * L; v += L; h += L; break;
        case 'X': score -= mismatch * L; v += L; h += L; break;
        case 'D': score -= std::min(o1 + e1 * L, o2 + e2 * L); v += L; break;
        default: score -= std::min(o1 + e1 * L, o2 + e2 * L); h += L; break;
      }
      if (best < score) { best = score; best_i = j - 1; best_v = v; best_h = h; }
      i = j;
    }
  } else {
    char last = 0;
    for (int i = 0; i < n; ++i) {
      const char op = ops[i];
      switch (op) {
        case 'M': score -= mscore; ++v; ++h; break;
        case 'X': score -= mismatch; ++v; ++h; break;
        default: {
          const int cost = m == gap_linear ? o1 : e1 + (last == op ? 0 : o1);
          score -= cost;
          if (op == 'D') ++v; else ++h;
        }
      }
      last = op;
      if (best < score) { best = score; best_i = i; best_v = v; best_h = h; }
    }
  }
  t.trimmed = best_i != n - 1;
  if (best == 0) {
    t.keep = 0;  // cigar_clear: score INT32_MIN, end position (-1, -1)
  } else {
    t.keep = best_i + 1;
    t.score = best;
    t.end_v = best_v;
    t.end_h = best_h;
  }
  return t;
}

bool all_acgt(const wfa_pair_t* pairs, int n) {
  // built once, thread-safely (serving lanes call this concurrently)
  static const struct Lut {
    bool v[256];
    Lut() : v() { v['A'] = v['C'] = v['G'] = v['T'] = true; }
  } lut_table;
  const bool* lut = lut_table.v;
  auto scan = [&](int i0, int i1, const std::atomic<bool>* stop) {
    for (int i = i0; i < i1; ++i) {
      if (stop && (i & 1023) == 0 && stop->load(std::memory_order_relaxed)) return false;
      const unsigned char* p = reinterpret_cast<const unsigned char*>(pairs[i].pattern);
      for (int j = 0; j < pairs[i].pattern_length; ++j) if (!lut[p[j]]) return false;
      const unsigned char* t = reinterpret_cast<const unsigned char*>(pairs[i].text);
      for (int j = 0; j < pairs[i].text_length; ++j) if (!lut[t[j]]) return false;
    }
    return true;
  };
  int64_t symbols = 0;
  for (int i = 0; i < n; ++i) symbols += (int64_t)pairs[i].pattern_length + pairs[i].text_length;
  // large batches are scanned by up to 16 threads, one range of pairs each
  const int nt = (int)std::min<int64_t>({16, (int64_t)std::max(1u, std::thread::hardware_concurrency()),
                                         symbols / (int64_t(1) << 22) + 1, (int64_t)std::max(n, 1)});
  if (nt <= 1) return scan(0, n, nullptr);
  std::atomic<bool> bad(false);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      if (!scan((int)((int64_t)n * t / nt), (int)((int64_t)n * (t + 1) / nt), &bad)) bad.store(true);
    });
  for (auto& x : th) x.join();
  return !bad.load();
}

inline int64_t nwords(int len, int bits) { return len / (64 / bits) + 2; }

void pack_seq(uint64_t* dst, const char* src, int len, int bits) {
  const int64_t nw = nwords(len, bits);
  memset(dst, 0, nw * sizeof(uint64_t));
  if (bits == 8) {
    memcpy(dst, src, len);
    return;
  }
  // A,C,G,T -> 0,1,3,2 ((c >> 1) & 3); any bijection works for equality tests.
  for (int i = 0; i < len; ++i) {
    const uint64_t code = (uint64_t)((static_cast<unsigned char>(src[i]) >> 1) & 3);
    dst[i >> 5] |= code << (2 * (i & 31));
  }
}

// LDS ring budget per workgroup of the one-wavefront family (16-bit ring
// elements): ~8.5 KB with the sequences keeps ~18 wavefronts per CU resident.
constexpr int kSmallLdsBudget = 8 * 1024;

// Ring width of the generic LDS family: the width the budget allows, and
// when that is under 256 diagonals, up to twice the budget for a ring of at
// most 256 (measured on 100-300 bp reads, gap-affine (0,5,7,3), ten slots:
// 11.1 M pairs/s at 128 diagonals, 15.3 M at 256, 9.0 M at 512: wider rings
// re-run fewer pairs, larger ones cost occupancy).  WFA_SMALL_LDS (bytes)
// replaces the budget for tuning.
int small_wcap(const Penalties& pen, int ncomp) {
  static const int budget = [] {
    const char* v = getenv("WFA_SMALL_LDS");
    return v && atoi(v) > 0 ? atoi(v) : kSmallLdsBudget;
  }();
  (void)ncomp;
  const int per = wfa::ring_slots(pen) * 2;  // bytes per diagonal
  int w = budget / per;
  if (w < 256) w = std::min(256, 2 * budget / per);
  w = (w / 32) * 32;
  return w;
}

// Lean family (gap-affine, 2-bit sequences up to kLeanMaxLen): ring width per
// slot for pairs up to short_len and for longer pairs, rounded up to a width
// with a kernel instance (224, 384, 768).  A pair whose wavefront outgrows its
// ring is re-run with the next width, then on the HBM-ring family.
// WFA_LEAN=0 turns the family off; WFA_LEAN_WCAP="short,long[,short_len]"
// overrides the widths.
constexpr int kLeanMaxLen = 1024;
struct LeanTuning {
  bool enabled = true;
  int wcap_short = 224;
  int wcap_long = 384;
  int short_len = 320;
};
const LeanTuning& lean_tuning() {
  static const LeanTuning t = [] {
    LeanTuning r;
    const char* on = getenv("WFA_LEAN");
    if (on && on[0] == '0') r.enabled = false;
    const char* w = getenv("WFA_LEAN_WCAP");
    int a = 0, b = 0, c = 0;
    const int n = w ? sscanf(w, "%d,%d,%d", &a, &b, &c) : 0;
    if (n >= 2 && wfa::lean_width(a) > 0 && wfa::lean_width(b) > 0) { r.wcap_short = a; r.wcap_long = b; }
    if (n == 3 && c > 0) r.short_len = c;
    return r;
  }();
  return t;
}

}  // namespace

// ------------------------------------------------------------ planning

namespace {

PlanGroup make_large(Engine* e, const JobSet& js, std::vector<int> ids) {
  PlanGroup g = take_group(e);
  g.family = wfa::kLarge;
  int pmax = 0, tmax = 0;
  for (int i : ids) { pmax = std::max(pmax, js.h_plen[i]); tmax = std::max(tmax, js.h_tlen[i]); }
  g.wcap = pmax + tmax + 8;
  // Sequences in LDS when both fit in 64 KB (2-bit sequences up to ~130 kbp).
  const int spw = 64 / js.bits;
  const int lw = std::max(pmax, tmax) / spw + 3;
  g.lmax_words = 2 * (size_t)lw * sizeof(uint64_t) <= kLargeSeqLdsBytes ? lw : 0;
  g.block = wfa::pick_block(wfa::kLarge, pmax + tmax);
  g.ids = std::move(ids);
  return g;
}

LaunchShape shape_of(Engine* e, const PlanGroup& g, int bits, const Penalties& pen, int cut) {
  LaunchShape shape;
  memset(&shape, 0, sizeof(shape));  // padding too: shapes are compared as bytes (graph keys)
  shape.family = g.family;
  shape.block = g.block;
  int per_cu = 0;
  if (g.family == wfa::kLean) {
    shape.lds_bytes = wfa::lean_lds_bytes(pen, g.wcap, g.lmax_words);
    per_cu = wfa::lean_max_blocks_per_cu(pen, g.wcap, shape.lds_bytes);
  } else {
    shape.lds_bytes = wfa::forward_lds_bytes(g.family, wfa::ring_slots(pen), g.wcap, g.lmax_words);
    per_cu = wfa::forward_max_blocks_per_cu(g.family, g.block, pen.ncomp, bits, cut, shape.lds_bytes);
  }
  if (per_cu <= 0) per_cu = 1;
  shape.grid = (int)std::min<int64_t>((int64_t)g.ids.size(), (int64_t)per_cu * e->num_cus);
  if (g.max_grid > 0) shape.grid = std::min(shape.grid, g.max_grid);
  if (e->na && g.family == wfa::kLarge) {
    // no-allocation run: as many workgroups as rings fit in its ring share
    const size_t per_block = (size_t)wfa::ring_slots(pen) * g.wcap * sizeof(int);
    shape.grid = (int)std::min<size_t>((size_t)shape.grid, e->na->ring_bytes / per_block);
  }
  if (e->na) {
    // ... and at most half the record arena in record chunks (one per
    // resident workgroup), so that most workgroups get one
    const size_t chunks = e->rec.cap / ((size_t)wfa::kChunkUnits * 16) / 2;
    shape.grid = (int)std::max<size_t>(1, std::min<size_t>((size_t)shape.grid, chunks));
  }
  return shape;
}

// The kernel arguments of one group's forward and backtrace launches.
// A group that takes jobs 0..n-1 in order passes no order array.
void group_args(wavefront_aligner_t* wf, Engine* e, const JobSet& js, const PlanGroup& grp, const Penalties& pen,
                int records, bool ident, ForwardArgs* pfa, BacktraceArgs* pba) {
  const int n = (int)grp.ids.size();
  wfa::Counters* c = counters_of(e);
  ForwardArgs fa{};
  fa.n_pairs = n;
  fa.order = ident ? nullptr : grp.d_ids.as<int>();
  fa.plen = js.plen;
  fa.tlen = js.tlen;
  fa.words = js.words;
  fa.pat_pos = js.ppos;
  fa.txt_pos = js.tpos;
  fa.comp_begin_p = js.cbeg;
  fa.comp_end_p = js.cend;
  fa.out_status = js.status;
  fa.out_score = js.score;
  fa.out_dsteps = js.dsteps;
  fa.out_last_block = js.last;
  fa.rec_arena = e->rec.as<uint4>();
  fa.rec_units = records ? std::min<unsigned long long>(e->rec.cap / 16, 0xFFFFFFF0ull) : 0;
  fa.rec_top = &c->rec_top;
  fa.ring_global = grp.family == wfa::kLarge ? e->ring.as<int>() : nullptr;
  fa.ring_stride = grp.family == wfa::kLarge ? (long long)wfa::ring_slots(pen) * grp.wcap : 0;
  fa.queue = &c->queue;
  fa.dsteps_total = &c->dsteps;
  fa.retry_count = &c->retry;
  fa.fail_count = &c->fail;
  fa.wcap = grp.wcap;
  fa.lmax_words = grp.lmax_words;
  fa.max_steps = wf->system.max_alignment_steps > 0 ? wf->system.max_alignment_steps : INT_MAX;
  fa.records = records;
  fa.comp_begin = wfa::kM;
  fa.comp_end = wfa::kM;
  fa.defer_shift = kLargeDeferShift;  // L^2 >> shift units of 16 B: L^2/2 bytes of records
  fa.endsfree = js.ef.on;
  fa.pbf = js.ef.pbf;
  fa.pef = js.ef.pef;
  fa.tbf = js.ef.tbf;
  fa.tef = js.ef.tef;
  fa.out_end_k = js.end_k;
  fa.out_end_off = js.end_off;
  fa.out_end_rec = js.end_rec;
  fa.heur = js.heur;
  fa.pen = pen;
  *pfa = fa;
  BacktraceArgs ba{};
  ba.n_pairs = n;
  ba.order = fa.order;
  ba.plen = js.plen;
  ba.tlen = js.tlen;
  ba.status = js.status;
  ba.score = js.score;
  ba.last_block = js.last;
  ba.rec_arena = e->rec.as<uint4>();
  ba.rec_units = fa.rec_units;
  ba.ops = js.ops;
  ba.ops_off = js.ops_off;
  ba.ops_begin = js.ops_begin;
  ba.fail_count = &c->fail;
  ba.comp_end_p = js.cend;
  ba.comp_begin_p = js.cbeg;
  ba.comp_begin = wfa::kM;
  ba.comp_end = wfa::kM;
  ba.words = js.words;
  ba.pat_pos = js.ppos;
  ba.txt_pos = js.tpos;
  ba.bits = js.bits;
  ba.endsfree = js.ef.on;
  ba.end_k = js.end_k;
  ba.end_off = js.end_off;
  ba.end_rec = js.end_rec;
  ba.pen = pen;
  *pba = ba;
}

// Forward (and, with records, backtrace) launches of one group, with the
// group's events around them when it has events.
hipError_t launch_group(const PlanGroup& grp, const LaunchShape& shape, const ForwardArgs& fa,
                        const BacktraceArgs& ba, int ncomp, int bits, int records, hipStream_t st) {
  if (grp.ev[0]) (void)hipEventRecord(grp.ev[0], st);
  hipError_t err = grp.family == wfa::kLean ? wfa::launch_forward_lean(fa, shape, st)
                                            : wfa::launch_forward(fa, ncomp, bits, shape, st);
  if (err != hipSuccess) return err;
  if (grp.ev[1]) (void)hipEventRecord(grp.ev[1], st);
  if (records) {
    err = wfa::launch_backtrace(ba, st);
    if (err != hipSuccess) return err;
  }
  if (grp.ev[2]) (void)hipEventRecord(grp.ev[2], st);
  return hipSuccess;
}

// Enqueues forward + backtrace of one group on the aligner's stream; with
// fresh, the run counters were just cleared and need no reset of their own.
int enqueue_group(wavefront_aligner_t* wf, Engine* e, const JobSet& js, PlanGroup& grp, const Penalties& pen,
                  int records, bool fresh) {
  const int n = (int)grp.ids.size();
  if (n == 0) return WF_STATUS_OK;
  hipStream_t st = e->stream;
  const LaunchShape shape = shape_of(e, grp, js.bits, pen, js.heur.cut());
  bool ident = true;
  for (int i = 0; i < n && ident; ++i) ident = grp.ids[i] == i;
  if (!ident && !grp.ids_on_device) {
    if (!grp.d_ids.ensure((size_t)n * 4)) {
      fprintf(stderr, "[WFA] cannot allocate the launch order of %d pairs\n", n);
      return WF_STATUS_OOM;
    }
    if (!hip_ok(hipMemcpyAsync(grp.d_ids.p, grp.ids.data(), (size_t)n * 4, hipMemcpyHostToDevice, st), "plan upload"))
      return WF_STATUS_OOM;
    grp.ids_on_device = true;
  }
  if (!e->na)
    for (auto& ev : grp.ev)
      if (!ev && !hip_ok(hipEventCreate(&ev), "hipEventCreate")) return WF_STATUS_OOM;
  ForwardArgs fa;
  BacktraceArgs ba;
  group_args(wf, e, js, grp, pen, records, ident, &fa, &ba);
  if (!fresh && !hip_ok(wfa::launch_reset_counters(counters_of(e), 0, st), "reset")) return WF_STATUS_OOM;
  if (!hip_ok(launch_group(grp, shape, fa, ba, pen.ncomp, js.bits, records, st), "forward/backtrace launch"))
    return WF_STATUS_OOM;
  return WF_STATUS_OK;
}

// Device memory every group of the plan needs, allocated before any launch.
bool reserve_for(Engine* e, std::vector<PlanGroup>& plan, int bits, const Penalties& pen, int cut, int records) {
  if (records && !e->rec.ensure(e->rec_bytes, e->rec_cap)) {
    fprintf(stderr, "[WFA] cannot allocate the %zu MB record arena\n", e->rec_bytes >> 20);
    return false;
  }
  size_t ring = 0;
  for (const PlanGroup& g : plan) {
    if (g.family != wfa::kLarge) continue;
    const LaunchShape sh = shape_of(e, g, bits, pen, cut);
    ring = std::max(ring, (size_t)wfa::ring_slots(pen) * g.wcap * sh.grid * sizeof(int));
  }
  if (ring && !e->ring.ensure(ring)) {
    fprintf(stderr, "[WFA] cannot allocate %zu MB of wavefront rings\n", ring >> 20);
    return false;
  }
  return true;
}

}  // namespace

// Splits the jobs over kernel families: short jobs on the LDS family, the
// rest on the HBM-ring family by workgroup size.
std::vector<PlanGroup> make_plan(const JobSet& js, const std::vector<int>& ids, const Penalties& pen) {
  const int ncomp = pen.ncomp;
  const int wcap_small = small_wcap(pen, ncomp);
  const int spw = 64 / js.bits;
  std::vector<int> small_ids, cls[3];
  int lmax_small = 0;
  const LeanTuning& lt = lean_tuning();
  // Ends-free pairs take the wider ring in both length classes (their
  // wavefronts grow over the free flanks: measured on 100-300 bp reads with
  // flanks of up to 30 bases, the narrow ring re-ran enough pairs to cost
  // 20%); the score-0 wavefront spans pbf + tbf + 1 diagonals and must fit.
  const bool ef_fits = !js.ef.on || (int64_t)js.ef.pbf + js.ef.tbf + 3 <= wfa::lean_width(lt.wcap_long);
  // of the heuristics, the lean family runs all but the drops
  const bool heur_ok = js.heur.cut() <= 1;
  const bool lean_ok = lt.enabled && ncomp == 3 && js.bits == 2 && !js.cbeg && !js.cend && ef_fits && heur_ok &&
                       wfa::lean_supported(pen);
  std::vector<int> lean_ids[2];
  int lean_lmax[2] = {0, 0};
  for (int i : ids) {
    const int L = std::max(js.h_plen[i], js.h_tlen[i]);
    if (lean_ok && L <= kLeanMaxLen) {
      const int c = L <= lt.short_len ? 0 : 1;
      lean_ids[c].push_back(i);
      lean_lmax[c] = std::max<int>(lean_lmax[c], L / spw + 3);
      continue;
    }
    if (wcap_small >= 64 && L <= 1024) {
      small_ids.push_back(i);
      lmax_small = std::max<int>(lmax_small, L / spw + 3);
    } else {
      const int blk = wfa::pick_block(wfa::kLarge, js.h_plen[i] + js.h_tlen[i]);
      cls[blk == 64 ? 0 : (blk == 256 ? 1 : 2)].push_back(i);
    }
  }
  std::vector<PlanGroup> plan;
  for (int c = 0; c < 2; ++c) {
    if (lean_ids[c].empty()) continue;
    PlanGroup g;
    g.family = wfa::kLean;
    g.block = 64;
    g.wcap = wfa::lean_width(c == 0 && !js.ef.on ? lt.wcap_short : lt.wcap_long);
    g.lmax_words = lean_lmax[c];
    g.lclass = c;
    g.ids = std::move(lean_ids[c]);
    plan.push_back(std::move(g));
  }
  if (!small_ids.empty()) {
    PlanGroup g;
    g.family = wfa::kSmall;
    g.block = 64;
    g.wcap = wcap_small;
    g.lmax_words = lmax_small;
    g.ids = std::move(small_ids);
    plan.push_back(std::move(g));
  }
  for (auto& c : cls)
    if (!c.empty()) plan.push_back(make_large(nullptr, js, std::move(c)));
  return plan;
}

int run_jobs(wavefront_aligner_t* wf, Engine* e, const JobSet& js, std::vector<PlanGroup>& plan, int records,
             RunTotals* tot, std::vector<int>& h_status) {
  const Penalties pen = make_penalties(wf->penalties);
  hipStream_t st = e->stream;
  wfa::Counters host_c{};
  std::vector<PlanGroup> retry_groups;
  std::vector<PlanGroup>* pending = &plan;
  if (!hip_ok(wfa::launch_reset_counters(counters_of(e), 2, st), "reset")) return WF_STATUS_OOM;
  if (records) {
    // Rough footprint: one byte per diagonal-step plus 16 per score step.
    // Diagonal-steps grow as L^2 e^2 (e the error rate): about L^2/16 for
    // short reads of up to 10% error; long pairs take L^2 (see
    // kLargeRecordShareOfL2); plus a 64 KB chunk per resident workgroup.
    double want = 0;
    for (PlanGroup& g : plan) {
      if (g.est_rec_bytes < 0) {
        const double per_l2 = g.family == wfa::kLarge ? 1.0 / kLargeRecordShareOfL2 : 1.0 / 16;
        double w = 0;
        for (int i : g.ids) {
          const double L = std::max(js.h_plen[i], js.h_tlen[i]);
          w += L * L * per_l2 + 8 * L;
        }
        g.est_rec_bytes = w + (double)shape_of(e, g, js.bits, pen, js.heur.cut()).grid * wfa::kChunkUnits * 16;
      }
      want += g.est_rec_bytes;
    }
    const size_t w = (size_t)std::min<double>(want, (double)e->rec_cap);
    if (w > e->rec_bytes) e->rec_bytes = ((w >> 26) + 1) << 26;  // 64 MB steps
    e->rec_bytes = std::min(e->rec_bytes, e->rec_cap);
  }
  for (int round = 0;; ++round) {
    if (round > 256) {
      fprintf(stderr, "[WFA] batch did not converge (record arena too small?)\n");
      return WF_STATUS_OOM;
    }
    // no-allocation run: the order arrays of the groups run before are done with
    if (e->na && round > 0) e->na->order.used = 0;
    if (!reserve_for(e, *pending, js.bits, pen, js.heur.cut(), records)) return WF_STATUS_OOM;
    // the reset before the first round clears every counter
    if (round > 0 && !hip_ok(wfa::launch_reset_counters(counters_of(e), 1, st), "reset")) return WF_STATUS_OOM;
    bool fresh = true;
    for (PlanGroup& g : *pending) {
      const int rc = enqueue_group(wf, e, js, g, pen, records, fresh);
      if (rc != WF_STATUS_OK) return rc;
      if (!g.ids.empty()) fresh = false;
      ++tot->launches;
    }
    if (!hip_ok(hipMemcpyAsync(&host_c, e->counters.p, sizeof(host_c), hipMemcpyDeviceToHost, st), "counters") ||
        !hip_ok(hipStreamSynchronize(st), "alignment"))
      return WF_STATUS_OOM;
    for (PlanGroup& g : *pending) {
      float f = 0, t = 0;
      if (g.ids.empty() || !g.ev[0]) continue;
      (void)hipEventElapsedTime(&f, g.ev[0], g.ev[1]);
      (void)hipEventElapsedTime(&t, g.ev[1], g.ev[2]);
      tot->forward_ms += f;
      tot->backtrace_ms += t;
    }
    if (host_c.retry == 0) break;
    // Some jobs overflowed the LDS ring or the record arena: re-run only them.
    h_status.resize(js.n);
    // copies on the aligner's stream, never the null stream (which would wait
    // for the streams of other aligners)
    if (!hip_ok(hipMemcpyAsync(h_status.data(), js.status, (size_t)js.n * 4, hipMemcpyDeviceToHost, st), "status") ||
        !hip_ok(hipStreamSynchronize(st), "status"))
      return WF_STATUS_OOM;
    std::vector<PlanGroup> next;
    for (PlanGroup& g : *pending) {
      std::vector<int> ring_over, arena_full;
      for (int i : g.ids) {
        if (h_status[i] == wfa::kDevRingOverflow) ring_over.push_back(i);
        else if (h_status[i] == wfa::kDevArenaFull) arena_full.push_back(i);
      }
      if (!ring_over.empty()) {
        if (g.family == wfa::kLean && wfa::lean_width(g.wcap + 1) > 0) {
          PlanGroup h = take_group(e);
          h.family = wfa::kLean;
          h.block = 64;
          h.wcap = wfa::lean_width(g.wcap + 1);
          h.lmax_words = g.lmax_words;
          h.ids = std::move(ring_over);
          next.push_back(std::move(h));
        } else {
          next.push_back(make_large(e, js, std::move(ring_over)));
        }
      }
      if (!arena_full.empty()) {
        int max_grid = g.max_grid;
        if (arena_full.size() == g.ids.size()) {
          // not a single job of the group fitted: grow the arena; when it
          // cannot grow, run the jobs one workgroup at a time (jobs that ran
          // at once can fill the arena together although each fits alone)
          size_t free_b = 0, total_b = 0;
          (void)hipMemGetInfo(&free_b, &total_b);
          if (e->rec_bytes * 2 > std::min(e->rec_cap, e->rec.cap + free_b / 2) && g.max_grid != 1 &&
              g.ids.size() > 1) {
            max_grid = 1;
          } else if (e->rec_bytes * 2 > std::min(e->rec_cap, e->rec.cap + free_b / 2)) {
            fprintf(stderr, "[WFA] record arena exhausted (%zu MB)\n", e->rec_bytes >> 20);
            for (int i : arena_full) h_status[i] = WF_STATUS_OOM;
            if (!hip_ok(hipMemcpyAsync(js.status, h_status.data(), (size_t)js.n * 4, hipMemcpyHostToDevice, st),
                        "status") ||
                !hip_ok(hipStreamSynchronize(st), "status"))
              return WF_STATUS_OOM;
            host_c.fail += (unsigned)arena_full.size();
            arena_full.clear();
          } else {
            e->rec.release();
            e->rec_bytes *= 2;
          }
        }
        if (!arena_full.empty()) {
          PlanGroup h = take_group(e);
          h.family = g.family;
          h.block = g.block;
          h.wcap = g.wcap;
          h.lmax_words = g.lmax_words;
          h.max_grid = max_grid;
          h.ids = std::move(arena_full);
          next.push_back(std::move(h));
        }
      }
    }
    give_back_groups(e, retry_groups);
    retry_groups = std::move(next);
    pending = &retry_groups;
    if (retry_groups.empty()) break;
  }
  give_back_groups(e, retry_groups);
  tot->dsteps += host_c.dsteps;
  tot->fail += host_c.fail;
  return WF_STATUS_OK;
}

bool download_results(Engine* e, wfa_device_batch_t* b, char* cigar_dst) {
  hipStream_t st = e->stream;
  const size_t n = (size_t)b->n, nn = std::max<size_t>(n, 1);
  const bool res = !b->host_results && n > 0;
  // status, score, ops_begin, and end_k, end_off when the end cell varies
  const bool ends = (b->ef.on || b->heur.zdrop) && !b->bi_mode;
  const size_t bytes = nn * 4 * (ends ? 5 : 3);
  bool ok = true;
  if (res) {
    ok = b->h_res.ensure(bytes) &&
         hip_ok(hipMemcpyAsync(b->h_res.p, b->d_res.p, bytes, hipMemcpyDeviceToHost, st), "download");
  }
  const bool ops_copy = cigar_dst && b->total_ops > 0 && !b->h_ops;
  if (cigar_dst && b->total_ops > 0 && b->h_ops) memcpy(cigar_dst, b->h_ops, (size_t)b->total_ops);
  if (ok && ops_copy)
    ok = hip_ok(hipMemcpyAsync(cigar_dst, b->d_ops.p, b->total_ops, hipMemcpyDeviceToHost, st), "cigar download");
  if (!res && !ops_copy) return ok;
  ok = ok && hip_ok(hipStreamSynchronize(st), "download sync");
  if (ok && !b->host_results) {
    b->status.resize(n);
    b->score.resize(n);
    b->ops_begin.resize(n);
    b->end_k.resize(n);
    b->end_off.resize(n);
    if (n > 0) {
      const char* h = b->h_res.as<char>();
      memcpy(b->status.data(), h, n * 4);
      memcpy(b->score.data(), h + nn * 4, n * 4);
      memcpy(b->ops_begin.data(), h + 2 * nn * 4, n * 4);
      if (ends) {
        memcpy(b->end_k.data(), h + 3 * nn * 4, n * 4);
        memcpy(b->end_off.data(), h + 4 * nn * 4, n * 4);
      }
    }
    b->host_results = true;
  }
  return ok;
}

}  // namespace wfa_host

using namespace wfa_host;

wfa_host::JobSet wfa_device_batch_t::jobs() const {
  JobSet js;
  js.n = n;
  js.bits = bits;
  js.h_plen = plen.data();
  js.h_tlen = tlen.data();
  js.words = words.as<uint64_t>();
  js.plen = d_plen.as<int>();
  js.tlen = d_tlen.as<int>();
  js.ppos = d_patw.as<int64_t>();
  js.tpos = d_txtw.as<int64_t>();
  js.status = d_status.as<int>();
  js.score = d_score.as<int>();
  js.last = d_last.as<uint32_t>();
  js.dsteps = d_dsteps.as<long long>();
  js.ops = d_ops.as<char>();
  js.ops_off = d_ops_off.as<int64_t>();
  js.ops_begin = d_ops_begin.as<int>();
  js.end_k = d_end_k.as<int>();
  js.end_off = d_end_off.as<int>();
  js.end_rec = d_end_rec.as<int>();
  js.ef = ef;
  js.heur = heur;
  return js;
}

wfa_device_batch_t::~wfa_device_batch_t() {
  h_words.release();
  h_res.release();
  for (auto& g : plan) g.release();
  for (wfa_host::DevBuf* x : {&words, &d_res, &d_last, &d_dsteps, &d_ops, &d_end_rec, &rev_words, &d_rev_ppos,
                              &d_rev_tpos})
    x->release();
}

// ------------------------------------------------------------ lifecycle

extern "C" wavefront_aligner_t* wavefront_aligner_new(wavefront_aligner_attr_t* attributes) {
  if (attributes == nullptr) attributes = &wavefront_aligner_attr_default;
  wavefront_aligner_t* wf = static_cast<wavefront_aligner_t*>(calloc(1, sizeof(wavefront_aligner_t)));
  const bool bi = attributes->memory_mode == wavefront_memory_ultralow;
  wf->align_mode = bi ? wf_align_biwfa : wf_align_regular;
  wf->align_status.status = WF_STATUS_OK;
  wf->alignment_scope = attributes->alignment_scope;
  wf->alignment_form = attributes->alignment_form;
  set_penalties(&wf->penalties, attributes);
  wf->heuristic = attributes->heuristic;
  if (wf->heuristic.strategy == wf_heuristic_none) memset(&wf->heuristic, 0, sizeof(wf->heuristic));
  wf->memory_mode = attributes->memory_mode;
  wf->component_begin = affine2p_matrix_M;
  wf->component_end = affine2p_matrix_M;
  wf->system = attributes->system;
  wf->cigar = static_cast<cigar_t*>(calloc(1, sizeof(cigar_t)));
  wf->cigar->max_operations = 2 * (1000 + 1000);
  wf->cigar->operations = static_cast<char*>(malloc(wf->cigar->max_operations));
  wf->cigar->cigar_buffer = static_cast<uint32_t*>(malloc(sizeof(uint32_t) * wf->cigar->max_operations));
  wf->cigar->score = INT32_MIN;
  wf->cigar->end_v = wf->cigar->end_h = -1;
  wf->device = new (std::nothrow) Engine();
  return wf;
}

extern "C" void wavefront_aligner_reap(wavefront_aligner_t* const wf) {
  Engine* e = engine_of(wf);
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  delete e->scratch;
  e->scratch = nullptr;
  e->rec.release();
  e->ring.release();
}

extern "C" void wavefront_aligner_delete(wavefront_aligner_t* const wf) {
  if (!wf) return;
  engine_free(engine_of(wf));
  free(wf->sequences.pattern);
  free(wf->sequences.text);
  if (wf->cigar) {
    free(wf->cigar->operations);
    free(wf->cigar->cigar_buffer);
    free(wf->cigar);
  }
  free(wf);
}

// ------------------------------------------------------------ batches

extern "C" size_t wavefront_align_batch_arena_bytes(const wfa_pair_t* pairs, int num_pairs) {
  size_t b = 0;
  for (int i = 0; i < num_pairs; ++i) b += (size_t)pairs[i].pattern_length + (size_t)pairs[i].text_length;
  return b;
}

bool wfa_device_batch_t::load(Engine* e, const wfa_pair_t* pairs, int count, bool cigars, bool bucketed,
                              size_t* upload_bytes) {
  h_ops = nullptr;
  runs_ready = false;
  n = count;
  bits = all_acgt(pairs, n) ? 2 : 8;
  plen.resize(n);
  tlen.resize(n);
  ops_off.resize(n);
  pat_pos.resize(n);
  txt_pos.resize(n);
  std::vector<int64_t> patw(n), txtw(n);  // word index of each sequence
  int64_t nw = 0, ops = 0;
  for (int i = 0; i < n; ++i) {
    plen[i] = pairs[i].pattern_length;
    tlen[i] = pairs[i].text_length;
    patw[i] = nw;
    nw += nwords(plen[i], bits);
    txtw[i] = nw;
    nw += nwords(tlen[i], bits);
    ops_off[i] = ops;
    ops += (int64_t)plen[i] + tlen[i];
  }
  total_ops = ops;
  words_used = nw;
  // per-pair inputs after the words: pat_pos, txt_pos, ops_off (int64), plen, tlen (int32)
  const size_t nn = std::max(n, 1);
  size_t meta_off = (size_t)std::max<int64_t>(nw, 1) * 8;
  if (bucketed) {
    size_t cw = 64;
    while ((int64_t)cw < nw) cw *= 2;
    meta_off = cw * 8;
  }
  const size_t meta_bytes = nn * (3 * 8 + 2 * 4);
  if (!h_words.ensure(meta_off + meta_bytes)) {
    fprintf(stderr, "[WFA] pinned host allocation failed for a batch of %d pairs\n", n);
    return false;
  }
  uint64_t* host_words = h_words.as<uint64_t>();
  const int64_t per_thread = 1 << 16;  // symbols: small batches are packed inline
  const int nt = std::max(1, std::min<int>({16, (int)std::thread::hardware_concurrency(),
                                            (int)std::min<int64_t>(nw * (64 / bits) / per_thread + 1, 16)}));
  auto pack = [&](int t) {
    for (int i = t; i < n; i += nt) {
      pack_seq(host_words + patw[i], pairs[i].pattern, plen[i], bits);
      pack_seq(host_words + txtw[i], pairs[i].text, tlen[i], bits);
    }
  };
  if (nt == 1) {
    pack(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back(pack, t);
    for (auto& x : th) x.join();
  }
  // Longest pairs first, ties in input order: a counting sort over
  // plen + tlen when the longest is within a few times the pair count, else
  // a stable sort.
  sorted.resize(n);
  int64_t lmax = 0;
  for (int i = 0; i < n; ++i) lmax = std::max<int64_t>(lmax, (int64_t)plen[i] + tlen[i]);
  if (lmax <= std::max<int64_t>(4096, 8 * (int64_t)n)) {
    std::vector<int> start((size_t)lmax + 2, 0);
    for (int i = 0; i < n; ++i) ++start[(size_t)(lmax - (plen[i] + tlen[i]))];
    int acc = 0;
    for (auto& c : start) {
      const int cnt = c;
      c = acc;
      acc += cnt;
    }
    for (int i = 0; i < n; ++i) sorted[(size_t)start[(size_t)(lmax - (plen[i] + tlen[i]))]++] = i;
  } else {
    std::iota(sorted.begin(), sorted.end(), 0);
    std::stable_sort(sorted.begin(), sorted.end(), [&](int x, int y) {
      return (int64_t)plen[x] + tlen[x] > (int64_t)plen[y] + tlen[y];
    });
  }
  plan_key.clear();  // new pairs: re-plan on the next run
  have_reverse = false;
  have_results = host_results = ops_on_device = false;
  bool ok = words.ensure(meta_off + meta_bytes) && d_res.ensure(nn * 5 * 4) && d_last.ensure(nn * 4) &&
            d_dsteps.ensure(nn * 8) && d_end_rec.ensure(nn * 4) &&
            (!cigars || d_ops.ensure(std::max<int64_t>(total_ops, 1)));
  if (!ok) {
    fprintf(stderr, "[WFA] device allocation failed for a batch of %d pairs\n", n);
    return false;
  }
  // Symbol positions (word index times symbols per word) for the kernels.
  const int64_t spw = 64 / bits;
  for (int i = 0; i < n; ++i) {
    pat_pos[i] = patw[i] * spw;
    txt_pos[i] = txtw[i] * spw;
  }
  char* dm = words.as<char>() + meta_off;
  d_patw.p = dm;
  d_txtw.p = dm + nn * 8;
  d_ops_off.p = dm + 2 * nn * 8;
  d_plen.p = dm + 3 * nn * 8;
  d_tlen.p = dm + 3 * nn * 8 + nn * 4;
  char* dr = d_res.as<char>();
  d_status.p = dr;
  d_score.p = dr + nn * 4;
  d_ops_begin.p = dr + 2 * nn * 4;
  d_end_k.p = dr + 3 * nn * 4;
  d_end_off.p = dr + 4 * nn * 4;
  char* hm = h_words.as<char>() + meta_off;
  if (n > 0) {
    memcpy(hm, pat_pos.data(), (size_t)n * 8);
    memcpy(hm + nn * 8, txt_pos.data(), (size_t)n * 8);
    memcpy(hm + 2 * nn * 8, ops_off.data(), (size_t)n * 8);
    memcpy(hm + 3 * nn * 8, plen.data(), (size_t)n * 4);
    memcpy(hm + 3 * nn * 8 + nn * 4, tlen.data(), (size_t)n * 4);
  }
  hipStream_t st = e->stream;
  if (upload_bytes) *upload_bytes = meta_off + meta_bytes;
  if (bucketed) return true;
  ok = n == 0 || hip_ok(hipMemcpyAsync(words.p, host_words, meta_off + meta_bytes, hipMemcpyHostToDevice, st),
                        "upload");
  // No wait here: the runs are ordered after the copies on the same stream,
  // and the host arrays stay untouched until the next load.
  return ok;
}

extern "C" wfa_device_batch_t* wfa_device_batch_upload(wavefront_aligner_t* const wf, const wfa_pair_t* pairs,
                                                      int n) {
  Engine* e = engine_of(wf);
  if (!e || !engine_init(e) || n < 0) return nullptr;
  wfa_device_batch_t* b = new (std::nothrow) wfa_device_batch_t();
  if (!b) return nullptr;
  if (!b->load(e, pairs, n, wf->alignment_scope == compute_alignment)) {
    delete b;
    return nullptr;
  }
  return b;
}

extern "C" void wfa_device_batch_free(wfa_device_batch_t* b) { delete b; }

namespace {

// ------------------------------------------------------------ autotuner
// Launch shapes chosen by measurement on the workload itself and cached per
// process: for a key (configuration, kernel family, length bin, error-rate
// bin) the candidate -- the lean family's ring width, or the HBM-ring
// family's threads per pair -- with the least kernel time per diagonal-step.
// A group with no cached choice and enough pairs is split into strided
// slices, one per candidate; each slice runs (all of its retries included)
// as a run of its own, so every pair is aligned once and the slices see the
// same mix of lengths and error rates.  The error-rate bin is measured on
// the slices' results (log2 of 64 x the mean wavefront score per base); a
// later plan predicts it as the one last measured for the length bin.
// WFA_AUTOTUNE=0 turns the tuner off (fixed shapes); WFA_AUTOTUNE_LOG=1
// prints each measurement.
struct TuneKey {
  std::vector<int> conf;
  int family, lbin, ebin;
  bool operator<(const TuneKey& o) const {
    return std::tie(conf, family, lbin, ebin) < std::tie(o.conf, o.family, o.lbin, o.ebin);
  }
};
struct TuneState {
  std::mutex mu;
  std::map<TuneKey, int> choice;     // full key -> candidate
  std::map<TuneKey, int> last_ebin;  // key with ebin = -1 -> error bin last measured
};
TuneState& tuner() {
  static TuneState t;
  return t;
}
bool env_on(const char* name, bool dflt) {
  const char* v = getenv(name);
  return v ? v[0] != '0' : dflt;
}
// Pairs per candidate slice: enough to fill the GPU, so that a slice
// measures throughput rather than the latency of its longest pairs (a slice
// of 50-100 long pairs favoured the widest workgroups and lost 6% against
// the fixed rule on 1-20 kbp pairs).
constexpr int kLeanSliceMin = 16384;   // lean family: about two waves per SIMD
constexpr int kLargeSliceMin = 512;    // HBM-ring family: 2 x the 1024-thread pairs in flight
constexpr int kLargeBlocks[3] = {64, 256, 1024};

// HBM-ring family length bins: log2(L / 512), up to 10.
int large_lbin(int L) {
  int b = 0;
  while (b < 10 && (512 << (b + 1)) <= L) ++b;
  return b;
}

// One candidate slice of a group being measured.
struct TuneTrial {
  TuneKey base;   // ebin = -1
  int candidate;
  PlanGroup group;
};

int predicted_ebin(TuneState& T, const TuneKey& base) {
  auto it = T.last_ebin.find(base);
  return it == T.last_ebin.end() ? -1 : it->second;
}

// Applies cached choices to the plan and moves groups that need measuring
// into trials.
void tune_plan(Engine* e, const JobSet& js, const std::vector<int>& conf, std::vector<PlanGroup>& plan,
               std::vector<TuneTrial>* trials) {
  TuneState& T = tuner();
  std::lock_guard<std::mutex> lk(T.mu);
  std::vector<PlanGroup> out;
  std::vector<int> large_ids;
  for (PlanGroup& g : plan) {
    if (g.family == wfa::kLarge) {
      large_ids.insert(large_ids.end(), g.ids.begin(), g.ids.end());
      continue;
    }
    if (g.family != wfa::kLean || g.lclass < 0 || js.ef.on) {
      out.push_back(std::move(g));
      continue;
    }
    const int cands[2] = {g.lclass == 0 ? 224 : 384, g.lclass == 0 ? 384 : 768};
    TuneKey base{conf, wfa::kLean, g.lclass, -1};
    const int eb = predicted_ebin(T, base);
    TuneKey key = base;
    key.ebin = eb;
    auto it = eb >= 0 ? T.choice.find(key) : T.choice.end();
    if (it != T.choice.end()) {
      g.wcap = it->second;
      out.push_back(std::move(g));
    } else if ((int)g.ids.size() >= 2 * kLeanSliceMin) {
      for (int c = 0; c < 2; ++c) {
        TuneTrial t{base, cands[c], take_group(e)};
        t.group.family = wfa::kLean;
        t.group.block = 64;
        t.group.wcap = cands[c];
        t.group.lmax_words = g.lmax_words;
        t.group.lclass = g.lclass;
        for (size_t i = c; i < g.ids.size(); i += 2) t.group.ids.push_back(g.ids[i]);
        trials->push_back(std::move(t));
      }
      g.ids.clear();
    } else {
      out.push_back(std::move(g));
    }
  }
  // HBM-ring family: per length bin, the cached threads per pair, a
  // measurement, or the fixed rule; bins with one block size share a launch.
  std::map<int, std::vector<int>> by_block;
  std::map<int, std::vector<int>> bins;
  for (int i : large_ids) bins[large_lbin(std::max(js.h_plen[i], js.h_tlen[i]))].push_back(i);
  for (auto& kv : bins) {
    std::vector<int>& ids = kv.second;
    TuneKey base{conf, wfa::kLarge, kv.first, -1};
    const int eb = predicted_ebin(T, base);
    TuneKey key = base;
    key.ebin = eb;
    auto it = eb >= 0 ? T.choice.find(key) : T.choice.end();
    if (it != T.choice.end()) {
      auto& v = by_block[it->second];
      v.insert(v.end(), ids.begin(), ids.end());
    } else if ((int)ids.size() >= 3 * kLargeSliceMin) {
      // 64 threads per pair only up to 2 kbp: on longer pairs its slice runs
      // for seconds (one wave steps through thousands of diagonals per step)
      const int c0 = kv.first <= 1 ? 0 : 1, nc = 3 - c0;
      for (int c = c0; c < 3; ++c) {
        std::vector<int> sl;
        for (size_t i = c - c0; i < ids.size(); i += nc) sl.push_back(ids[i]);
        TuneTrial t{base, kLargeBlocks[c], make_large(e, js, std::move(sl))};
        t.group.block = kLargeBlocks[c];
        trials->push_back(std::move(t));
      }
    } else {
      for (int i : ids) by_block[wfa::pick_block(wfa::kLarge, js.h_plen[i] + js.h_tlen[i])].push_back(i);
    }
  }
  for (auto& kv : by_block) {
    std::vector<int>& ids = kv.second;
    std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) {
      return (int64_t)js.h_plen[x] + js.h_tlen[x] > (int64_t)js.h_plen[y] + js.h_tlen[y];
    });
    PlanGroup g = make_large(e, js, std::move(ids));
    g.block = kv.first;
    out.push_back(std::move(g));
  }
  give_back_groups(e, plan);
  plan = std::move(out);
}

// Runs every trial slice as a run of its own, then stores the choice of each
// measured key.
int run_trials(wavefront_aligner_t* wf, Engine* e, const JobSet& js, int records, std::vector<TuneTrial>& trials,
               RunTotals* tot) {
  struct Cost {
    double ms = 0;
    unsigned long long dsteps = 0;
  };
  std::vector<Cost> costs(trials.size());
  std::vector<int> scratch;
  for (size_t t = 0; t < trials.size(); ++t) {
    std::vector<PlanGroup> one;
    one.push_back(std::move(trials[t].group));
    RunTotals rt;
    const int rc = run_jobs(wf, e, js, one, records, &rt, scratch);
    trials[t].group = std::move(one[0]);
    if (rc != WF_STATUS_OK) return rc;
    costs[t].ms = rt.forward_ms + rt.backtrace_ms;
    costs[t].dsteps = rt.dsteps;
    tot->forward_ms += rt.forward_ms;
    tot->backtrace_ms += rt.backtrace_ms;
    tot->launches += rt.launches;
    tot->dsteps += rt.dsteps;
    tot->fail += rt.fail;
  }
  // error-rate bins from the slices' scores
  std::vector<int> score(js.n);
  if (!hip_ok(hipMemcpyAsync(score.data(), js.score, (size_t)js.n * 4, hipMemcpyDeviceToHost, e->stream), "scores") ||
      !hip_ok(hipStreamSynchronize(e->stream), "scores"))
    return WF_STATUS_OOM;
  const bool log = env_on("WFA_AUTOTUNE_LOG", false);
  TuneState& T = tuner();
  std::lock_guard<std::mutex> lk(T.mu);
  for (size_t t = 0; t < trials.size();) {
    size_t u = t;
    double sum = 0;
    long long cnt = 0;
    while (u < trials.size() && !(trials[t].base < trials[u].base) && !(trials[u].base < trials[t].base)) {
      for (int i : trials[u].group.ids) {
        sum += (double)score[i] / std::max(1, std::max(js.h_plen[i], js.h_tlen[i]));
        ++cnt;
      }
      ++u;
    }
    const double mean = cnt ? sum / cnt : 0;
    int eb = 0;
    while (eb < 12 && 64.0 * mean >= (double)(2 << eb)) ++eb;
    int best = trials[t].candidate;
    double best_cost = 1e300;
    for (size_t v = t; v < u; ++v) {
      const double c = costs[v].dsteps ? costs[v].ms / (double)costs[v].dsteps : 1e300;
      if (log)
        fprintf(stderr, "[WFA] autotune family %d length bin %d error bin %d: candidate %d, %zu pairs, %.3f ms, "
                "%.3e ms per diagonal-step\n", trials[v].base.family, trials[v].base.lbin, eb, trials[v].candidate,
                trials[v].group.ids.size(), costs[v].ms, c);
      // lean family: a wider ring (fewer waves per CU) only if it is
      // measurably faster, not on a tie within the slices' noise
      if (trials[v].base.family == wfa::kLean && v > t ? c * 1.03 < best_cost : c < best_cost) {
        best_cost = c;
        best = trials[v].candidate;
      }
    }
    TuneKey key = trials[t].base;
    key.ebin = eb;
    T.choice[key] = best;
    T.last_ebin[trials[t].base] = eb;
    if (log) fprintf(stderr, "[WFA] autotune: chose %d\n", best);
    t = u;
  }
  std::vector<PlanGroup> done;
  for (auto& tr : trials) done.push_back(std::move(tr.group));
  give_back_groups(e, done);
  return WF_STATUS_OK;
}

}  // namespace

namespace {

// Final per-pair statuses of a run with failures: internal codes the host
// reports differently, and the first failing status.
int settle_statuses(wavefront_aligner_t* wf, std::vector<int>& status) {
  int first_fail = WF_STATUS_ALG_COMPLETED;
  for (size_t i = 0; i < status.size(); ++i) {
    if (status[i] == wfa::kDevBadTrace) {
      fprintf(stderr, "[WFA] internal error: inconsistent backtrace for pair %zu\n", i);
      status[i] = WF_STATUS_UNATTAINABLE;
    }
    if (status[i] == wfa::kDevStaleTrace) {
      // The reference's backtrace (wavefront_backtrace.c:64-219) would read
      // cells of a wavefront the cut-off emptied, which the compute step
      // ignored; that trace is not replayed.
      if (wf->system.verbose >= 0)
        fprintf(stderr, "[WFA] pair %zu: backtrace through a wavefront emptied by the heuristic cut-off "
                "is not supported\n", i);
      status[i] = WF_STATUS_UNSUPPORTED;
    }
    // an end made unreachable (or z-dropped) by the cut-off is a partial
    // result, not a failure
    if (status[i] == wfa::kDevUnreachable || status[i] == wfa::kDevZdropped) continue;
    if (status[i] != WF_STATUS_ALG_COMPLETED && first_fail == WF_STATUS_ALG_COMPLETED) first_fail = status[i];
  }
  return first_fail;
}

// wfa_device_batch_run, and its no-allocation form when the engine's
// buffers are views into a fixed region (e->na set).
int batch_run(wavefront_aligner_t* wf, wfa_device_batch_t* b, Engine* e) {
  const Penalties pen = make_penalties(wf->penalties);
  const int records = (wf->alignment_scope == compute_alignment) ? 1 : 0;
  b->records = records;
  b->ef = ends_free_of(wf);
  b->heur = heuristic_of(wf);
  b->stats = wfa_run_stats_t{};
  b->have_results = false;
  b->host_results = false;
  b->h_ops = nullptr;
  b->runs_ready = false;
  b->first_fail = WF_STATUS_ALG_COMPLETED;
  if (b->n == 0) {
    b->have_results = true;
    b->host_results = true;
    return WF_STATUS_ALG_COMPLETED;
  }
  if (records && !e->na && !b->d_ops.ensure(std::max<int64_t>(b->total_ops, 1))) {
    fprintf(stderr, "[WFA] cannot allocate %lld bytes of CIGAR output\n", (long long)b->total_ops);
    return WF_STATUS_OOM;
  }
  b->ops_on_device = records != 0;
  RunTotals tot;
  int rc;
  b->bi_mode = wf->memory_mode == wavefront_memory_ultralow;
  if (b->bi_mode) {
    rc = bialign_run(wf, e, b, &tot);
    if (rc == WF_STATUS_OOM) return rc;
  } else if (e->na) {
    // a plan of its own, whose order arrays come from the region
    std::vector<PlanGroup> plan = make_plan(b->jobs(), b->sorted, pen);
    for (PlanGroup& g : plan) g.d_ids.region = &e->na->order;
    std::vector<int> scratch;
    rc = run_jobs(wf, e, b->jobs(), plan, records, &tot, scratch);
    give_back_groups(e, plan);
    if (rc != WF_STATUS_OK) return rc;
  } else {
    const std::vector<int> key = {pen.metric, pen.x, pen.o1, pen.e1, pen.o2, pen.e2, b->bits, b->ef.on,
                                  b->heur.cut(), records};
    if (b->plan_key != key) {
      std::vector<PlanGroup> plan = make_plan(b->jobs(), b->sorted, pen);
      std::vector<TuneTrial> trials;
      if (env_on("WFA_AUTOTUNE", true)) tune_plan(e, b->jobs(), key, plan, &trials);
      if (!trials.empty()) {
        // measure the candidates on their slices (aligning those pairs)
        rc = run_trials(wf, e, b->jobs(), records, trials, &tot);
        if (rc != WF_STATUS_OK) return rc;
      }
      recycle_plan(e, b->plan, plan);
      b->plan = std::move(plan);
      // after a measurement the next run plans again, with the choices
      b->plan_key = trials.empty() ? key : std::vector<int>();
    }
    std::vector<int> scratch;
    rc = run_jobs(wf, e, b->jobs(), b->plan, records, &tot, scratch);
    if (rc != WF_STATUS_OK) return rc;
  }
  b->stats.diagonal_steps = (int64_t)tot.dsteps;
  b->stats.forward_ms = tot.forward_ms;
  b->stats.backtrace_ms = tot.backtrace_ms;
  b->stats.forward_launches = tot.launches;
  b->stats.record_arena_bytes = (int64_t)e->rec.cap;
  b->have_results = true;
  b->host_results = false;
  if (tot.fail == 0) return WF_STATUS_ALG_COMPLETED;
  if (e->na) {
    // statuses through pageable host memory (no pinned staging), settled,
    // and written back for the later download
    std::vector<int> st(b->n);
    if (!hip_ok(hipMemcpyAsync(st.data(), b->d_status.p, (size_t)b->n * 4, hipMemcpyDeviceToHost, e->stream),
                "status") ||
        !hip_ok(hipStreamSynchronize(e->stream), "status"))
      return WF_STATUS_OOM;
    b->first_fail = settle_statuses(wf, st);
    if (!hip_ok(hipMemcpyAsync(b->d_status.p, st.data(), (size_t)b->n * 4, hipMemcpyHostToDevice, e->stream),
                "status") ||
        !hip_ok(hipStreamSynchronize(e->stream), "status"))
      return WF_STATUS_OOM;
    return b->first_fail;
  }
  if (!download_results(e, b)) return WF_STATUS_OOM;
  b->first_fail = settle_statuses(wf, b->status);
  return b->first_fail;
}

}  // namespace

extern "C" int wfa_device_batch_run(wavefront_aligner_t* const wf, wfa_device_batch_t* b) {
  if (!b) return WF_STATUS_OOM;
  presets_check(wf, b->plen.data(), b->tlen.data(), b->n);
  const int sup = check_supported(wf);
  if (sup != WF_STATUS_OK) return sup;
  Engine* e = engine_of(wf);
  if (!e || !engine_init(e)) return WF_STATUS_OOM;
  return batch_run(wf, b, e);
}

namespace wfa_host {

// The no-allocation run (wfa_device_batch_run_noalloc, wfa_noalloc.hip): the
// region [scratch, scratch + bytes) is split into the run counters, the plan
// groups' order arrays (8 bytes per pair covers the first plan and every
// retry plan), the HBM-ring share and the record arena; the engine's own
// buffers are swapped out for views of these parts during the run and
// swapped back afterwards.
int device_batch_run_in(wavefront_aligner_t* wf, wfa_device_batch_t* b, char* scratch, size_t bytes) {
  constexpr size_t kRingShare = 64ull << 20;
  constexpr size_t kMinArena = 32ull << 20;
  if (!b) return WF_STATUS_OOM;
  presets_check(wf, b->plen.data(), b->tlen.data(), b->n);
  const int sup = check_supported(wf);
  if (sup != WF_STATUS_OK) return sup;
  Engine* e = engine_of(wf);
  if (!e || !e->ok) {
    fprintf(stderr, "[WFA] wfa_device_batch_run_noalloc: the batch must be uploaded with this aligner first\n");
    return WF_STATUS_OOM;
  }
  if (wf->memory_mode == wavefront_memory_ultralow) {
    fprintf(stderr, "[WFA] wfa_device_batch_run_noalloc: memory mode ultralow (BiWFA) is not available "
            "without allocation; use wfa_device_batch_run\n");
    return WF_STATUS_UNSUPPORTED;
  }
  const int records = (wf->alignment_scope == compute_alignment) ? 1 : 0;
  if (records && b->d_ops.cap < (size_t)std::max<int64_t>(b->total_ops, 1)) {
    fprintf(stderr, "[WFA] wfa_device_batch_run_noalloc: the batch has no CIGAR output buffer (upload it with "
            "an aligner that computes alignments)\n");
    return WF_STATUS_OOM;
  }
  const size_t order_bytes = ((size_t)b->n * 8 + 4096 + 255) & ~(size_t)255;
  if (256 + order_bytes + kRingShare + kMinArena > bytes) {
    fprintf(stderr, "[WFA] wfa_device_batch_run_noalloc: %d pairs exceed the %zu MB scratch array "
            "(at most %zu pairs per call)\n", b->n, bytes >> 20,
            (size_t)((bytes - 256 - kRingShare - kMinArena - 4096) / 8));
    return WF_STATUS_OOM;
  }
  const Penalties pen = make_penalties(wf->penalties);
  for (int i = 0; i < b->n; ++i) {
    const size_t ring = (size_t)wfa::ring_slots(pen) * ((size_t)b->plen[i] + b->tlen[i] + 8) * sizeof(int);
    if (ring > kRingShare) {
      fprintf(stderr, "[WFA] wfa_device_batch_run_noalloc: pair %d (%d + %d symbols) needs a %zu MB wavefront "
              "ring, more than the %zu MB share of the scratch array\n", i, b->plen[i], b->tlen[i], ring >> 20,
              kRingShare >> 20);
      return WF_STATUS_OOM;
    }
  }
  // counters | order arrays | ring share | record arena
  Engine::NoAlloc na;
  na.order.base = scratch + 256;
  na.order.bytes = order_bytes;
  na.ring_bytes = kRingShare;
  ScratchRegion fixed;  // binds the views below: they never allocate or free
  DevBuf counters, ring, rec;
  counters.p = scratch;
  counters.cap = 256;
  counters.region = &fixed;
  ring.p = scratch + 256 + order_bytes;
  ring.cap = kRingShare;
  ring.region = &fixed;
  rec.p = scratch + 256 + order_bytes + kRingShare;
  rec.cap = (bytes - 256 - order_bytes - kRingShare) & ~(size_t)255;
  rec.region = &fixed;
  std::swap(e->counters, counters);
  std::swap(e->ring, ring);
  std::swap(e->rec, rec);
  const size_t rec_bytes = e->rec_bytes, rec_cap = e->rec_cap;
  e->rec_bytes = e->rec_cap = e->rec.cap;
  e->na = &na;
  const int rc = batch_run(wf, b, e);
  e->na = nullptr;
  e->rec_bytes = rec_bytes;
  e->rec_cap = rec_cap;
  std::swap(e->counters, counters);
  std::swap(e->ring, ring);
  std::swap(e->rec, rec);
  return rc;
}

}  // namespace wfa_host

extern "C" int wfa_device_batch_stats(wfa_device_batch_t* b, wfa_run_stats_t* stats) {
  if (!b || !stats || !b->have_results) return -1;
  *stats = b->stats;
  return 0;
}

extern "C" int wfa_device_batch_download(wavefront_aligner_t* const wf, wfa_device_batch_t* b,
                                         wfa_result_t* results, char* cigar_arena,
                                         size_t cigar_arena_bytes) {
  if (!b || !b->have_results) return WF_STATUS_UNATTAINABLE;
  Engine* e = engine_of(wf);
  const bool want_ops = cigar_arena != nullptr && b->records && b->ops_on_device;
  if (want_ops && cigar_arena_bytes < (size_t)b->total_ops) {
    fprintf(stderr, "[WFA] cigar arena too small (%zu < %lld bytes)\n", cigar_arena_bytes,
            (long long)b->total_ops);
    return WF_STATUS_OOM;
  }
  if (!download_results(e, b, want_ops ? cigar_arena : nullptr)) return WF_STATUS_OOM;
  const bool extension = wf->alignment_form.extension && b->records && !b->bi_mode;
  if ((extension || (b->heur.zdrop && b->records && !b->bi_mode)) && !want_ops) {
    fprintf(stderr, "[WFA] the extension form and z-drop need the CIGAR arena (they trim the CIGAR)\n");
    return WF_STATUS_OOM;
  }
  int first_fail = WF_STATUS_ALG_COMPLETED;
  for (int i = 0; i < b->n; ++i) {
    wfa_result_t& r = results[i];
    int st = b->status[i];
    if (st == wfa::kDevBadTrace) st = WF_STATUS_UNATTAINABLE;
    r.status = st;
    r.end_v = b->plen[i];
    r.end_h = b->tlen[i];
    r.cigar_offset = b->ops_off[i];
    r.cigar_length = 0;
    if (st == WF_STATUS_ALG_COMPLETED) {
      if (b->records) {
        r.cigar_offset = b->ops_off[i] + b->ops_begin[i];
        r.cigar_length = b->plen[i] + b->tlen[i] - b->ops_begin[i];
      }
      if (b->bi_mode) {
        r.score = b->bi_score[i];
      } else if (!b->records) {
        // compute_score: the score at the full lengths (wavefront_unialign.c:159-166)
        r.score = classic_score(wf->penalties, b->plen[i], b->tlen[i], b->score[i]);
      } else {
        // wavefront_unialign_terminate (wavefront_unialign.c:167-236): the
        // end cell gives the end position and the score's lengths; the
        // extension form then keeps the best-scoring CIGAR prefix.
        const int k = b->ef.on ? b->end_k[i] : b->tlen[i] - b->plen[i];
        const int off = b->ef.on ? b->end_off[i] : b->tlen[i];
        r.end_v = off - k;
        r.end_h = off;
        if (!extension) {
          r.score = classic_score(wf->penalties, r.end_v, r.end_h, b->score[i]);
        } else {
          const Trim t = maxtrim(wf->penalties, cigar_arena + r.cigar_offset, r.cigar_length);
          r.score = b->score[i];  // the backtrace's raw score (wavefront_backtrace.c:96, :527)
          r.end_v = r.end_h = -1;
          if (t.examined) {
            r.cigar_length = t.keep;
            r.score = t.score;
            r.end_v = t.end_v;
            r.end_h = t.end_h;
          }
          if (t.trimmed) r.status = st = WF_STATUS_ALG_PARTIAL;
        }
      }
    } else if (st == WF_STATUS_MAX_STEPS_REACHED && !b->bi_mode) {
      r.score = -wf->system.max_alignment_steps;  // wavefront_unialign.c:102-107
    } else if (b->bi_mode) {
      // what the BiWFA recursion left in cigar->score, and the CIGAR it had
      // appended before the failure
      r.score = b->bi_score[i];
      if (b->records) {
        r.cigar_offset = b->ops_off[i] + b->ops_begin[i];
        r.cigar_length = b->plen[i] + b->tlen[i] - b->ops_begin[i];
      }
    } else if (st == wfa::kDevUnreachable || st == wfa::kDevZdropped) {
      // WF_STATUS_END_UNREACHABLE (wavefront_unialign.c:147-236) at the end
      // cell a z-drop recorded, or at none (alignment_end_pos keeps
      // k = INT_MAX and offset = NULL, wavefront_aligner.c:411-414).
      // Score-only takes the classic score at that end with the wrap-around
      // of 32-bit arithmetic.  With a CIGAR, a z-drop's backtrace (traced on
      // the device) is max-trimmed, keeping the backtrace's raw score when
      // max-trim does not examine it; without an end cell, max-trim of the
      // empty CIGAR leaves the cleared score and end (cigar.c:79-90).
      const bool zdrop = st == wfa::kDevZdropped;
      r.status = st = WF_STATUS_ALG_PARTIAL;
      const uint32_t off = zdrop ? (uint32_t)b->end_off[i] : (uint32_t)wfa::kOffsetNull;
      const uint32_t kk = zdrop ? (uint32_t)b->end_k[i] : (uint32_t)INT_MAX;
      if (!b->records) {
        r.end_v = (int)(off - kk);
        r.end_h = (int)off;
        const int swg_match = -wf->penalties.match;
        if (wf->penalties.distance_metric <= edit || swg_match == 0) {
          r.score = classic_score(wf->penalties, 0, 0, b->score[i]);
        } else {
          const uint32_t num = (uint32_t)swg_match * ((uint32_t)r.end_v + (uint32_t)r.end_h) - (uint32_t)b->score[i];
          r.score = (int)num / 2;
        }
      } else if (!zdrop) {
        r.score = INT32_MIN;
        r.end_v = r.end_h = -1;
      } else {
        r.cigar_offset = b->ops_off[i] + b->ops_begin[i];
        r.cigar_length = b->plen[i] + b->tlen[i] - b->ops_begin[i];
        const Trim t = maxtrim(wf->penalties, cigar_arena + r.cigar_offset, r.cigar_length);
        r.score = b->score[i];  // the backtrace's raw score (wavefront_backtrace.c:317, :527)
        r.end_v = r.end_h = -1;
        if (t.examined) {
          r.cigar_length = t.keep;
          r.score = t.score;
          r.end_v = t.end_v;
          r.end_h = t.end_h;
        }
      }
    } else {
      r.score = INT32_MIN;
    }
    if (st != WF_STATUS_ALG_COMPLETED && st != WF_STATUS_ALG_PARTIAL && first_fail == WF_STATUS_ALG_COMPLETED)
      first_fail = st;
  }
  return first_fail;
}

namespace {

// Single-pair graphs can be turned off with WFA_PLAIN_GRAPH=0 (read at
// every call, so a process can compare both forms).
bool plain_graphs_enabled() {
  const char* v = getenv("WFA_PLAIN_GRAPH");
  return !(v && v[0] == '0');
}

template <typename T>
void key_add(std::vector<unsigned char>& k, const T& v) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
  k.insert(k.end(), p, p + sizeof(T));
}

// One pair through a captured graph (see Engine::PlainGraph).  Returns
// WF_STATUS_OK when the pair's results are ready in the batch (on the host),
// a failing status on a HIP error, or 1 when the pair must take the general
// path (a configuration with more than one launch group, or a retry).
int plain_run(wavefront_aligner_t* wf, Engine* e, wfa_device_batch_t* b, const wfa_pair_t* pair) {
  const int records = (wf->alignment_scope == compute_alignment) ? 1 : 0;
  size_t upload = 0;
  if (!b->load(e, pair, 1, records != 0, true, &upload)) return WF_STATUS_OOM;
  const Penalties pen = make_penalties(wf->penalties);
  b->records = records;
  b->ef = ends_free_of(wf);
  b->heur = heuristic_of(wf);
  b->stats = wfa_run_stats_t{};
  b->bi_mode = false;
  b->first_fail = WF_STATUS_ALG_COMPLETED;
  hipStream_t st = e->stream;
  const JobSet js = b->jobs();
  std::vector<PlanGroup> plan = make_plan(js, b->sorted, pen);
  auto general = [&]() {
    // the staged inputs go up uncaptured; the caller runs the general path
    return hip_ok(hipMemcpyAsync(b->words.p, b->h_words.p, upload, hipMemcpyHostToDevice, st), "upload")
               ? 1 : WF_STATUS_OOM;
  };
  if (plan.size() != 1) return general();
  PlanGroup& g = plan[0];
  if (g.family != wfa::kLarge) g.lmax_words = (int)(upload / 8);  // the bucket's words: a capacity, not a length
  if (!reserve_for(e, plan, js.bits, pen, js.heur.cut(), records)) return WF_STATUS_OOM;
  size_t ops_cap = 1024;
  while ((int64_t)ops_cap < b->total_ops) ops_cap *= 2;
  if ((records && !b->d_ops.ensure(ops_cap)) || !e->plain_out.ensure(256 + ops_cap)) return WF_STATUS_OOM;
  const LaunchShape shape = shape_of(e, g, js.bits, pen, js.heur.cut());
  ForwardArgs fa;
  BacktraceArgs ba;
  group_args(wf, e, js, g, pen, records, true, &fa, &ba);
  char* out = e->plain_out.as<char>();
  std::vector<unsigned char> key;
  key_add(key, fa);
  key_add(key, ba);
  key_add(key, shape);
  key_add(key, g.family);
  key_add(key, js.bits);
  key_add(key, records);
  key_add(key, upload);
  key_add(key, ops_cap);
  key_add(key, b->words.p);
  key_add(key, b->h_words.p);
  key_add(key, b->d_res.p);
  key_add(key, b->d_ops.p);
  key_add(key, out);
  Engine::PlainGraph& pg = e->plain;
  if (pg.key != key) {
    if (pg.seen != key) {
      pg.seen = key;  // first sighting: run uncaptured
      return general();
    }
    // second sighting in a row: capture it
    if (pg.exec) (void)hipGraphExecDestroy(pg.exec);
    if (pg.graph) (void)hipGraphDestroy(pg.graph);
    pg.exec = nullptr;
    pg.graph = nullptr;
    pg.key.clear();
    hipGraph_t graph = nullptr;
    if (!hip_ok(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed), "graph capture")) {
      pg.off = true;
      return general();
    }
    bool ok = hipMemcpyAsync(b->words.p, b->h_words.p, upload, hipMemcpyHostToDevice, st) == hipSuccess &&
              wfa::launch_reset_counters(counters_of(e), 2, st) == hipSuccess &&
              launch_group(g, shape, fa, ba, pen.ncomp, js.bits, records, st) == hipSuccess &&
              hipMemcpyAsync(out, b->d_res.p, 5 * sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
              hipMemcpyAsync(out + 64, e->counters.p, sizeof(wfa::Counters), hipMemcpyDeviceToHost, st) ==
                  hipSuccess &&
              (!records || hipMemcpyAsync(out + 256, b->d_ops.p, ops_cap, hipMemcpyDeviceToHost, st) == hipSuccess);
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    hipGraphExec_t exec = nullptr;
    if (ok && ec == hipSuccess && graph && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
      pg.graph = graph;
      pg.exec = exec;
      pg.key = key;
    } else {
      if (graph) (void)hipGraphDestroy(graph);
      (void)hipGetLastError();
      fprintf(stderr, "[WFA] single-pair graph capture failed; single pairs run uncaptured\n");
      pg.off = true;
      return general();
    }
  }
  if (!hip_ok(hipGraphLaunch(pg.exec, st), "graph launch") || !hip_ok(hipStreamSynchronize(st), "graph"))
    return WF_STATUS_OOM;
  ++pg.replays;
  wfa::Counters c;
  memcpy(&c, out + 64, sizeof(c));
  if (c.retry != 0) return 1;  // a ring or the arena overflowed: the general path retries (inputs are up)
  b->status.assign(1, 0);
  b->score.assign(1, 0);
  b->ops_begin.assign(1, 0);
  b->end_k.assign(1, 0);
  b->end_off.assign(1, 0);
  memcpy(b->status.data(), out, 4);
  memcpy(b->score.data(), out + 4, 4);
  memcpy(b->ops_begin.data(), out + 8, 4);
  memcpy(b->end_k.data(), out + 12, 4);
  memcpy(b->end_off.data(), out + 16, 4);
  b->host_results = true;
  b->have_results = true;
  b->runs_ready = false;
  b->ops_on_device = records != 0;
  b->h_ops = records ? out + 256 : nullptr;
  b->stats.diagonal_steps = (int64_t)c.dsteps;
  b->stats.forward_launches = 1;
  b->stats.record_arena_bytes = (int64_t)e->rec.cap;
  if (c.fail) b->first_fail = settle_statuses(wf, b->status);
  return WF_STATUS_OK;
}

}  // namespace

namespace {

// Builds the run-length CIGARs of the batch's last run on the device (once
// per run) and brings the per-pair run offsets to the host.
int make_runs(wavefront_aligner_t* wf, Engine* e, wfa_device_batch_t* b) {
  if (b->runs_ready) return WF_STATUS_OK;
  if (!b->have_results) return WF_STATUS_UNATTAINABLE;
  const bool extension = wf->alignment_form.extension && !b->bi_mode;
  if (extension || (b->heur.zdrop && !b->bi_mode)) {
    fprintf(stderr, "[WFA] run-length CIGARs are not made for the extension form or z-drop (their CIGARs are "
            "trimmed on the host): use wfa_device_batch_download\n");
    return WF_STATUS_UNSUPPORTED;
  }
  const int n = b->n;
  b->h_run_off.assign((size_t)n + 1, 0);
  b->total_runs = 0;
  if (n == 0 || !b->records || !b->ops_on_device) {
    b->runs_ready = true;
    return WF_STATUS_OK;
  }
  hipStream_t st = e->stream;
  wfa::CigarRunsArgs a{};
  a.n = n;
  a.keep_failed = b->bi_mode ? 1 : 0;
  a.ops = b->d_ops.as<char>();
  a.ops_off = b->d_ops_off.as<int64_t>();
  a.ops_begin = b->d_ops_begin.as<int>();
  a.plen = b->d_plen.as<int>();
  a.tlen = b->d_tlen.as<int>();
  a.status = b->d_status.as<int>();
  if (!b->d_run_counts.ensure((size_t)n * 8) || !b->d_run_off.ensure(((size_t)n + 1) * 8)) return WF_STATUS_OOM;
  a.counts = b->d_run_counts.as<long long>();
  a.offsets = b->d_run_off.as<long long>();
  size_t temp = 0;
  if (!hip_ok(wfa::launch_cigar_runs_count(a, nullptr, &temp, st), "scan size") ||
      !b->d_scan_temp.ensure(std::max<size_t>(temp, 256)))
    return WF_STATUS_OOM;
  temp = b->d_scan_temp.cap;
  if (!hip_ok(wfa::launch_cigar_runs_count(a, b->d_scan_temp.p, &temp, st), "run count") ||
      !hip_ok(hipMemcpyAsync(b->h_run_off.data(), b->d_run_off.p, ((size_t)n + 1) * 8, hipMemcpyDeviceToHost, st),
              "run offsets") ||
      !hip_ok(hipStreamSynchronize(st), "run offsets"))
    return WF_STATUS_OOM;
  b->total_runs = b->h_run_off[n];
  if (b->total_runs > 0) {
    if (!b->d_runs.ensure((size_t)b->total_runs * 4)) return WF_STATUS_OOM;
    a.runs = b->d_runs.as<uint32_t>();
    if (!hip_ok(wfa::launch_cigar_runs_write(a, st), "run write")) return WF_STATUS_OOM;
  }
  b->runs_ready = true;
  return WF_STATUS_OK;
}

}  // namespace

extern "C" int64_t wfa_device_batch_cigar_runs(wavefront_aligner_t* const wf, wfa_device_batch_t* b) {
  if (!wf || !b) return WF_STATUS_OOM;
  Engine* e = engine_of(wf);
  if (!e || !e->ok) return WF_STATUS_OOM;
  const int rc = make_runs(wf, e, b);
  return rc == WF_STATUS_OK ? (int64_t)b->total_runs : (int64_t)rc;
}

extern "C" int wfa_device_batch_download_runs(wavefront_aligner_t* const wf, wfa_device_batch_t* b,
                                              wfa_result_t* results, uint32_t* runs, size_t max_runs) {
  if (!wf || !b || !b->have_results) return WF_STATUS_UNATTAINABLE;
  Engine* e = engine_of(wf);
  if (!e || !e->ok) return WF_STATUS_OOM;
  int rc = make_runs(wf, e, b);
  if (rc != WF_STATUS_OK) return rc;
  if (b->total_runs > 0 && (runs == nullptr || max_runs < (size_t)b->total_runs)) {
    fprintf(stderr, "[WFA] wfa_device_batch_download_runs: runs must hold %lld entries\n",
            (long long)b->total_runs);
    return WF_STATUS_OOM;
  }
  // the runs go down on the stream while the host settles the results
  if (b->total_runs > 0 &&
      !hip_ok(hipMemcpyAsync(runs, b->d_runs.p, (size_t)b->total_runs * 4, hipMemcpyDeviceToHost, e->stream),
              "run download"))
    return WF_STATUS_OOM;
  rc = wfa_device_batch_download(wf, b, results, nullptr, 0);
  if (!hip_ok(hipStreamSynchronize(e->stream), "run download")) return WF_STATUS_OOM;
  for (int i = 0; i < b->n; ++i) {
    wfa_result_t& r = results[i];
    const long long cnt = b->h_run_off[i + 1] - b->h_run_off[i];
    r.cigar_offset = b->h_run_off[i];
    r.cigar_length = r.cigar_length > 0 ? (int)cnt : 0;
  }
  return rc;
}

extern "C" int wavefront_align_batch(wavefront_aligner_t* const wf, const wfa_pair_t* pairs, int num_pairs,
                                     wfa_result_t* results, char* cigar_arena, size_t cigar_arena_bytes) {
  if (num_pairs > 0) presets_check(wf, pairs, num_pairs);
  const int sup = check_supported(wf);
  if (sup != WF_STATUS_OK) {
    for (int i = 0; i < num_pairs; ++i) {
      results[i] = wfa_result_t{sup, INT32_MIN, -1, -1, 0, 0};
    }
    return sup;
  }
  if (num_pairs <= 0) return WF_STATUS_ALG_COMPLETED;
  // Every early failure leaves each result with that status and no CIGAR.
  auto fail_all = [&](int st) {
    for (int i = 0; i < num_pairs; ++i) results[i] = wfa_result_t{st, INT32_MIN, -1, -1, 0, 0};
    return st;
  };
  if (wf->alignment_scope == compute_alignment &&
      (cigar_arena == nullptr || cigar_arena_bytes < wavefront_align_batch_arena_bytes(pairs, num_pairs))) {
    fprintf(stderr, "[WFA] wavefront_align_batch: cigar_arena must hold %zu bytes\n",
            wavefront_align_batch_arena_bytes(pairs, num_pairs));
    return fail_all(WF_STATUS_OOM);
  }
  // The aligner's resident batch: its device buffers, pinned staging and
  // plan buffers are reused from call to call (grown when too small).
  Engine* e = engine_of(wf);
  if (!e || !engine_init(e)) return fail_all(WF_STATUS_OOM);
  if (!e->scratch) e->scratch = new (std::nothrow) wfa_device_batch_t();
  wfa_device_batch_t* b = e->scratch;
  if (!b) return fail_all(WF_STATUS_OOM);
  int rc;
  if (num_pairs == 1 && wf->memory_mode != wavefront_memory_ultralow && !e->plain.off && plain_graphs_enabled()) {
    rc = plain_run(wf, e, b, pairs);
    if (rc == WF_STATUS_OK) return wfa_device_batch_download(wf, b, results, cigar_arena, cigar_arena_bytes);
    if (rc != 1) return fail_all(rc);
    rc = batch_run(wf, b, e);  // inputs already on the device
  } else {
    if (!b->load(e, pairs, num_pairs, wf->alignment_scope == compute_alignment)) return fail_all(WF_STATUS_OOM);
    rc = wfa_device_batch_run(wf, b);
  }
  if (rc == WF_STATUS_OOM && !b->have_results) return fail_all(rc);
  const int drc = wfa_device_batch_download(wf, b, results, cigar_arena, cigar_arena_bytes);
  if (drc == WF_STATUS_OOM && !b->host_results) return fail_all(drc);
  return drc;
}

// ------------------------------------------------------------ single pair

namespace {

void remember_sequences(wavefront_aligner_t* wf, const char* p, int plen, const char* t, int tlen) {
  free(wf->sequences.pattern);
  free(wf->sequences.text);
  wf->sequences.pattern = static_cast<char*>(malloc(plen + 1));
  wf->sequences.text = static_cast<char*>(malloc(tlen + 1));
  memcpy(wf->sequences.pattern, p, plen);
  wf->sequences.pattern[plen] = '\0';
  memcpy(wf->sequences.text, t, tlen);
  wf->sequences.text[tlen] = '\0';
  wf->sequences.pattern_length = plen;
  wf->sequences.text_length = tlen;
}

int align_one(wavefront_aligner_t* wf) {
  cigar_t* cigar = wf->cigar;
  const int plen = wf->sequences.pattern_length, tlen = wf->sequences.text_length;
  cigar->begin_offset = cigar->end_offset = 0;
  cigar->score = INT32_MIN;
  cigar->end_v = cigar->end_h = -1;
  cigar->cigar_length = 0;  // drop the cached SAM CIGAR
  wf->align_status.dropped = false;
  presets_check(wf, &plen, &tlen, 1);
  const int sup = check_supported(wf);
  if (sup != WF_STATUS_OK) {
    wf->align_status.status = sup;
    return sup;
  }
  wfa_pair_t pr{wf->sequences.pattern, plen, wf->sequences.text, tlen};
  const bool cig = wf->alignment_scope == compute_alignment;
  if (cig && cigar->max_operations < plen + tlen + 1) {
    free(cigar->operations);
    cigar->max_operations = 2 * (plen + tlen) + 1;
    cigar->operations = static_cast<char*>(malloc(cigar->max_operations));
    free(cigar->cigar_buffer);
    cigar->cigar_buffer = static_cast<uint32_t*>(malloc(sizeof(uint32_t) * cigar->max_operations));
  }
  wfa_result_t r{WF_STATUS_OOM, INT32_MIN, -1, -1, 0, 0};
  const int rc = wavefront_align_batch(wf, &pr, 1, &r, cig ? cigar->operations : nullptr,
                                       cig ? (size_t)cigar->max_operations : 0);
  (void)rc;
  wf->align_status.status = r.status;
  wf->align_status.score = r.score;
  const bool done = r.status == WF_STATUS_ALG_COMPLETED || r.status == WF_STATUS_ALG_PARTIAL;
  if (done || r.status == WF_STATUS_MAX_STEPS_REACHED) {
    cigar->score = r.score;
  }
  if (done) {
    cigar->end_v = r.end_v;
    cigar->end_h = r.end_h;
    if (cig) {
      cigar->begin_offset = (int)r.cigar_offset;
      cigar->end_offset = (int)r.cigar_offset + r.cigar_length;
    }
  }
  return r.status;
}

}  // namespace

extern "C" int wavefront_align(wavefront_aligner_t* const wf, const char* const pattern, const int plen,
                               const char* const text, const int tlen) {
  remember_sequences(wf, pattern, plen, text, tlen);
  return align_one(wf);
}

extern "C" int wavefront_align_packed2bits(wavefront_aligner_t* const wf, const uint8_t* const pattern,
                                           const int plen, const uint8_t* const text, const int tlen) {
  // Decodes every base (the layout of wavefront_sequences.c:102-140).
  static const char dec[4] = {'A', 'C', 'G', 'T'};
  std::vector<char> p(plen), t(tlen);
  for (int i = 0; i < plen; ++i) p[i] = dec[(pattern[i >> 2] >> (2 * (i & 3))) & 3];
  for (int i = 0; i < tlen; ++i) t[i] = dec[(text[i >> 2] >> (2 * (i & 3))) & 3];
  remember_sequences(wf, p.data(), plen, t.data(), tlen);
  return align_one(wf);
}

extern "C" int wavefront_align_lambda(wavefront_aligner_t* const wf, alignment_match_funct_t const,
                                      void*, const int, const int) {
  fprintf(stderr, "[WFA] Unsupported configuration on the device path: lambda sequences\n");
  wf->align_status.status = WF_STATUS_UNSUPPORTED;
  return WF_STATUS_UNSUPPORTED;
}

// wavefront_align.c:245-266: continue a pair paused by the step limit.  The
// device recomputes it from score 0 under the current limits, which yields
// the same result as continuing.
extern "C" int wavefront_align_resume(wavefront_aligner_t* const wf) {
  if (wf->align_status.status != WF_STATUS_MAX_STEPS_REACHED || wf->memory_mode == wavefront_memory_ultralow) {
    fprintf(stderr, "[WFA] Alignment cannot be resumed\n");
    exit(1);
  }
  return align_one(wf);
}
