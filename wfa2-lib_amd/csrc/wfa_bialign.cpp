// BiWFA (memory mode ultralow) on the device, for whole batches.
//
// wavefront_bialign_alignment (wavefront_bialign.c:1145-1222) recurses depth
// first; every sub-problem depends only on its own bounds and components, so
// here the recursion runs breadth first: all sub-problems of one level, from
// every pair of the batch, go to one breakpoint-kernel launch; their halves
// form the next level.  Base cases (score_remaining <= 250, or the whole pair
// when max(n,m) <= 100, or an end reached at score <= 500) are collected and
// aligned in one pass of the unidirectional kernels, with their begin/end
// components; trivial halves (an empty side) are runs of D or I.  A final
// gather kernel concatenates the pieces of every pair in sequence order.
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "wfa_internal.h"

namespace wfa_host {
namespace {

// wavefront_bialign.c:52-54
constexpr int kFallbackMinScore = 250;
constexpr int kFallbackMinLength = 100;
constexpr int kRecoveryMinScore = 500;

struct Node {
  wfa::BiProblem prob;
  int score_remaining;
  int level;
};

struct Leaf {
  int pb, tb;      // start of the piece in the pair's sequences
  char op;         // run operation (D or I) when base < 0
  int len;         // run length
  int base;        // index of the base-case job, or -1
};

struct BaseScore {
  int pb, tb, score;
};

struct PairState {
  int status = WF_STATUS_ALG_COMPLETED;
  int fail_pb = INT_MAX, fail_tb = INT_MAX;  // leftmost failing sub-problem (depth-first order)
  bool fail_base = false;  // ... and whether it is a base alignment, with its score
  int fail_score = INT32_MIN;
  int score = INT32_MIN;   // cigar->score
  std::vector<Leaf> leaves;
  std::vector<BaseScore> bases;  // completed base alignments
};

// The reference's recursion stops at the first failing sub-problem in
// depth-first order, which is the leftmost one.
void fail_pair(PairState& ps, int status, int pb, int tb, bool base = false, int score = INT32_MIN) {
  if (pb < ps.fail_pb || (pb == ps.fail_pb && tb < ps.fail_tb)) {
    ps.fail_pb = pb;
    ps.fail_tb = tb;
    ps.status = status;
    ps.fail_base = base;
    ps.fail_score = score;
  }
}

// cigar->score of a failed pair: every base alignment sets it
// (wavefront_bialign.c:181-182), so it is the failing base's own score, or
// that of the last base alignment run before the failing search, if any.
int failed_score(const PairState& ps) {
  if (ps.fail_base) return ps.fail_score;
  int best = INT32_MIN, bpb = INT_MIN, btb = INT_MIN;
  for (const BaseScore& b : ps.bases) {
    const bool before = b.pb < ps.fail_pb || (b.pb == ps.fail_pb && b.tb < ps.fail_tb);
    if (before && (b.pb > bpb || (b.pb == bpb && b.tb > btb))) {
      bpb = b.pb;
      btb = b.tb;
      best = b.score;
    }
  }
  return best;
}

bool ensure_reverse(Engine* e, wfa_device_batch_t* b) {
  if (b->have_reverse) return true;
  const int n = b->n;
  if (!b->rev_words.ensure(std::max<int64_t>(b->words_used, 1) * 8)) return false;
  std::vector<int64_t> pos(2 * (size_t)n);
  std::vector<int> len(2 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    pos[i] = b->pat_pos[i];
    pos[n + i] = b->txt_pos[i];
    len[i] = b->plen[i];
    len[n + i] = b->tlen[i];
  }
  DevBuf& d_pos = e->bi.pos;
  DevBuf& d_len = e->bi.len;
  if (!d_pos.ensure(pos.size() * 8) || !d_len.ensure(len.size() * 4)) return false;
  hipStream_t st = e->stream;
  bool ok = hip_ok(hipMemcpyAsync(d_pos.p, pos.data(), pos.size() * 8, hipMemcpyHostToDevice, st), "reverse") &&
            hip_ok(hipMemcpyAsync(d_len.p, len.data(), len.size() * 4, hipMemcpyHostToDevice, st), "reverse") &&
            hip_ok(wfa::launch_reverse_pack(b->words.as<uint64_t>(), d_pos.as<int64_t>(), d_pos.as<int64_t>(),
                                            d_len.as<int>(), 2 * n, b->bits, b->rev_words.as<uint64_t>(), st),
                   "reverse pack") &&
            hip_ok(hipStreamSynchronize(st), "reverse pack");
  b->have_reverse = ok;
  return ok;
}

int bp_block(int size) { return size <= 512 ? 64 : (size <= 8192 ? 256 : 1024); }

// Bound of every wait of a paired search: 4 s of the 100 MHz device clock.
// WFA_PAIR_WAIT_TICKS sets another (tests force the one-workgroup fallback
// with a bound of a few ticks).
unsigned long long pair_wait_ticks() {
  const char* v = getenv("WFA_PAIR_WAIT_TICKS");
  return v ? strtoull(v, nullptr, 10) : 400000000ull;
}

// WFA_PAIR_SEARCH=0 turns paired-workgroup breakpoint searches off.
bool pair_searches_enabled() {
  const char* v = getenv("WFA_PAIR_SEARCH");
  return !v || v[0] != '0';
}

// LDS for the four sequences of a breakpoint problem (2-bit sequences of a
// 100 kbp pair take 100 KB of gfx950's 160 KB per CU).
constexpr size_t kBreakpointSeqLdsBytes = 120 * 1024;

// Runs the breakpoint search of `nodes` (grouped by size) and returns the outcomes.
bool run_breakpoints(wavefront_aligner_t* wf, Engine* e, wfa_device_batch_t* b, const std::vector<Node>& nodes,
                     std::vector<wfa::BiOutcome>& out, RunTotals* tot) {
  const wfa::Penalties pen = make_penalties(wf->penalties);
  const int ncomp = pen.ncomp;
  out.assign(nodes.size(), wfa::BiOutcome{});
  if (nodes.empty()) return true;
  // groups by workgroup size
  std::vector<int> groups[3];
  for (int i = 0; i < (int)nodes.size(); ++i) {
    const wfa::BiProblem& p = nodes[i].prob;
    const int size = (p.pe - p.pb) + (p.te - p.tb);
    const int bl = bp_block(size);
    groups[bl == 64 ? 0 : (bl == 256 ? 1 : 2)].push_back(i);
  }
  hipStream_t st = e->stream;
  const wfa::HeurArgs heur = heuristic_of(wf);
  for (int gi = 0; gi < 3; ++gi) {
    std::vector<int> ids = groups[gi];
    if (ids.empty()) continue;
    const int block = gi == 0 ? 64 : (gi == 1 ? 256 : 1024);
    // One launch over the sub-problems `sel` (indices into nodes); paired:
    // two workgroups per sub-problem.  Outcomes go to out[].
    auto launch = [&](const std::vector<int>& sel, bool paired) -> bool {
      std::vector<wfa::BiProblem> probs(sel.size());
      int wmax = 0;
      for (size_t j = 0; j < sel.size(); ++j) {
        probs[j] = nodes[sel[j]].prob;
        wmax = std::max(wmax, (probs[j].pe - probs[j].pb) + (probs[j].te - probs[j].tb));
      }
      const int wcap = wmax + 8;
      // The four sequences of the launch's longest problem in LDS when they fit.
      const int spw = 64 / b->bits;
      int lw = 0;
      for (const wfa::BiProblem& p : probs) lw = std::max(lw, std::max(p.pe - p.pb, p.te - p.tb) / spw + 3);
      if (wfa::breakpoint_lds_bytes(lw) > kBreakpointSeqLdsBytes) lw = 0;
      int grid = 0;
      wfa::Penalties lpen = pen;
      if (paired) {
        lpen.nslots = pen.nslots + 1;  // room for the speculative reverse step
        grid = 2 * (int)probs.size();
      } else {
        int per_cu = wfa::breakpoint_max_blocks_per_cu(ncomp, b->bits, block, heur.cut() != 0,
                                                       wfa::breakpoint_lds_bytes(lw));
        if (per_cu <= 0 && lw > 0) {
          lw = 0;
          per_cu = wfa::breakpoint_max_blocks_per_cu(ncomp, b->bits, block, heur.cut() != 0);
        }
        if (per_cu <= 0) per_cu = 1;
        grid = (int)std::min<int64_t>((int64_t)probs.size(), (int64_t)per_cu * e->num_cus);
      }
      const long long stride = 2LL * lpen.nslots * ncomp * wcap;
      const size_t ring_units = paired ? probs.size() : (size_t)grid;  // teams or workgroups
      DevBuf& d_probs = e->bi.probs;
      DevBuf& d_out = e->bi.out;
      const size_t hand_ints = wfa::breakpoint_pair_handover_ints(ncomp);
      if (!d_probs.ensure(probs.size() * sizeof(wfa::BiProblem)) ||
          !d_out.ensure(probs.size() * sizeof(wfa::BiOutcome)) ||
          !e->ring.ensure((size_t)stride * ring_units * sizeof(int)) ||
          (paired && (!e->bi.mail.ensure(probs.size() * 4 * sizeof(unsigned long long)) ||
                      !e->bi.hand.ensure(probs.size() * hand_ints * sizeof(int))))) {
        fprintf(stderr, "[WFA] cannot allocate BiWFA breakpoint buffers\n");
        return false;
      }
      wfa::Counters* c = counters_of(e);
      wfa::BiArgs a{};
      a.n = (int)probs.size();
      a.probs = d_probs.as<wfa::BiProblem>();
      a.words = b->words.as<uint64_t>();
      a.rev_words = b->rev_words.as<uint64_t>();
      a.fwd_ppos = b->d_patw.as<int64_t>();
      a.fwd_tpos = b->d_txtw.as<int64_t>();
      a.rev_ppos = b->d_patw.as<int64_t>();
      a.rev_tpos = b->d_txtw.as<int64_t>();
      a.plen_total = b->d_plen.as<int>();
      a.tlen_total = b->d_tlen.as<int>();
      a.out = d_out.as<wfa::BiOutcome>();
      a.ring = e->ring.as<int>();
      a.ring_stride = stride;
      a.wcap = wcap;
      a.queue = &c->queue;
      a.dsteps_total = &c->dsteps;
      a.max_steps = wf->system.max_alignment_steps > 0 ? wf->system.max_alignment_steps : INT_MAX;
      a.lmax_words = lw;
      a.heur = heur;  // the breakpoint aligners inherit the heuristic (wavefront_bialigner.c:53)
      a.pen = lpen;
      a.mail = paired ? e->bi.mail.as<unsigned long long>() : nullptr;
      a.handover = paired ? e->bi.hand.as<int>() : nullptr;
      a.pair_wait_ticks = pair_wait_ticks();
      for (auto& ev : e->bi.ev)
        if (!ev && !hip_ok(hipEventCreate(&ev), "hipEventCreate")) return false;
      hipEvent_t e0 = e->bi.ev[0], e1 = e->bi.ev[1];
      std::vector<wfa::BiOutcome> res(probs.size());
      wfa::Counters host_c{};
      bool ok = hip_ok(hipMemcpyAsync(d_probs.p, probs.data(), probs.size() * sizeof(wfa::BiProblem),
                                      hipMemcpyHostToDevice, st), "bi upload") &&
                hip_ok(wfa::launch_reset_counters(c, 2, st), "reset") &&
                (!paired || hip_ok(hipMemsetAsync(a.mail, 0, probs.size() * 4 * sizeof(unsigned long long), st),
                                   "mailboxes"));
      (void)hipEventRecord(e0, st);
      ok = ok && hip_ok(paired ? wfa::launch_breakpoint_pair(a, ncomp, b->bits, block, st)
                               : wfa::launch_breakpoint(a, ncomp, b->bits, block, grid, st),
                        "breakpoint launch");
      (void)hipEventRecord(e1, st);
      ok = ok && hip_ok(hipMemcpyAsync(res.data(), d_out.p, res.size() * sizeof(wfa::BiOutcome),
                                       hipMemcpyDeviceToHost, st), "bi download") &&
           hip_ok(hipMemcpyAsync(&host_c, e->counters.p, sizeof(host_c), hipMemcpyDeviceToHost, st), "counters") &&
           hip_ok(hipStreamSynchronize(st), "breakpoint");
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (!ok) return false;
      tot->forward_ms += ms;
      tot->launches += 1;
      tot->dsteps += host_c.dsteps;
      for (size_t j = 0; j < sel.size(); ++j) out[sel[j]] = res[j];
      return true;
    };
    // Paired workgroups when every team fits on the device at once beside
    // nothing else of this launch (large sub-problems, fewer than half the
    // workgroups the CUs hold); not with a heuristic cut-off, or where the
    // extra ring slot would pass the 64 slots of the slot metadata.
    bool paired = false;
    if (block == 1024 && heur.cut() == 0 && pen.nslots + 1 <= 64 && pair_searches_enabled()) {
      const int spw = 64 / b->bits;
      int lw = 0;
      for (int i : ids) {
        const wfa::BiProblem& p = nodes[i].prob;
        lw = std::max(lw, std::max(p.pe - p.pb, p.te - p.tb) / spw + 3);
      }
      if (wfa::breakpoint_lds_bytes(lw) > kBreakpointSeqLdsBytes) lw = 0;
      const int per_cu = wfa::breakpoint_pair_max_blocks_per_cu(ncomp, b->bits, block, wfa::breakpoint_lds_bytes(lw));
      paired = per_cu > 0 && 2 * (int64_t)ids.size() <= (int64_t)per_cu * e->num_cus;
    }
    if (!launch(ids, paired)) return false;
    if (paired) {
      // a team whose partner did not answer in time: search again on one workgroup
      std::vector<int> again;
      for (int i : ids)
        if (out[i].status == wfa::kBiPairAbort) again.push_back(i);
      if (!again.empty()) {
        fprintf(stderr, "[WFA] %zu paired breakpoint searches timed out; searching them again on one workgroup\n",
                again.size());
        if (!launch(again, false)) return false;
      }
    }
  }
  return true;
}

}  // namespace

int bialign_run(wavefront_aligner_t* wf, Engine* e, wfa_device_batch_t* b, RunTotals* tot) {
  const int n = b->n;
  const int records = b->records;
  if (!ensure_reverse(e, b)) return WF_STATUS_OOM;
  std::vector<PairState> ps(n);
  std::vector<Node> nodes, next;
  std::vector<wfa::BiOutcome> outc;
  std::vector<Node> base_nodes;
  for (int i = 0; i < n; ++i) {
    Node nd;
    nd.prob = wfa::BiProblem{i, 0, b->plen[i], 0, b->tlen[i], wfa::kM, wfa::kM, 0};
    nd.level = 0;
    // wavefront_bialign, wavefront_bialign.c:1276-1284
    nd.score_remaining = std::max(b->plen[i], b->tlen[i]) <= kFallbackMinLength ? 0 : INT_MAX;
    nodes.push_back(nd);
  }
  if (!records) {
    // wavefront_bialign_compute_score (:1226-1262): the top-level breakpoint only.
    if (!run_breakpoints(wf, e, b, nodes, outc, tot)) return WF_STATUS_OOM;
    for (int i = 0; i < n; ++i) {
      const wfa::BiOutcome& o = outc[i];
      if (o.status == wfa::kBiOk || o.status == wfa::kBiEndReached) {
        const int s = (o.status == wfa::kBiOk) ? o.score : (o.fwd_end_score >= 0 ? o.fwd_end_score : o.rev_end_score);
        ps[i].score = classic_score(wf->penalties, b->plen[i], b->tlen[i], s);
      } else {
        ps[i].status = (o.status == wfa::kDevMaxSteps) ? WF_STATUS_MAX_STEPS_REACHED : WF_STATUS_UNATTAINABLE;
      }
    }
  } else {
    while (!nodes.empty()) {
      next.clear();
      std::vector<Node> bp_nodes;
      for (const Node& nd : nodes) {
        PairState& st = ps[nd.prob.pair];
        // After a failure the reference's depth-first recursion has still
        // run everything to its left (whose base alignments set the score).
        if (st.status != WF_STATUS_ALG_COMPLETED &&
            !(nd.prob.pb < st.fail_pb || (nd.prob.pb == st.fail_pb && nd.prob.tb < st.fail_tb)))
          continue;
        const wfa::BiProblem& p = nd.prob;
        const int pl = p.pe - p.pb, tl = p.te - p.tb;
        if (tl == 0) {
          st.leaves.push_back(Leaf{p.pb, p.tb, 'D', pl, -1});
        } else if (pl == 0) {
          st.leaves.push_back(Leaf{p.pb, p.tb, 'I', tl, -1});
        } else if (nd.score_remaining <= kFallbackMinScore) {
          base_nodes.push_back(nd);
        } else {
          bp_nodes.push_back(nd);
        }
      }
      if (!run_breakpoints(wf, e, b, bp_nodes, outc, tot)) return WF_STATUS_OOM;
      for (size_t j = 0; j < bp_nodes.size(); ++j) {
        const Node& nd = bp_nodes[j];
        const wfa::BiProblem& p = nd.prob;
        const wfa::BiOutcome& o = outc[j];
        PairState& st = ps[p.pair];
        if (o.status == wfa::kBiOk) {
          const int bp_h = o.offset_forward;
          const int bp_v = o.offset_forward - o.k_forward;
          Node left, right;
          left.prob = wfa::BiProblem{p.pair, p.pb, p.pb + bp_v, p.tb, p.tb + bp_h, p.comp_begin, o.component, 0};
          left.score_remaining = o.score_forward;
          left.level = nd.level + 1;
          right.prob = wfa::BiProblem{p.pair, p.pb + bp_v, p.pe, p.tb + bp_h, p.te, o.component, p.comp_end, 0};
          right.score_remaining = o.score_reverse;
          right.level = nd.level + 1;
          next.push_back(left);
          next.push_back(right);
          if (nd.level == 0) st.score = classic_score(wf->penalties, p.pe - p.pb, p.te - p.tb, o.score);
        } else if (o.status == wfa::kBiEndReached) {
          // wavefront_bialign_find_breakpoint_exception (:1084-1111)
          const int reached = o.fwd_end_score >= 0 ? o.fwd_end_score : o.rev_end_score;
          if (reached <= kRecoveryMinScore) base_nodes.push_back(nd);
          else fail_pair(st, WF_STATUS_UNATTAINABLE, p.pb, p.tb);
        } else if (o.status == wfa::kDevMaxSteps) {
          fail_pair(st, WF_STATUS_MAX_STEPS_REACHED, p.pb, p.tb);
        } else if (o.status == wfa::kDevUnreachable) {
          // a side the cut-off left unreachable (wavefront_bialign.c:1084-1111, 1286-1294)
          fail_pair(st, WF_STATUS_UNATTAINABLE, p.pb, p.tb);
        } else {
          if (o.status != wfa::kDevMaxSteps)
            fprintf(stderr, "[WFA] BiWFA breakpoint search failed (status %d) for pair %d\n", o.status, p.pair);
          fail_pair(st, WF_STATUS_UNATTAINABLE, p.pb, p.tb);
        }
      }
      nodes.swap(next);
    }
    // Base cases: one pass of the unidirectional kernels over all of them.
    const int nb = (int)base_nodes.size();
    std::vector<int> bpl(nb), btl(nb), bstatus, bscore(nb), bbegin(nb);
    std::vector<int64_t> bpp(nb), btp(nb), boff(nb);
    std::vector<signed char> bcb(nb), bce(nb);
    int64_t bops = 0;
    for (int j = 0; j < nb; ++j) {
      const wfa::BiProblem& p = base_nodes[j].prob;
      bpl[j] = p.pe - p.pb;
      btl[j] = p.te - p.tb;
      bpp[j] = b->pat_pos[p.pair] + p.pb;
      btp[j] = b->txt_pos[p.pair] + p.tb;
      bcb[j] = (signed char)p.comp_begin;
      bce[j] = (signed char)p.comp_end;
      boff[j] = bops;
      bops += (int64_t)bpl[j] + btl[j];
    }
    Engine::BiBuffers& bb = e->bi;
    DevBuf &d_pl = bb.pl, &d_tl = bb.tl, &d_pp = bb.pp, &d_tp = bb.tp, &d_cb = bb.cb, &d_ce = bb.ce,
           &d_st = bb.st, &d_sc = bb.sc, &d_last = bb.last, &d_ds = bb.ds, &d_off = bb.off, &d_beg = bb.beg,
           &d_ops = bb.ops;
    hipStream_t st = e->stream;
    if (nb > 0) {
      const size_t m = (size_t)nb;
      bool ok = d_pl.ensure(m * 4) && d_tl.ensure(m * 4) && d_pp.ensure(m * 8) && d_tp.ensure(m * 8) &&
                d_cb.ensure(m) && d_ce.ensure(m) && d_st.ensure(m * 4) && d_sc.ensure(m * 4) && d_last.ensure(m * 4) &&
                d_ds.ensure(m * 8) && d_off.ensure(m * 8) && d_beg.ensure(m * 4) &&
                d_ops.ensure(std::max<int64_t>(bops, 1));
      if (!ok) return WF_STATUS_OOM;
      ok = hip_ok(hipMemcpyAsync(d_pl.p, bpl.data(), m * 4, hipMemcpyHostToDevice, st), "base upload") &&
           hip_ok(hipMemcpyAsync(d_tl.p, btl.data(), m * 4, hipMemcpyHostToDevice, st), "base upload") &&
           hip_ok(hipMemcpyAsync(d_pp.p, bpp.data(), m * 8, hipMemcpyHostToDevice, st), "base upload") &&
           hip_ok(hipMemcpyAsync(d_tp.p, btp.data(), m * 8, hipMemcpyHostToDevice, st), "base upload") &&
           hip_ok(hipMemcpyAsync(d_cb.p, bcb.data(), m, hipMemcpyHostToDevice, st), "base upload") &&
           hip_ok(hipMemcpyAsync(d_ce.p, bce.data(), m, hipMemcpyHostToDevice, st), "base upload") &&
           hip_ok(hipMemcpyAsync(d_off.p, boff.data(), m * 8, hipMemcpyHostToDevice, st), "base upload");
      if (!ok) return WF_STATUS_OOM;
      JobSet js;
      js.n = nb;
      js.bits = b->bits;
      js.h_plen = bpl.data();
      js.h_tlen = btl.data();
      js.words = b->words.as<uint64_t>();
      js.plen = d_pl.as<int>();
      js.tlen = d_tl.as<int>();
      js.ppos = d_pp.as<int64_t>();
      js.tpos = d_tp.as<int64_t>();
      js.cbeg = d_cb.as<signed char>();
      js.cend = d_ce.as<signed char>();
      js.status = d_st.as<int>();
      js.score = d_sc.as<int>();
      js.last = d_last.as<uint32_t>();
      js.dsteps = d_ds.as<long long>();
      js.ops = d_ops.as<char>();
      js.ops_off = d_off.as<int64_t>();
      js.ops_begin = d_beg.as<int>();
      std::vector<int> ids(nb);
      for (int j = 0; j < nb; ++j) ids[j] = j;
      std::sort(ids.begin(), ids.end(), [&](int x, int y) { return bpl[x] + btl[x] > bpl[y] + btl[y]; });
      std::vector<PlanGroup> plan = make_plan(js, ids, make_penalties(wf->penalties));
      RunTotals bt;
      const int rc = run_jobs(wf, e, js, plan, 1, &bt, bstatus);
      give_back_groups(e, plan);
      if (rc != WF_STATUS_OK) return rc;
      tot->forward_ms += bt.forward_ms;
      tot->backtrace_ms += bt.backtrace_ms;
      tot->launches += bt.launches;
      tot->dsteps += bt.dsteps;
      bstatus.resize(m);
      ok = hip_ok(hipMemcpyAsync(bstatus.data(), d_st.p, m * 4, hipMemcpyDeviceToHost, st), "base download") &&
           hip_ok(hipMemcpyAsync(bscore.data(), d_sc.p, m * 4, hipMemcpyDeviceToHost, st), "base download") &&
           hip_ok(hipMemcpyAsync(bbegin.data(), d_beg.p, m * 4, hipMemcpyDeviceToHost, st), "base download") &&
           hip_ok(hipStreamSynchronize(st), "base download");
      if (!ok) return WF_STATUS_OOM;
      for (int j = 0; j < nb; ++j) {
        const Node& nd = base_nodes[j];
        const wfa::BiProblem& p = nd.prob;
        PairState& s = ps[p.pair];
        if (bstatus[j] != WF_STATUS_ALG_COMPLETED) {
          // wavefront_bialign_base (:159-190): anything but completed is
          // unattainable; the base's own score stays (the step limit's)
          fail_pair(s, WF_STATUS_UNATTAINABLE, p.pb, p.tb, true,
                    bstatus[j] == WF_STATUS_MAX_STEPS_REACHED ? -wf->system.max_alignment_steps : INT32_MIN);
          continue;
        }
        s.leaves.push_back(Leaf{p.pb, p.tb, 0, bpl[j] + btl[j] - bbegin[j], j});
        s.bases.push_back(BaseScore{p.pb, p.tb, classic_score(wf->penalties, bpl[j], btl[j], bscore[j])});
        // the base call sets cigar->score; at level 0 nothing overrides it
        if (nd.level == 0) s.score = classic_score(wf->penalties, bpl[j], btl[j], bscore[j]);
      }
    }
    // Assemble every pair right-aligned in its slot of the batch's operations:
    // all its pieces, or for a failed pair the ones the reference's recursion
    // appended before the failure (the cigar it leaves, wavefront_bialign.c:
    // 159-190, 1145-1222).
    std::vector<wfa::OpsPiece> pieces;
    std::vector<int> begin(n, 0);
    for (int i = 0; i < n; ++i) {
      PairState& s = ps[i];
      const int cap = b->plen[i] + b->tlen[i];
      if (s.status != WF_STATUS_ALG_COMPLETED) {
        std::vector<Leaf> kept;
        for (const Leaf& l : s.leaves)
          if (l.pb < s.fail_pb || (l.pb == s.fail_pb && l.tb < s.fail_tb)) kept.push_back(l);
        s.leaves.swap(kept);
      }
      std::sort(s.leaves.begin(), s.leaves.end(),
                [](const Leaf& x, const Leaf& y) { return x.pb != y.pb ? x.pb < y.pb : x.tb < y.tb; });
      int64_t total = 0;
      for (const Leaf& l : s.leaves) total += l.len;
      if (total > cap) {
        fprintf(stderr, "[WFA] internal error: BiWFA CIGAR longer than its slot for pair %d\n", i);
        s.status = WF_STATUS_UNATTAINABLE;
        begin[i] = cap;
        continue;
      }
      begin[i] = (int)(cap - total);
      int64_t dst = b->ops_off[i] + begin[i];
      for (const Leaf& l : s.leaves) {
        if (l.len == 0) continue;
        wfa::OpsPiece pc{};
        pc.dst = dst;
        pc.len = l.len;
        if (l.base >= 0) {
          pc.src = boff[l.base] + bbegin[l.base];
          pc.op = 0;
        } else {
          pc.src = -1;
          pc.op = l.op;
        }
        pieces.push_back(pc);
        dst += l.len;
      }
    }
    if (!pieces.empty()) {
      DevBuf& d_pieces = e->bi.pieces;
      if (!d_pieces.ensure(pieces.size() * sizeof(wfa::OpsPiece))) return WF_STATUS_OOM;
      bool ok = hip_ok(hipMemcpyAsync(d_pieces.p, pieces.data(), pieces.size() * sizeof(wfa::OpsPiece),
                                      hipMemcpyHostToDevice, st), "pieces") &&
                hip_ok(wfa::launch_gather_ops(d_pieces.as<wfa::OpsPiece>(), (int)pieces.size(),
                                              d_ops.p ? d_ops.as<char>() : nullptr, b->d_ops.as<char>(), st), "gather") &&
                hip_ok(hipStreamSynchronize(st), "gather");
      if (!ok) return WF_STATUS_OOM;
    }
    if (!hip_ok(hipMemcpyAsync(b->d_ops_begin.p, begin.data(), (size_t)n * 4, hipMemcpyHostToDevice, e->stream),
                "results") ||
        !hip_ok(hipStreamSynchronize(e->stream), "results"))
      return WF_STATUS_OOM;
  }
  // Per-pair results into the batch (status on the device, scores on the host).
  std::vector<int> status(n);
  b->bi_score.assign(n, INT32_MIN);
  for (int i = 0; i < n; ++i) {
    status[i] = ps[i].status;
    b->bi_score[i] = (records && status[i] != WF_STATUS_ALG_COMPLETED) ? failed_score(ps[i]) : ps[i].score;
    if (status[i] != WF_STATUS_ALG_COMPLETED) tot->fail++;
  }
  if (n > 0 && (!hip_ok(hipMemcpyAsync(b->d_status.p, status.data(), (size_t)n * 4, hipMemcpyHostToDevice, e->stream),
                        "results") ||
                 !hip_ok(hipStreamSynchronize(e->stream), "results")))
    return WF_STATUS_OOM;
  return WF_STATUS_OK;
}

}  // namespace wfa_host
