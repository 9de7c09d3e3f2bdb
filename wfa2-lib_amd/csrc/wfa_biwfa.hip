// BiWFA (memory mode ultralow) breakpoint search on the device.
//
// One workgroup per sub-problem runs wavefront_bialign_find_breakpoint
// (wavefront_bialign.c:975-1083): a forward and a reverse score-only
// wavefront (rings of the last max_score_scope scores, in HBM) advanced in the
// reference's alternation order, phase 1 until the furthest antidiagonals of
// the two sides meet (wavefront_extend_end2end_max, wavefront_extend.c:173-211),
// then phase 2 with the overlap scans of wavefront_bialign_overlap
// (:878-956): for each of the last scores of the other side, D2, I2, D1, I1
// (gap-opening adjusted) and then M, each taking the first diagonal k0 in
// ascending order (breakpoint_indel2indel :477-573, breakpoint_m2m :798-874).
// Each scan is a block-parallel min-reduction over k0.  The host recurses on
// the two halves, level by level across all pairs (wfa_bialign.cpp).
#include "wfa_common.h"

namespace wfa {
namespace {

template <int NC, bool CUT>
struct Side {
  MetaRegs<NC, CUT> meta;
  int steps_wait;          // cut-off countdown (wavefront_heuristic.c:506-521)
  int nulls;               // null steps in a row (align_status.num_null_steps)
  DropState xs;           // x-drop's best score (wavefront_heuristic.c:114-121)
  bool dead;               // indel / edit: the cut-off emptied the wavefront
  int cur;                 // ring slot of the latest multiple-of-g score
  int* ring;
  const uint64_t* P;       // sequences in global memory ...
  int pb;
  const uint64_t* T;
  int tb;
  int lds_p, lds_t;        // ... or, when >= 0, at these words of the LDS sequence area
  int cb, ce;              // begin / end component of this direction
  int end_score;           // score at which this side reached its end, -1 if not
};

struct Bp {  // wf_bialign_breakpoint_t
  int score, score_forward, score_reverse;
  int k_forward, k_reverse, offset_forward, offset_reverse;
  int component;
};

// Dynamic LDS of the breakpoint kernels: the sequences of the sub-problem.
extern __shared__ __align__(16) unsigned char bp_smem[];

// Extension of diagonal k from offset h on one side.  With the sequences in
// LDS the windows are read through an LDS pointer (ds_read); a generic
// pointer would make every window a flat load that waits on both counters.
template <int BITS, int NC, bool CUT>
__device__ __forceinline__ int side_lcp(const Side<NC, CUT>& sd, int plen, int tlen, int k, int h) {
  if (sd.lds_p >= 0) {
    const uint64_t* W = reinterpret_cast<const uint64_t*>(bp_smem);
    return lcp_extend<BITS>(W + sd.lds_p, sd.pb, W + sd.lds_t, sd.tb, plen, tlen, k, h);
  }
  return lcp_extend<BITS>(sd.P, sd.pb, sd.T, sd.tb, plen, tlen, k, h);
}

// The stored wavefront of component c at score s of a side, if present:
// its trimmed range and the ring offset of diagonal 0.
template <int NC>
struct WfRef {
  bool present;
  int lo, hi, base;  // value at k is ring[base + k]
};

template <int NC, bool CUT>
__device__ __forceinline__ WfRef<NC> wf_ref(const Side<NC, CUT>& sd, int s, int c, const Penalties& pen, int wcap) {
  WfRef<NC> r{false, 1, 0, 0};
  if (s < 0 || (s % pen.g) != 0) return r;
  // the slot of score s, counted back from the latest multiple of g (slot cur)
  const int tag_cur = lane_of(sd.meta.tag, sd.cur);
  const int back = (tag_cur - s) / pen.g;
  if (back < 0 || back >= pen.nslots) return r;
  int slot = sd.cur - back;
  if (slot < 0) slot += pen.nslots;
  slot = uniform(slot);
  if (lane_of(sd.meta.tag, slot) != s) return r;
  const unsigned fl = (unsigned)lane_of((int)sd.meta.flags, slot);
  if (!(fl & (1u << c))) return r;
  r.present = true;
  // the overlap scans the range a cut-off left (wavefront_bialign.c:477-573)
  if constexpr (CUT) {
    r.lo = lane_of(sd.meta.cut.lo[c], slot);
    r.hi = lane_of(sd.meta.cut.hi[c], slot);
  } else {
    r.lo = lane_of(sd.meta.lo[c], slot);
    r.hi = lane_of(sd.meta.hi[c], slot);
  }
  r.base = (slot * NC + c) * wcap - lane_of(sd.meta.alo, slot);
  return r;
}

// Heuristic cut-off of the side's M wavefront of score s in ring slot
// `slot`, trimmed to [lo, hi] with array origin alo
// (wavefront_heuristic_cufoff, wavefront_heuristic.c:506-561): wf-adaptive
// or wfmash, x-drop or z-drop, then the band (static clamp or adaptive
// window), one countdown for all; when M changed, I1/D1/I2/D2 are narrowed
// to M's range; a z-drop returns true (the side's end is then unreachable);
// indel / edit then read M only inside the cut range (their compute step
// NULLs the two cells beyond each end, wavefront_compute_edit.c:355-358).
template <int NC, int BLOCK, bool CUT>
__device__ bool side_cutoff(Side<NC, CUT>& sd, const HeurArgs& heur, BlockRed<BLOCK>& red, int slot, int alo, int lo,
                            int hi, int s, int plen, int tlen, int wcap, int metric) {
  if constexpr (CUT) {
    --sd.steps_wait;
    const int* m = sd.ring + (slot * NC + kM) * wcap - alo;
    int clo = lo, chi = hi;
    if (heur.wfadaptive && sd.steps_wait <= 0 && hi - lo + 1 >= heur.min_len) {
      wfadaptive_cut<BLOCK>(m, lo, hi, plen, tlen, heur.thr, heur.wfadaptive == 2, red, &clo, &chi);
      sd.steps_wait = heur.steps;
    }
    if (heur.xdrop && sd.steps_wait <= 0) {
      xdrop_cut<BLOCK>(m, s, heur.swg_match, heur.xdrop_thr, red, sd.xs, &clo, &chi);
      sd.steps_wait = heur.steps;
    } else if (heur.zdrop && sd.steps_wait <= 0) {
      if (zdrop_check<BLOCK>(m, clo, chi, s, heur.swg_match, heur.zdrop_thr, red, sd.xs)) return true;
      sd.steps_wait = heur.steps;
    }
    if (heur.banded == 1) {
      clo = max(clo, heur.band_lo);
      chi = min(chi, heur.band_hi);
    } else if (heur.banded == 2) {
      banded_adaptive_cut(m, heur.band_lo, heur.band_hi, plen, tlen, sd.steps_wait, heur.steps, &clo, &chi);
    }
    // unchanged M: nothing else is narrowed either (wavefront_heuristic.c:540-541)
    if (clo == lo && chi == hi) return false;
    if ((threadIdx.x & 63) == slot) {
      unsigned fl = sd.meta.flags;
      sd.meta.cut.lo[kM] = clo;
      sd.meta.cut.hi[kM] = chi;
      if (clo > chi) fl |= 0x100u;
      if (metric <= kEdit) { sd.meta.lo[kM] = clo; sd.meta.hi[kM] = chi; }
#pragma unroll
      for (int c = 1; c < NC; ++c) {
        if ((fl & (1u << c)) && !(fl & (0x100u << c))) {
          sd.meta.cut.lo[c] = max(sd.meta.cut.lo[c], clo);
          sd.meta.cut.hi[c] = min(sd.meta.cut.hi[c], chi);
          if (sd.meta.cut.lo[c] > sd.meta.cut.hi[c]) fl |= 0x100u << c;
        }
      }
      sd.meta.flags = fl;
    }
  }
  return false;
}

// Compute score s of one side (no backtrace records) fused with its
// extension; returns 1 when the end is reached (wavefront_termination.c),
// 2 when the cut-off left it unreachable (more null steps in a row than the
// score scope, wavefront_extend.c:184-191), -1 on an internal error, 0
// otherwise.  *max_ak gets the largest antidiagonal 2*offset-k of the
// extended M wavefront (0 when absent).
// The per-wave ranges of a side step, double-buffered: consecutive computed
// side steps of a workgroup write alternate halves, so one barrier per step
// separates a wave's writes from every wave's merge (a wave already in the
// next step writes the other half).
template <int NC, int NW>
struct RedRows {
  int (*buf)[NW][2 * NC + 2];  // two halves
  int par;                     // computed steps so far (uniform)
};

template <int NC, int BLOCK, int BITS, bool CUT>
__device__ int side_step(Side<NC, CUT>& sd, int s, int plen, int tlen, const Penalties& pen, const HeurArgs& heur, int wcap,
                         bool want_max, int* max_ak, long long* dsteps, RedRows<NC, BLOCK / 64>& rr,
                         BlockRed<BLOCK>& red) {
  constexpr int NW = BLOCK / 64;
  const int tid = threadIdx.x, lane = tid & 63, wid = uniform(tid >> 6);
  const int R = pen.nslots, g = pen.g;
  *max_ak = 0;
  if (s % g) {  // only multiples of g can be non-null
    if constexpr (CUT) {
      if (++sd.nulls > pen.scope) return 2;
    }
    return 0;
  }
  sd.cur = (sd.cur + 1 == R) ? 0 : sd.cur + 1;
  const int cur = sd.cur;
  const int dq_x = pen.x / g, dq_o1 = (pen.o1 + pen.e1) / g, dq_e1 = pen.e1 / g;
  Src mx{1, -1, kFarDiagonal, 0, 0}, i1 = mx, d1 = mx, mo2 = mx, i2 = mx, d2 = mx;
  if (NC > 1 || pen.metric != kIndel) mx = get_src(sd.meta, s - pen.x, cur - dq_x, R, kM, wcap);
  const Src mo1 = get_src(sd.meta, s - pen.o1 - pen.e1, cur - dq_o1, R, kM, wcap);
  if constexpr (NC > 1) {
    i1 = get_src(sd.meta, s - pen.e1, cur - dq_e1, R, kI1, wcap);
    d1 = get_src(sd.meta, s - pen.e1, cur - dq_e1, R, kD1, wcap);
  }
  if constexpr (NC == 5) {
    const int dq_o2 = (pen.o2 + pen.e2) / g, dq_e2 = pen.e2 / g;
    mo2 = get_src(sd.meta, s - pen.o2 - pen.e2, cur - dq_o2, R, kM, wcap);
    i2  = get_src(sd.meta, s - pen.e2, cur - dq_e2, R, kI2, wcap);
    d2  = get_src(sd.meta, s - pen.e2, cur - dq_e2, R, kD2, wcap);
  }
  const bool any = mx.valid() || mo1.valid() || i1.valid() || d1.valid() || mo2.valid() || i2.valid() || d2.valid();
  // indel / edit: only a cut-off empties the previous wavefront, and every
  // later one is then empty too (wavefront_compute_edit.c:344-352)
  if (NC == 1 && pen.metric <= kEdit && !mo1.valid()) {
    if constexpr (!CUT) return -1;
    sd.dead = true;
    if (lane == cur) { sd.meta.tag = s; sd.meta.flags = 0u; }
    return 0;
  }
  if (!any) {
    if constexpr (CUT) {
      if (++sd.nulls > pen.scope) return 2;
    }
    if (lane == cur) { sd.meta.tag = s; sd.meta.flags = 0u; }
    return 0;
  }
  if constexpr (CUT) sd.nulls = 0;
  int lo, hi;
  if (NC == 1 && pen.metric <= kEdit) {
    lo = mo1.lim_lo - 1;
    hi = mo1.lim_hi + 1;
  } else {
    lo = min(mx.lim_lo, mo1.lim_lo - 1);
    hi = max(mx.lim_hi, mo1.lim_hi + 1);
  }
  if constexpr (NC > 1) {
    lo = min(lo, i1.lim_lo + 1);  hi = max(hi, i1.lim_hi + 1);
    lo = min(lo, d1.lim_lo - 1);  hi = max(hi, d1.lim_hi - 1);
  }
  if constexpr (NC == 5) {
    lo = min(lo, mo2.lim_lo - 1); hi = max(hi, mo2.lim_hi + 1);
    lo = min(lo, i2.lim_lo + 1);  hi = max(hi, i2.lim_hi + 1);
    lo = min(lo, d2.lim_lo - 1);  hi = max(hi, d2.lim_hi - 1);
  }
  lo = uniform(lo);
  hi = uniform(hi);
  if (hi - lo + 1 > wcap) return -1;
  *dsteps += hi - lo + 1;
  const bool prune_try = NC == 1 && pen.metric == kEdit && hi - lo + 1 >= kEditPruneWidth;
  int wmin = INT_MAX;
  const bool has_i1 = (NC > 1) && (mo1.valid() || i1.valid());
  const bool has_d1 = (NC > 1) && (mo1.valid() || d1.valid());
  const bool has_i2 = (NC == 5) && (mo2.valid() || i2.valid());
  const bool has_d2 = (NC == 5) && (mo2.valid() || d2.valid());
  int* ring = sd.ring;
  int* outM  = ring + (cur * NC + kM) * wcap - lo;
  int* outI1 = ring + (cur * NC + (NC > 1 ? kI1 : 0)) * wcap - lo;
  int* outD1 = ring + (cur * NC + (NC > 1 ? kD1 : 0)) * wcap - lo;
  int* outI2 = ring + (cur * NC + (NC == 5 ? kI2 : 0)) * wcap - lo;
  int* outD2 = ring + (cur * NC + (NC == 5 ? kD2 : 0)) * wcap - lo;
  WaveRange rM, rI1, rD1, rI2, rD2;
  rM.init(); rI1.init(); rD1.init(); rI2.init(); rD2.init();
  int best_ak = 0;
  for (int kb = lo; kb <= hi; kb += BLOCK) {
    const int k = kb + tid;
    bool vM = false, vI1 = false, vD1 = false, vI2 = false, vD2 = false;
    if constexpr (NC == 1) {
      if (k <= hi) {
        // indel / edit / gap-linear (wavefront_compute_edit.c:44-102,
        // wavefront_compute_linear.c:44-76)
        int m = max(rd(ring, mo1, k + 1), max(rd(ring, mo1, k - 1), rd(ring, mx, k)) + 1);
        if (!in_bounds(m, k, plen, tlen)) m = kOffsetNull;
        if (m != kOffsetNull) {
          if (prune_try) wmin = min(wmin, max(plen - (m - k), tlen - m));
          m = side_lcp<BITS>(sd, plen, tlen, k, m);
          best_ak = max(best_ak, 2 * m - k);
        }
        outM[k] = m;
        vM = (m != kOffsetNull);
      }
    } else if (k <= hi) {
      const int ins1 = max(rd(ring, mo1, k - 1), rd(ring, i1, k - 1)) + 1;
      const int del1 = max(rd(ring, mo1, k + 1), rd(ring, d1, k + 1));
      int m = max(del1, max(rd(ring, mx, k) + 1, ins1));
      int ins2 = kOffsetNull, del2 = kOffsetNull;
      if constexpr (NC == 5) {
        ins2 = max(rd(ring, mo2, k - 1), rd(ring, i2, k - 1)) + 1;
        del2 = max(rd(ring, mo2, k + 1), rd(ring, d2, k + 1));
        m = max(m, max(ins2, del2));
      }
      if (!in_bounds(m, k, plen, tlen)) m = kOffsetNull;
      if (m != kOffsetNull) {
        m = side_lcp<BITS>(sd, plen, tlen, k, m);
        best_ak = max(best_ak, 2 * m - k);
      }
      outM[k] = m;
      vM = (m != kOffsetNull);
      if (has_i1) { outI1[k] = ins1; vI1 = in_bounds(ins1, k, plen, tlen); }
      if (has_d1) { outD1[k] = del1; vD1 = in_bounds(del1, k, plen, tlen); }
      if constexpr (NC == 5) {
        if (has_i2) { outI2[k] = ins2; vI2 = in_bounds(ins2, k, plen, tlen); }
        if (has_d2) { outD2[k] = del2; vD2 = in_bounds(del2, k, plen, tlen); }
      }
    }
    const int kw = kb + wid * 64;
    rM.add(__ballot(vM), kw);
    if constexpr (NC > 1) {
      rI1.add(__ballot(vI1), kw);
      rD1.add(__ballot(vD1), kw);
    }
    if constexpr (NC == 5) {
      rI2.add(__ballot(vI2), kw);
      rD2.add(__ballot(vD2), kw);
    }
  }
  // The furthest antidiagonal rides along with the range merge (one block
  // barrier for both): a wave maximum first, then the merge over waves.
  if (want_max) {
    for (int o = 32; o > 0; o >>= 1) best_ak = max(best_ak, __shfl_xor(best_ak, o, 64));
    best_ak = uniform(best_ak);
  }
  if (prune_try) wmin = red.min_all(wmin);
  if constexpr (NW == 1) {
    if (want_max) *max_ak = best_ak;
  }
  if constexpr (NW > 1) {
    int (*sh_red)[2 * NC + 2] = rr.buf[rr.par & 1];
    if (lane == 0) {
      sh_red[wid][2 * NC] = best_ak;
      sh_red[wid][0] = rM.first;  sh_red[wid][1] = rM.last;
      if constexpr (NC > 1) {
        sh_red[wid][2] = rI1.first; sh_red[wid][3] = rI1.last;
        sh_red[wid][4] = rD1.first; sh_red[wid][5] = rD1.last;
      }
      if constexpr (NC == 5) {
        sh_red[wid][6] = rI2.first; sh_red[wid][7] = rI2.last;
        sh_red[wid][8] = rD2.first; sh_red[wid][9] = rD2.last;
      }
    }
    __syncthreads();
    rM.init(); rI1.init(); rD1.init(); rI2.init(); rD2.init();
    int ak_all = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      ak_all = max(ak_all, sh_red[w][2 * NC]);
      rM.merge(sh_red[w][0], sh_red[w][1]);
      if constexpr (NC > 1) {
        rI1.merge(sh_red[w][2], sh_red[w][3]);
        rD1.merge(sh_red[w][4], sh_red[w][5]);
      }
      if constexpr (NC == 5) {
        rI2.merge(sh_red[w][6], sh_red[w][7]);
        rD2.merge(sh_red[w][8], sh_red[w][9]);
      }
    }
    if (want_max) *max_ak = ak_all;
    ++rr.par;
  }
  // (one barrier: the one before the merge already makes this score's ring
  // writes visible to the whole workgroup)
  if constexpr (NW == 1) __syncthreads();
  if constexpr (NC == 1) {
    if (prune_try && !rM.null()) {
      const int first0 = rM.first, last0 = rM.last;
      edit_prune(rM, ring, mo1, wmin, plen, tlen);
      // The reference extends only the kept range (the prune runs in the
      // compute step, before the extension), so the furthest antidiagonal
      // is taken over that range (wavefront_extend.c:173-211).
      if (want_max && (rM.first != first0 || rM.last != last0)) {
        int b = 0;
        for (int k = rM.first + tid; k <= rM.last; k += BLOCK) {
          const int m = outM[k];
          if (m != kOffsetNull) b = max(b, 2 * m - k);
        }
        *max_ak = red.max_all(b);
      }
    }
  }
  if (lane == cur) {
    sd.meta.tag = s;
    sd.meta.alo = lo;
    unsigned fl = 1u | (rM.null() ? 0x100u : 0u);
    sd.meta.lo[kM] = rM.tlo(lo);
    sd.meta.hi[kM] = rM.thi(lo);
    if constexpr (NC > 1) {
      if (has_i1) { fl |= 2u | (rI1.null() ? 0x200u : 0u); sd.meta.lo[kI1] = rI1.tlo(lo); sd.meta.hi[kI1] = rI1.thi(lo); }
      if (has_d1) { fl |= 4u | (rD1.null() ? 0x400u : 0u); sd.meta.lo[kD1] = rD1.tlo(lo); sd.meta.hi[kD1] = rD1.thi(lo); }
    }
    if constexpr (NC == 5) {
      if (has_i2) { fl |= 8u | (rI2.null() ? 0x800u : 0u); sd.meta.lo[kI2] = rI2.tlo(lo); sd.meta.hi[kI2] = rI2.thi(lo); }
      if (has_d2) { fl |= 16u | (rD2.null() ? 0x1000u : 0u); sd.meta.lo[kD2] = rD2.tlo(lo); sd.meta.hi[kD2] = rD2.thi(lo); }
    }
    if constexpr (CUT) {
#pragma unroll
      for (int c = 0; c < NC; ++c) { sd.meta.cut.lo[c] = sd.meta.lo[c]; sd.meta.cut.hi[c] = sd.meta.hi[c]; }
    }
    sd.meta.flags = fl;
  }
  const int kend = tlen - plen;
  const int ce = NC == 1 ? (int)kM : sd.ce;
  bool pce = true;
  WaveRange rce = rM;
  if (ce == kI1) { pce = has_i1; rce = rI1; }
  if (ce == kD1) { pce = has_d1; rce = rD1; }
  if constexpr (NC == 5) {
    if (ce == kI2) { pce = has_i2; rce = rI2; }
    if (ce == kD2) { pce = has_d2; rce = rD2; }
  }
  if (pce && !rce.null() && kend >= rce.first && kend <= rce.last) {
    if (uniform(ring[(cur * NC + ce) * wcap + (kend - lo)]) >= tlen) {
      sd.end_score = s;
      return 1;
    }
  }
  if (!rM.null() && side_cutoff<NC, BLOCK>(sd, heur, red, cur, lo, rM.first, rM.last, s, plen, tlen, wcap, pen.metric))
    return 2;
  return 0;
}

// First k0 of the overlap of w0 (side 0) and w1 (side 1, diagonals inverted)
// with h0 + h1 >= tlen; indel breakpoints also require the forward side's
// cell inside the matrix.  Sets the breakpoint (uniform) when found.
template <int NC, int BLOCK, bool CUT>
__device__ void scan_breakpoint(const Side<NC, CUT>& s0d, const Side<NC, CUT>& s1d, const WfRef<NC>& w0, const WfRef<NC>& w1,
                                bool forward, bool indel, int s0, int s1, int gap_open, int comp,
                                int plen, int tlen, Bp& bp, BlockRed<BLOCK>& red) {
  const int K = tlen - plen;
  const int lo_0 = w0.lo, hi_0 = w0.hi;
  const int lo_1 = K - w1.hi, hi_1 = K - w1.lo;
  if (hi_1 < lo_0 || hi_0 < lo_1) return;
  const int min_hi = min(hi_0, hi_1), max_lo = max(lo_0, lo_1);
  const int sc = indel ? s0 + s1 - gap_open : s0 + s1;
  if (sc >= bp.score) return;
  const int* r0 = s0d.ring + w0.base;
  const int* r1 = s1d.ring + w1.base;
  int first = INT_MAX;
  for (int kb = max_lo; kb <= min_hi; kb += BLOCK) {
    const int k0 = kb + (int)threadIdx.x;
    if (k0 <= min_hi && first == INT_MAX) {
      const int k1 = K - k0;
      const int h0 = r0[k0], h1 = r1[k1];
      bool ok = h0 + h1 >= tlen;
      if (ok && indel) {
        const int fk = forward ? k0 : k1, fh = forward ? h0 : h1;
        if (fh - fk > plen || fh > tlen) ok = false;
      }
      if (ok) first = k0;
    }
  }
  first = red.min_all(first);
  if (first == INT_MAX) return;
  const int k0 = first, k1 = K - k0;
  const int h0 = uniform(r0[k0]), h1 = uniform(r1[k1]);
  if (forward) {
    bp.score_forward = s0; bp.score_reverse = s1;
    bp.k_forward = k0; bp.k_reverse = k1;
    bp.offset_forward = h0; bp.offset_reverse = h1;
  } else {
    bp.score_forward = s1; bp.score_reverse = s0;
    bp.k_forward = k1; bp.k_reverse = k0;
    bp.offset_forward = h1; bp.offset_reverse = h0;
  }
  bp.score = sc;
  bp.component = comp;
}

// wavefront_bialign_overlap (wavefront_bialign.c:878-956)
template <int NC, int BLOCK, bool CUT>
__device__ void overlap(const Side<NC, CUT>& a0, const Side<NC, CUT>& a1, int s0, int s1, bool forward, int plen, int tlen,
                        const Penalties& pen, int wcap, Bp& bp, BlockRed<BLOCK>& red) {
  const WfRef<NC> m0 = wf_ref(a0, s0, kM, pen, wcap);
  if (!m0.present) return;
  WfRef<NC> d1w0{false, 1, 0, 0}, i1w0{false, 1, 0, 0};
  if constexpr (NC > 1) {
    d1w0 = wf_ref(a0, s0, kD1, pen, wcap);
    i1w0 = wf_ref(a0, s0, kI1, pen, wcap);
  }
  WfRef<NC> d2w0{false, 1, 0, 0}, i2w0{false, 1, 0, 0};
  if constexpr (NC == 5) {
    d2w0 = wf_ref(a0, s0, kD2, pen, wcap);
    i2w0 = wf_ref(a0, s0, kI2, pen, wcap);
  }
  for (int i = 0; i < pen.scope; ++i) {
    const int si = s1 - i;
    if (si < 0) break;
    if constexpr (NC == 5) {
      if (s0 + si - pen.o2 >= bp.score) continue;
      const WfRef<NC> d2w1 = wf_ref(a1, si, kD2, pen, wcap);
      if (d2w0.present && d2w1.present)
        scan_breakpoint<NC, BLOCK, CUT>(a0, a1, d2w0, d2w1, forward, true, s0, si, pen.o2, kD2, plen, tlen, bp, red);
      const WfRef<NC> i2w1 = wf_ref(a1, si, kI2, pen, wcap);
      if (i2w0.present && i2w1.present)
        scan_breakpoint<NC, BLOCK, CUT>(a0, a1, i2w0, i2w1, forward, true, s0, si, pen.o2, kI2, plen, tlen, bp, red);
    }
    if constexpr (NC > 1) {  // I1/D1 breakpoints: gap-affine metrics only
      if (s0 + si - pen.o1 >= bp.score) continue;
      const WfRef<NC> d1w1 = wf_ref(a1, si, kD1, pen, wcap);
      if (d1w0.present && d1w1.present)
        scan_breakpoint<NC, BLOCK, CUT>(a0, a1, d1w0, d1w1, forward, true, s0, si, pen.o1, kD1, plen, tlen, bp, red);
      const WfRef<NC> i1w1 = wf_ref(a1, si, kI1, pen, wcap);
      if (i1w0.present && i1w1.present)
        scan_breakpoint<NC, BLOCK, CUT>(a0, a1, i1w0, i1w1, forward, true, s0, si, pen.o1, kI1, plen, tlen, bp, red);
    }
    if (s0 + si >= bp.score) continue;
    const WfRef<NC> mw1 = wf_ref(a1, si, kM, pen, wcap);
    if (mw1.present) scan_breakpoint<NC, BLOCK, CUT>(a0, a1, m0, mw1, forward, false, s0, si, 0, kM, plen, tlen, bp, red);
  }
}

// own_ring: this workgroup writes the side's score-0 cell (a paired
// workgroup sets up its partner's side in registers only).
template <int NC, bool CUT>
__device__ void side_init(Side<NC, CUT>& sd, int cb, int ce, int wcap, int steps, bool own_ring = true) {
  const int lane = threadIdx.x & 63;
  sd.meta.tag = (lane == 0) ? 0 : -1;
  sd.meta.alo = 0;
  sd.meta.flags = (lane == 0) ? (1u << cb) : 0u;
#pragma unroll
  for (int c = 0; c < NC; ++c) { sd.meta.lo[c] = 0; sd.meta.hi[c] = 0; }
  if constexpr (CUT) {
#pragma unroll
    for (int c = 0; c < NC; ++c) { sd.meta.cut.lo[c] = 0; sd.meta.cut.hi[c] = 0; }
  }
  sd.steps_wait = steps;  // wavefront_heuristic_clear (wavefront_heuristic.c:114-121)
  sd.nulls = 0;
  sd.xs.clear();
  sd.dead = false;
  sd.cur = 0;
  sd.cb = cb;
  sd.ce = ce;
  sd.end_score = -1;
  if (own_ring && threadIdx.x == 0) sd.ring[(0 * NC + cb) * wcap] = 0;
}

// Extension of the score-0 wavefront (only M can extend).
template <int NC, int BLOCK, int BITS, bool CUT>
__device__ int side_extend0(Side<NC, CUT>& sd, int plen, int tlen, int wcap, const Penalties& pen,
                            const HeurArgs& heur, BlockRed<BLOCK>& red, int* max_ak) {
  int v0 = 0;
  *max_ak = 0;
  if (sd.cb != kM) return 0;  // M absent at score 0: nothing to extend, no termination test
  v0 = uniform(side_lcp<BITS>(sd, plen, tlen, 0, 0));
  __syncthreads();
  if (threadIdx.x == 0) sd.ring[(0 * NC + kM) * wcap] = v0;
  __syncthreads();
  *max_ak = 2 * v0;
  if (sd.ce == kM && tlen - plen == 0 && v0 >= tlen) {
    sd.end_score = 0;
    return 1;
  }
  // (the first z-drop test only records the best score: no drop at score 0)
  side_cutoff<NC, BLOCK>(sd, heur, red, 0, 0, 0, 0, 0, plen, tlen, wcap, pen.metric);
  return 0;
}

// Sequences and rings of sub-problem pr for both sides (ring: this
// workgroup's or team's region, forward then reverse); with a.lmax_words > 0
// the four sequences (forward and reversed pattern and text) are copied into
// LDS, so extension windows cost an LDS round trip.  The caller separates the
// previous sub-problem's use of the LDS with a barrier.
template <int NC, int BLOCK, int BITS, bool CUT>
__device__ __forceinline__ void setup_sides(const BiArgs& a, const BiProblem& pr, int* ring, int ring_side,
                                            Side<NC, CUT>& f, Side<NC, CUT>& r) {
  constexpr int SPW = 64 / BITS;
  const int pair = pr.pair;
  const int plen = pr.pe - pr.pb, tlen = pr.te - pr.tb;
  const int plen_all = a.plen_total[pair], tlen_all = a.tlen_total[pair];
  f.ring = ring;
  r.ring = ring + ring_side;
  const long long pp = a.fwd_ppos[pair] + pr.pb, tp = a.fwd_tpos[pair] + pr.tb;
  f.P = a.words + pp / SPW; f.pb = (int)(pp % SPW);
  f.T = a.words + tp / SPW; f.tb = (int)(tp % SPW);
  const long long rp = a.rev_ppos[pair] + (plen_all - pr.pe), rt = a.rev_tpos[pair] + (tlen_all - pr.te);
  r.P = a.rev_words + rp / SPW; r.pb = (int)(rp % SPW);
  r.T = a.rev_words + rt / SPW; r.tb = (int)(rt % SPW);
  f.lds_p = f.lds_t = r.lds_p = r.lds_t = -1;
  if (a.lmax_words > 0) {
    const int nwp = (f.pb + plen) / SPW + 2, nwt = (f.tb + tlen) / SPW + 2;
    const int nrp = (r.pb + plen) / SPW + 2, nrt = (r.tb + tlen) / SPW + 2;
    uint64_t* s0 = reinterpret_cast<uint64_t*>(bp_smem);
    uint64_t* s1 = s0 + a.lmax_words;
    uint64_t* s2 = s1 + a.lmax_words;
    uint64_t* s3 = s2 + a.lmax_words;
    for (int i = threadIdx.x; i < nwp; i += BLOCK) s0[i] = f.P[i];
    for (int i = threadIdx.x; i < nwt; i += BLOCK) s1[i] = f.T[i];
    for (int i = threadIdx.x; i < nrp; i += BLOCK) s2[i] = r.P[i];
    for (int i = threadIdx.x; i < nrt; i += BLOCK) s3[i] = r.T[i];
    f.lds_p = 0;
    f.lds_t = a.lmax_words;
    r.lds_p = 2 * a.lmax_words;
    r.lds_t = 3 * a.lmax_words;
  }
}

// Phase 2 of wavefront_bialign_find_breakpoint (wavefront_bialign.c:
// 1046-1080): overlap scans against the other side's last scores, one step
// of the side scanned against, alternating from where phase 1 stopped.
// Returns the search status (kBiOk when the bound stops it).
template <int NC, int BLOCK, int BITS, bool CUT>
__device__ int find_phase2(Side<NC, CUT>& f, Side<NC, CUT>& r, int& sf, int& sr, bool last_forward, int plen,
                           int tlen, const Penalties& pen, const BiArgs& a, int wcap, int s_cap, Bp& bp,
                           long long* dsteps, RedRows<NC, BLOCK / 64>& rr, BlockRed<BLOCK>& red) {
  const int scope = pen.scope;
  // wavefront_bialign_overlap_gopen_adjust (wavefront_bialign.c:960-973)
  const int gap_open = (NC == 5) ? max(pen.o1, pen.o2) : (NC == 3 ? pen.o1 : 0);
  int ak, rc;
  while (true) {
    if (last_forward) {
      const int min_sr = (sr > scope - 1) ? sr - (scope - 1) : 0;
      if (sf + min_sr - gap_open >= bp.score) break;
      overlap<NC, BLOCK, CUT>(f, r, sf, sr, true, plen, tlen, pen, wcap, bp, red);
      ++sr;
      rc = side_step<NC, BLOCK, BITS>(r, sr, plen, tlen, pen, a.heur, wcap, false, &ak, dsteps, rr, red);
      if (rc < 0) return kDevRingOverflow;
      if (rc == 2) return kDevUnreachable;
      if (rc > 0) return kBiEndReached;
    }
    const int min_sf = (sf > scope - 1) ? sf - (scope - 1) : 0;
    if (min_sf + sr - gap_open >= bp.score) break;
    overlap<NC, BLOCK, CUT>(r, f, sr, sf, false, plen, tlen, pen, wcap, bp, red);
    ++sf;
    rc = side_step<NC, BLOCK, BITS>(f, sf, plen, tlen, pen, a.heur, wcap, false, &ak, dsteps, rr, red);
    if (rc < 0) return kDevRingOverflow;
    if (rc == 2) return kDevUnreachable;
    if (rc > 0) return kBiEndReached;
    if (sr + sf >= a.max_steps) return kDevMaxSteps;
    // both sides emptied and no breakpoint yet: nothing changes until the step limit
    if (f.dead && r.dead && bp.score == INT_MAX) return kDevMaxSteps;
    if (sf > s_cap) return kDevBadTrace;
    last_forward = true;
  }
  return kBiOk;
}

template <int NC, bool CUT>
__device__ __forceinline__ void write_outcome(BiOutcome* out, int status, const Bp& bp, int fwd_end, int rev_end,
                                              long long dsteps) {
  BiOutcome o;
  o.status = status;
  o.score = bp.score;
  o.score_forward = bp.score_forward;
  o.score_reverse = bp.score_reverse;
  o.k_forward = bp.k_forward;
  o.k_reverse = bp.k_reverse;
  o.offset_forward = bp.offset_forward;
  o.offset_reverse = bp.offset_reverse;
  o.component = bp.component;
  o.fwd_end_score = fwd_end;
  o.rev_end_score = rev_end;
  o.pad = 0;
  o.dsteps = dsteps;
  *out = o;
}

__device__ __forceinline__ void bp_clear(Bp& bp) {
  bp.score = INT_MAX;
  bp.score_forward = bp.score_reverse = bp.k_forward = bp.k_reverse = 0;
  bp.offset_forward = bp.offset_reverse = 0;
  bp.component = kM;
}

template <int NC, int BLOCK, int BITS, bool CUT>
__global__ void __launch_bounds__(BLOCK) wfa_breakpoint_kernel(BiArgs a) {
  __shared__ int sh_red[2][BLOCK / 64][2 * NC + 2];
  RedRows<NC, BLOCK / 64> rr{sh_red, 0};
  __shared__ int sh_min[BLOCK / 64 > 0 ? BLOCK / 64 : 1];
  __shared__ int sh_q;
  BlockRed<BLOCK> red{sh_min};
  const Penalties pen = a.pen;
  const int wcap = a.wcap;
  long long wg_dsteps = 0;
  for (;;) {
    if (threadIdx.x == 0) sh_q = (int)atomicAdd(a.queue, 1u);
    __syncthreads();
    const int q = uniform(sh_q);
    __syncthreads();
    if (q >= a.n) break;
    const BiProblem pr = a.probs[q];
    const int plen = pr.pe - pr.pb, tlen = pr.te - pr.tb;
    Side<NC, CUT> f, r;
    __syncthreads();  // the previous sub-problem is done with the rings and the LDS sequences
    setup_sides<NC, BLOCK, BITS>(a, pr, a.ring + (long long)blockIdx.x * a.ring_stride,
                                 pen.nslots * NC * wcap, f, r);
    side_init(f, pr.comp_begin, pr.comp_end, wcap, a.heur.steps);
    side_init(r, pr.comp_end, pr.comp_begin, wcap, a.heur.steps);
    __syncthreads();
    Bp bp;
    bp_clear(bp);
    int status = kBiOk;
    long long dsteps = 0;
    const int s_cap = (pen.x + pen.o1 + pen.e1 + (NC == 5 ? pen.o2 + pen.e2 : 0) + 2) * (plen + tlen + 2);
    int sf = 0, sr = 0, fmax = 0, rmax = 0, ak = 0;
    const int max_ad = plen + tlen - 1;
    bool last_forward = false;
    int rc;
    // phase 1 (wavefront_bialign.c:1003-1044)
    if (side_extend0<NC, BLOCK, BITS>(f, plen, tlen, wcap, pen, a.heur, red, &fmax)) { status = kBiEndReached; goto done; }
    if (side_extend0<NC, BLOCK, BITS>(r, plen, tlen, wcap, pen, a.heur, red, &rmax)) { status = kBiEndReached; goto done; }
    while (true) {
      if (fmax + rmax >= max_ad) break;
      ++sf;
      rc = side_step<NC, BLOCK, BITS>(f, sf, plen, tlen, pen, a.heur, wcap, true, &ak, &dsteps, rr, red);
      if (rc < 0) { status = kDevRingOverflow; goto done; }
      if (rc == 2) { status = kDevUnreachable; goto done; }
      if (fmax < ak) fmax = ak;
      last_forward = true;
      if (rc > 0) { status = kBiEndReached; goto done; }
      if (fmax + rmax >= max_ad) break;
      ++sr;
      rc = side_step<NC, BLOCK, BITS>(r, sr, plen, tlen, pen, a.heur, wcap, true, &ak, &dsteps, rr, red);
      if (rc < 0) { status = kDevRingOverflow; goto done; }
      if (rc == 2) { status = kDevUnreachable; goto done; }
      if (rmax < ak) rmax = ak;
      last_forward = false;
      if (rc > 0) { status = kBiEndReached; goto done; }
      if (sr + sf >= a.max_steps) { status = kDevMaxSteps; goto done; }
      // both sides emptied: nothing changes until the step limit
      if (f.dead && r.dead) { status = kDevMaxSteps; goto done; }
      if (sf > s_cap) { status = kDevBadTrace; goto done; }
    }
    status = find_phase2<NC, BLOCK, BITS>(f, r, sf, sr, last_forward, plen, tlen, pen, a, wcap, s_cap, bp, &dsteps,
                                          rr, red);
  done:
    if (threadIdx.x == 0) write_outcome<NC, CUT>(a.out + q, status, bp, f.end_score, r.end_score, dsteps);
    wg_dsteps += dsteps;
    __syncthreads();
  }
  if (threadIdx.x == 0 && wg_dsteps) atomicAdd(a.dsteps_total, (unsigned long long)wg_dsteps);
}

// ---------------------------------------------------------------------------
// Paired workgroups: the forward and the reverse side of one large
// sub-problem on two workgroups (a team: blocks 2t and 2t + 1, one
// sub-problem per team), for launches with fewer sub-problems than half the
// workgroups the CUs hold (the top levels of small batches, a serving lane's
// long requests).  Phase 1 advances both sides at once, one score each per
// round: round i computes forward score i and, speculatively, reverse score
// i, then the two exchange their furthest antidiagonals and step statuses
// and both apply the reference's order to the pair (forward step, its end
// test and the meeting test, then the reverse step and its tests).  When the
// meeting test stops phase 1 after the forward step, the reverse workgroup
// takes its speculative step back (its ring has one slot more than the score
// scope needs, so the step overwrote nothing phase 2 reads).  The reverse
// side's state then goes to the forward workgroup, which runs phase 2 alone.
//
// Exchanges use agent-scope atomics on a per-team mailbox (one 64-bit word
// per side and round parity: round number, step status, antidiagonal); the
// hand-over of the reverse side's rings and slot metadata is ordered by a
// device-scope release fence before its mailbox word and an acquire fence
// after it.  Every wait is bounded by BiArgs::pair_wait_ticks of the 100 MHz
// device clock (4 s unless the host asks for less); a team whose partner
// does not answer in time reports kBiPairAbort and the host searches that
// sub-problem again on one workgroup.

// Handover record of the reverse side: per lane of wave 0, its slot's
// metadata, then the side's scalars.
template <int NC>
struct PairHandover {
  static constexpr int kLaneInts = 3 + 2 * NC;
  static constexpr int kInts = 64 * kLaneInts + 8;
};

// Publishes `payload` for exchange `seq` and waits for the partner's word of
// the same exchange (parity slots: a partner can be at most one exchange
// ahead).  Returns false on timeout.  Every thread of the workgroup calls it.
__device__ __forceinline__ bool pair_exchange(unsigned long long* box, int role, unsigned seq, unsigned payload,
                                              unsigned* got, unsigned long long* sh, unsigned long long wait_ticks) {
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long* mine = box + 2 * role + (seq & 1u);
    unsigned long long* other = box + 2 * (role ^ 1) + (seq & 1u);
    __hip_atomic_store(mine, ((unsigned long long)seq << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
    unsigned long long w = 0;
    bool ok = false;
    for (;;) {
      w = __hip_atomic_load(other, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((unsigned)(w >> 32) == seq) { ok = true; break; }
      if (__builtin_amdgcn_s_memrealtime() - t0 > wait_ticks) break;
      __builtin_amdgcn_s_sleep(2);
    }
    *sh = ok ? (w & 0xFFFFFFFFull) : ~0ull;
  }
  __syncthreads();
  const unsigned long long w = *sh;
  __syncthreads();  // *sh is reused by the next exchange
  if (w == ~0ull) return false;
  *got = (unsigned)w;
  return true;
}

// payload: step status + 1 in the top 4 bits, antidiagonal below (< 2^28)
__device__ __forceinline__ unsigned pair_payload(int rc, int ak) {
  return ((unsigned)(rc + 1) << 28) | ((unsigned)ak & 0x0FFFFFFFu);
}
__device__ __forceinline__ int payload_rc(unsigned w) { return (int)(w >> 28) - 1; }
__device__ __forceinline__ int payload_ak(unsigned w) { return (int)(w & 0x0FFFFFFFu); }

template <int NC, int BLOCK, int BITS>
__global__ void __launch_bounds__(BLOCK) wfa_breakpoint_pair_kernel(BiArgs a) {
  constexpr bool CUT = false;
  __shared__ int sh_red[2][BLOCK / 64][2 * NC + 2];
  RedRows<NC, BLOCK / 64> rr{sh_red, 0};
  __shared__ int sh_min[BLOCK / 64 > 0 ? BLOCK / 64 : 1];
  __shared__ unsigned long long sh_word;
  BlockRed<BLOCK> red{sh_min};
  const Penalties pen = a.pen;  // nslots: one more than the score scope needs
  const int wcap = a.wcap;
  const int team = blockIdx.x >> 1, role = blockIdx.x & 1;  // role 0: forward side, 1: reverse side
  const int q = team;
  if (q >= a.n) return;
  unsigned long long* box = a.mail + 4 * team;
  int* hand = a.handover + (long long)team * PairHandover<NC>::kInts;
  unsigned seq = 0;
  unsigned got = 0;
  const BiProblem pr = a.probs[q];
  const int plen = pr.pe - pr.pb, tlen = pr.te - pr.tb;
  Side<NC, CUT> f, r;
  setup_sides<NC, BLOCK, BITS>(a, pr, a.ring + (long long)team * a.ring_stride, pen.nslots * NC * wcap, f, r);
  side_init(f, pr.comp_begin, pr.comp_end, wcap, a.heur.steps, role == 0);
  side_init(r, pr.comp_end, pr.comp_begin, wcap, a.heur.steps, role == 1);
  Side<NC, CUT>& me = role ? r : f;
  __syncthreads();
  Bp bp;
  bp_clear(bp);
  int status = kBiOk;
  long long dsteps = 0;
  const int s_cap = (pen.x + pen.o1 + pen.e1 + (NC == 5 ? pen.o2 + pen.e2 : 0) + 2) * (plen + tlen + 2);
  int sf = 0, sr = 0, fmax = 0, rmax = 0, ak = 0;
  int fwd_end = -1, rev_end = -1;
  const int max_ad = plen + tlen - 1;
  bool last_forward = false;
  int rc_me, rcf, rcr, akf, akr;
  // score 0 of both sides, then the reference's order: forward end test, reverse end test
  rc_me = side_extend0<NC, BLOCK, BITS>(me, plen, tlen, wcap, pen, a.heur, red, &ak);
  if (!pair_exchange(box, role, ++seq, pair_payload(rc_me, ak), &got, &sh_word, a.pair_wait_ticks)) { status = kBiPairAbort; goto done; }
  rcf = role ? payload_rc(got) : rc_me;  akf = role ? payload_ak(got) : ak;
  rcr = role ? rc_me : payload_rc(got);  akr = role ? ak : payload_ak(got);
  fmax = akf;
  if (rcf) { status = kBiEndReached; fwd_end = 0; goto done; }
  rmax = akr;
  if (rcr) { status = kBiEndReached; rev_end = 0; goto done; }
  // phase 1 (wavefront_bialign.c:1003-1044), one round per score of both sides
  while (true) {
    if (fmax + rmax >= max_ad) break;
    ++sf;
    ++sr;
    // the reverse side's state before its (speculative) step
    const int cur0 = r.cur, end0 = r.end_score;
    const long long ds0 = dsteps;
    rc_me = side_step<NC, BLOCK, BITS>(me, role ? sr : sf, plen, tlen, pen, a.heur, wcap, true, &ak, &dsteps, rr,
                                       red);
    if (!pair_exchange(box, role, ++seq, pair_payload(rc_me, ak), &got, &sh_word, a.pair_wait_ticks)) { status = kBiPairAbort; goto done; }
    rcf = role ? payload_rc(got) : rc_me;  akf = role ? payload_ak(got) : ak;
    rcr = role ? rc_me : payload_rc(got);  akr = role ? ak : payload_ak(got);
    if (rcf < 0) { status = kDevRingOverflow; goto done; }
    if (rcf == 2) { status = kDevUnreachable; goto done; }
    if (fmax < akf) fmax = akf;
    last_forward = true;
    if (rcf > 0) { status = kBiEndReached; fwd_end = sf; goto done; }
    if (fmax + rmax >= max_ad) {
      // the reverse step of this round never happened: take it back
      --sr;
      if (role) {
        r.cur = cur0;
        r.end_score = end0;
        dsteps = ds0;
      }
      break;
    }
    if (rcr < 0) { status = kDevRingOverflow; goto done; }
    if (rcr == 2) { status = kDevUnreachable; goto done; }
    if (rmax < akr) rmax = akr;
    last_forward = false;
    if (rcr > 0) { status = kBiEndReached; rev_end = sr; goto done; }
    if (sr + sf >= a.max_steps) { status = kDevMaxSteps; goto done; }
    if (sf > s_cap) { status = kDevBadTrace; goto done; }
  }
  // hand the reverse side to the forward workgroup, which runs phase 2
  if (role) {
    if (threadIdx.x < 64) {
      int* h = hand + threadIdx.x * PairHandover<NC>::kLaneInts;
      h[0] = r.meta.tag;
      h[1] = r.meta.alo;
      h[2] = (int)r.meta.flags;
#pragma unroll
      for (int c = 0; c < NC; ++c) { h[3 + c] = r.meta.lo[c]; h[3 + NC + c] = r.meta.hi[c]; }
    }
    if (threadIdx.x == 0) {
      int* hs = hand + 64 * PairHandover<NC>::kLaneInts;
      hs[0] = r.cur;
      hs[1] = r.end_score;
    }
    __syncthreads();
    if (threadIdx.x == 0) __threadfence();  // the rings and the record above, device-wide, before the word
    (void)pair_exchange(box, role, ++seq, 0u, &got, &sh_word, a.pair_wait_ticks);
    goto done;
  }
  if (!pair_exchange(box, role, ++seq, 0u, &got, &sh_word, a.pair_wait_ticks)) { status = kBiPairAbort; goto done; }
  if (threadIdx.x == 0) __threadfence();  // ... and after it, before reading them
  __syncthreads();
  {
    const int lane = threadIdx.x & 63;
    const int* h = hand + lane * PairHandover<NC>::kLaneInts;
    r.meta.tag = h[0];
    r.meta.alo = h[1];
    r.meta.flags = (unsigned)h[2];
#pragma unroll
    for (int c = 0; c < NC; ++c) { r.meta.lo[c] = h[3 + c]; r.meta.hi[c] = h[3 + NC + c]; }
    const int* hs = hand + 64 * PairHandover<NC>::kLaneInts;
    r.cur = uniform(hs[0]);
    r.end_score = uniform(hs[1]);
  }
  status = find_phase2<NC, BLOCK, BITS>(f, r, sf, sr, last_forward, plen, tlen, pen, a, wcap, s_cap, bp, &dsteps,
                                        rr, red);
  fwd_end = f.end_score;
  rev_end = r.end_score;
done:
  if (role == 0 && threadIdx.x == 0) write_outcome<NC, CUT>(a.out + q, status, bp, fwd_end, rev_end, dsteps);
  if (threadIdx.x == 0 && dsteps) atomicAdd(a.dsteps_total, (unsigned long long)dsteps);
}

template <int NC, int BLOCK, int BITS, bool CUT>
const void* bp_fn() {
  return reinterpret_cast<const void*>(&wfa_breakpoint_kernel<NC, BLOCK, BITS, CUT>);
}

const void* select_breakpoint_pair(int ncomp, int bits, int block) {
#define WFA_PSEL(NC, BL, BI) \
  if (ncomp == NC && block == BL && bits == BI) return reinterpret_cast<const void*>(&wfa_breakpoint_pair_kernel<NC, BL, BI>);
  WFA_PSEL(1, 1024, 2) WFA_PSEL(1, 1024, 8) WFA_PSEL(3, 1024, 2) WFA_PSEL(3, 1024, 8) WFA_PSEL(5, 1024, 2)
  WFA_PSEL(5, 1024, 8)
#undef WFA_PSEL
  return nullptr;
}

const void* select_breakpoint(int ncomp, int bits, int block, bool cut) {
#define WFA_BSEL(NC, BL, BI, CU) \
  if (ncomp == NC && block == BL && bits == BI && cut == CU) return bp_fn<NC, BL, BI, CU>();
#define WFA_BSEL_NC(BL, CU) \
  WFA_BSEL(1, BL, 2, CU) WFA_BSEL(1, BL, 8, CU) WFA_BSEL(3, BL, 2, CU) WFA_BSEL(3, BL, 8, CU) \
  WFA_BSEL(5, BL, 2, CU) WFA_BSEL(5, BL, 8, CU)
  WFA_BSEL_NC(64, false)
  WFA_BSEL_NC(256, false)
  WFA_BSEL_NC(1024, false)
  WFA_BSEL_NC(64, true)
  WFA_BSEL_NC(256, true)
  WFA_BSEL_NC(1024, true)
#undef WFA_BSEL_NC
#undef WFA_BSEL
  return nullptr;
}

// Reversed copies of packed sequences: sequence j (patterns, then texts) of
// length len[j] at symbol src_pos[j] of `words` is written reversed at symbol
// dst_pos[j] (word aligned) of `out`, zero padded like the forward packing.
template <int BITS>
__global__ void __launch_bounds__(256) reverse_pack_kernel(const uint64_t* words, const int64_t* src_pos,
                                                            const int64_t* dst_pos, const int* len, int n,
                                                            uint64_t* out) {
  constexpr int SPW = 64 / BITS;
  constexpr uint64_t MASK = (BITS == 64) ? ~0ull : ((1ull << BITS) - 1);
  const int j = blockIdx.x;
  if (j >= n) return;
  const int L = len[j];
  const long long sp = src_pos[j];
  uint64_t* dst = out + dst_pos[j] / SPW;
  const int nw = L / SPW + 2;
  for (int w = threadIdx.x; w < nw; w += blockDim.x) {
    uint64_t v = 0;
    for (int b = 0; b < SPW; ++b) {
      const int i = w * SPW + b;
      if (i >= L) break;
      const long long src = sp + (L - 1 - i);
      const uint64_t sym = (words[src / SPW] >> ((src % SPW) * BITS)) & MASK;
      v |= sym << (b * BITS);
    }
    dst[w] = v;
  }
}

// Copies CIGAR pieces (or writes runs of one operation) to their final place.
__global__ void __launch_bounds__(256) gather_ops_kernel(const OpsPiece* pieces, int n, const char* src, char* dst) {
  const int j = blockIdx.x;
  if (j >= n) return;
  const OpsPiece pc = pieces[j];
  for (int i = threadIdx.x; i < pc.len; i += blockDim.x)
    dst[pc.dst + i] = pc.src >= 0 ? src[pc.src + i] : pc.op;
}

}  // namespace

hipError_t launch_reverse_pack(const uint64_t* words, const int64_t* src_pos, const int64_t* dst_pos, const int* len,
                               int n, int bits, uint64_t* out, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  if (bits == 2)
    hipLaunchKernelGGL(reverse_pack_kernel<2>, dim3(n), dim3(256), 0, stream, words, src_pos, dst_pos, len, n, out);
  else
    hipLaunchKernelGGL(reverse_pack_kernel<8>, dim3(n), dim3(256), 0, stream, words, src_pos, dst_pos, len, n, out);
  return hipGetLastError();
}

hipError_t launch_gather_ops(const OpsPiece* pieces, int n, const char* src, char* dst, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_ops_kernel, dim3(n), dim3(256), 0, stream, pieces, n, src, dst);
  return hipGetLastError();
}

size_t breakpoint_lds_bytes(int lmax_words) { return (size_t)4 * lmax_words * sizeof(uint64_t); }

int breakpoint_max_blocks_per_cu(int ncomp, int bits, int block, bool cut, size_t lds_bytes) {
  const void* f = select_breakpoint(ncomp, bits, block, cut);
  if (!f) return 0;
  if (lds_bytes > 48 * 1024 &&
      hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess)
    return 0;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, block, lds_bytes) != hipSuccess) return 0;
  return n;
}

int breakpoint_pair_max_blocks_per_cu(int ncomp, int bits, int block, size_t lds_bytes) {
  const void* f = select_breakpoint_pair(ncomp, bits, block);
  if (!f) return 0;
  if (lds_bytes > 48 * 1024 &&
      hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess)
    return 0;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, block, lds_bytes) != hipSuccess) return 0;
  return n;
}

size_t breakpoint_pair_handover_ints(int ncomp) {
  return ncomp == 1 ? PairHandover<1>::kInts : (ncomp == 3 ? PairHandover<3>::kInts : PairHandover<5>::kInts);
}

hipError_t launch_breakpoint_pair(const BiArgs& a, int ncomp, int bits, int block, hipStream_t stream) {
  const void* f = select_breakpoint_pair(ncomp, bits, block);
  if (!f || a.heur.cut() != 0 || !a.mail || !a.handover) return hipErrorInvalidValue;
  const size_t lds = breakpoint_lds_bytes(a.lmax_words);
  if (lds > 48 * 1024) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  void* args[] = {const_cast<BiArgs*>(&a)};
  return hipLaunchKernel(f, dim3(2 * a.n), dim3(block), args, lds, stream);
}

hipError_t launch_breakpoint(const BiArgs& a, int ncomp, int bits, int block, int grid, hipStream_t stream) {
  const void* f = select_breakpoint(ncomp, bits, block, a.heur.cut() != 0);
  if (!f) return hipErrorInvalidValue;
  const size_t lds = breakpoint_lds_bytes(a.lmax_words);
  if (lds > 48 * 1024) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  void* args[] = {const_cast<BiArgs*>(&a)};
  return hipLaunchKernel(f, dim3(grid), dim3(block), args, lds, stream);
}

}  // namespace wfa
