// Device helpers shared by the unidirectional (wfa_kernels.hip) and the
// bidirectional (wfa_biwfa.hip) kernels.  Included by .hip files only.
#pragma once
#include <climits>
#include <type_traits>
#include "wfa_device.h"

namespace wfa {

// Ring metadata lives in registers: lane r of every wave holds the state of
// ring slot r (score tag, computed-range origin, trimmed [lo,hi] of each
// component, present/null flags).  A source is read with v_readlane at a
// uniform slot index, so no memory round trip sits on the per-step critical
// path; every wave of a workgroup keeps an identical copy.
// With a heuristic cut-off (CUT) a slot also keeps, per component, the range
// the cut-off left: it sets the limits of the wavefronts computed from it,
// while reads still see the whole trimmed range (the reference narrows
// lo/hi but not the initialised range that wavefront_compute_init_ends,
// wavefront_compute.c:498-575, fills with NULLs).
template <int NC, bool CUT>
struct CutRange {
  int lo[NC];
  int hi[NC];
};
template <int NC>
struct CutRange<NC, false> {};

template <int NC, bool CUT = false>
struct MetaRegs {
  int tag;             // score held by the slot, -1 if none
  int alo;             // lo of the computed range: array origin of all components
  unsigned flags;      // bit c: component present; bit 8+c: component null (empty after trim / cut-off)
  int lo[NC];          // trimmed range (readable)
  int hi[NC];
  CutRange<NC, CUT> cut;
};

constexpr int kFarDiagonal = 0x40000000;

// A compute input.  lim_lo/lim_hi are the limits it contributes to the output
// range (the shared null wavefront of the reference has lo = 1, hi = -1).
// Reads at k in [lo, lo+span] return ring[base + k - lo], else NULL; an
// invalid source gets an empty read range far from any diagonal.
struct Src {
  int lim_lo, lim_hi;
  int lo, span, base;
  __device__ __forceinline__ bool valid() const { return lim_lo <= lim_hi; }
  __device__ __forceinline__ bool readable() const { return lo != kFarDiagonal && span >= 0; }
};

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int lane_of(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// The input wavefront of component c at score s_src, held in ring slot
// `slot`.  A null wavefront is no input (invalid); with keep_null, one that
// the cut-off emptied still gives its readable cells (the reference's
// backtrace probes them, wavefront_backtrace.c:64-219).
template <int NC, bool CUT>
__device__ __forceinline__ Src get_src(const MetaRegs<NC, CUT>& m, int s_src, int slot, int R, int c, int wcap,
                                       bool keep_null = false) {
  Src r{1, -1, kFarDiagonal, 0, 0};
  if (s_src < 0) return r;
  if (slot < 0) slot += R;
  slot = uniform(slot);
  const int tag = lane_of(m.tag, slot);
  const unsigned fl = (unsigned)lane_of((int)m.flags, slot);
  if (tag != s_src) return r;
  if (!(fl & (1u << c)) || (!keep_null && (fl & (0x100u << c)))) return r;
  const int lo = lane_of(m.lo[c], slot), span = lane_of(m.hi[c], slot) - lo;
  // a trimmed-empty wavefront has nothing to read (rd needs span >= 0)
  if (span < 0) return r;
  r.lo = lo;
  r.span = span;
  if constexpr (CUT) {
    r.lim_lo = lane_of(m.cut.lo[c], slot);
    r.lim_hi = lane_of(m.cut.hi[c], slot);
  } else {
    r.lim_lo = r.lo;
    r.lim_hi = r.lo + r.span;
  }
  r.base = (slot * NC + c) * wcap + (r.lo - lane_of(m.alo, slot));
  return r;
}

// Ring layout of the generic forward kernel: M in nslots slots, then I1 and
// D1 in r1 slots each, then I2 and D2 in r2 slots each (a gap wavefront is
// only read e/g steps later).  RingPos: the current step's slot in each.
struct RingPos {
  int cur, cur1, cur2;
  __device__ __forceinline__ void advance(const Penalties& pen) {
    cur = (cur + 1 == pen.nslots) ? 0 : cur + 1;
    cur1 = (cur1 + 1 >= pen.r1) ? 0 : cur1 + 1;
    cur2 = (cur2 + 1 >= pen.r2) ? 0 : cur2 + 1;
  }
};

// Element offset of component c's ring slot `back` steps before the current
// one (back < the component's ring depth).
__device__ __forceinline__ int ring_base(const Penalties& pen, int wcap, int c, const RingPos& p, int back) {
  if (c == kM) {
    int x = p.cur - back;
    if (x < 0) x += pen.nslots;
    return x * wcap;
  }
  const bool one = c == kI1 || c == kD1;
  const int rc = one ? pen.r1 : pen.r2;
  int x = (one ? p.cur1 : p.cur2) - back;
  if (x < 0) x += rc;
  const int first = pen.nslots + (c == kD1 ? pen.r1 : 0) + (c == kI2 ? 2 * pen.r1 : 0) +
                    (c == kD2 ? 2 * pen.r1 + pen.r2 : 0);
  return (first + x) * wcap;
}

// get_src over that layout: rbase is ring_base of the source.
template <int NC, bool CUT>
__device__ __forceinline__ Src get_src_at(const MetaRegs<NC, CUT>& m, int s_src, int slot, int R, int c, int rbase,
                                          bool keep_null = false) {
  Src r{1, -1, kFarDiagonal, 0, 0};
  if (s_src < 0) return r;
  if (slot < 0) slot += R;
  slot = uniform(slot);
  const int tag = lane_of(m.tag, slot);
  const unsigned fl = (unsigned)lane_of((int)m.flags, slot);
  if (tag != s_src) return r;
  if (!(fl & (1u << c)) || (!keep_null && (fl & (0x100u << c)))) return r;
  const int lo = lane_of(m.lo[c], slot), span = lane_of(m.hi[c], slot) - lo;
  if (span < 0) return r;
  r.lo = lo;
  r.span = span;
  if constexpr (CUT) {
    r.lim_lo = lane_of(m.cut.lo[c], slot);
    r.lim_hi = lane_of(m.cut.hi[c], slot);
  } else {
    r.lim_lo = r.lo;
    r.lim_hi = r.lo + r.span;
  }
  r.base = rbase + (r.lo - lane_of(m.alo, slot));
  return r;
}

// Branch-free predicated read: the load address is always inside the ring.
template <typename RingPtr>
__device__ __forceinline__ int rd(RingPtr ring, const Src& s, int k) {
  const int i = k - s.lo;
  const bool in = (unsigned)i <= (unsigned)s.span;
  // base + i is a non-negative element of the ring: an unsigned index lets
  // a global ring be read with a 32-bit offset from its uniform base
  const int v = (int)ring[(unsigned)(s.base + (in ? i : 0))];
  return in ? v : kOffsetNull;
}

// The choices of cell k over its sources, as one key: the winning M offset,
// the winning M source and the gap-extension flags (the record byte's bits
// 0..6), ties as the compute step breaks them.  Reads below 0 count as the
// one NULL offset, so that two ways of holding NULL compare equal.
template <int NC, typename RingPtr>
__device__ __forceinline__ unsigned long long choice_key(RingPtr ring, const Src& mx, const Src& mo1, const Src& i1,
                                                         const Src& d1, const Src& mo2, const Src& i2, const Src& d2,
                                                         int k) {
  auto rdn = [&](const Src& s, int kk) {
    const int v = rd(ring, s, kk);
    return v < 0 ? kOffsetNull : v;
  };
  int best = rdn(mx, k) + 1;
  uint32_t src = kSrcX, bits = 0;
  if constexpr (NC == 1) {
    const int del = rdn(mo1, k + 1), ins = rdn(mo1, k - 1) + 1;
    if (del > best) { best = del; src = kSrcD1; }
    if (ins > best) { best = ins; src = kSrcI1; }
  } else {
    const int oml = rdn(mo1, k - 1), omr = rdn(mo1, k + 1), eil = rdn(i1, k - 1), edr = rdn(d1, k + 1);
    bits = (eil >= oml ? kRecI1Ext : 0u) | (edr >= omr ? kRecD1Ext : 0u);
    int ins2 = kOffsetNull;
    if constexpr (NC == 5) {
      const int o2l = rdn(mo2, k - 1), o2r = rdn(mo2, k + 1), e2l = rdn(i2, k - 1), e2r = rdn(d2, k + 1);
      bits |= (e2l >= o2l ? kRecI2Ext : 0u) | (e2r >= o2r ? kRecD2Ext : 0u);
      ins2 = max(o2l, e2l) + 1;
      const int del2 = max(o2r, e2r);
      if (del2 > best) { best = del2; src = kSrcD2; }
    }
    const int del1 = max(omr, edr), ins1 = max(oml, eil) + 1;
    if (del1 > best) { best = del1; src = kSrcD1; }
    if constexpr (NC == 5) { if (ins2 > best) { best = ins2; src = kSrcI2; } }
    if (ins1 > best) { best = ins1; src = kSrcI1; }
  }
  return ((unsigned long long)(unsigned)best << 32) | (src | bits);
}

// Record byte of cell k for a step with an input the cut-off emptied: the
// reference's backtrace probes that input's readable cells too
// (wavefront_backtrace.c:64-219), so the byte holds the choices over the
// probed sources (b*).  The match run it then writes before the cell is the
// extended offset m minus the winning probed offset; a winning offset past m
// (possible only through such a cell) takes operations back instead, which
// the replay does not do: kRecStale marks the cell.  A NULL cell (m < 0) is
// compared by its computed best (over the compute sources c*) instead.
template <int NC, typename RingPtr>
__device__ __forceinline__ uint32_t probed_record(RingPtr ring, const Src& bmx, const Src& bmo1, const Src& bi1,
                                                  const Src& bd1, const Src& bmo2, const Src& bi2, const Src& bd2,
                                                  const Src& cmx, const Src& cmo1, const Src& ci1, const Src& cd1,
                                                  const Src& cmo2, const Src& ci2, const Src& cd2, int k, int m) {
  const unsigned long long kb = choice_key<NC>(ring, bmx, bmo1, bi1, bd1, bmo2, bi2, bd2, k);
  const unsigned long long kc = choice_key<NC>(ring, cmx, cmo1, ci1, cd1, cmo2, ci2, cd2, k);
  const int q = (int)(unsigned)(kb >> 32), best = (int)(unsigned)(kc >> 32);
  uint32_t r = (uint32_t)kb & 0x7Fu;
  if (q >= 0 && q > max(m, best)) r |= kRecStale;
  return r;
}

// The sources of score s (ring slot cur) as the compute step reads them (c:
// mx, mo1, i1, d1, mo2, i2, d2) and as the reference's backtrace probes them
// (b: also the readable cells of a wavefront the cut-off emptied,
// wavefront_backtrace.c:64-219).  Returns whether the two differ.
template <int NC, bool CUT>
__device__ __forceinline__ bool probe_sources(const MetaRegs<NC, CUT>& meta, int s, const RingPos& rp,
                                              const Penalties& pen, int wcap, Src* c, Src* b) {
  const int R = pen.nslots, g = pen.g, cur = rp.cur;
  bool stale = false;
  auto both = [&](int i, int s_src, int back, int comp) {
    const int rb = ring_base(pen, wcap, comp, rp, back);
    c[i] = get_src_at(meta, s_src, cur - back, R, comp, rb);
    b[i] = c[i];
    if (!c[i].valid()) {
      b[i] = get_src_at(meta, s_src, cur - back, R, comp, rb, true);
      stale = stale || b[i].readable();
    }
  };
  const Src none{1, -1, kFarDiagonal, 0, 0};
  for (int i = 0; i < 7; ++i) c[i] = b[i] = none;
  if (NC > 1 || pen.metric != kIndel) both(0, s - pen.x, pen.x / g, kM);
  both(1, s - pen.o1 - pen.e1, (pen.o1 + pen.e1) / g, kM);
  if constexpr (NC > 1) {
    both(2, s - pen.e1, pen.e1 / g, kI1);
    both(3, s - pen.e1, pen.e1 / g, kD1);
  }
  if constexpr (NC == 5) {
    both(4, s - pen.o2 - pen.e2, (pen.o2 + pen.e2) / g, kM);
    both(5, s - pen.e2, pen.e2 / g, kI2);
    both(6, s - pen.e2, pen.e2 / g, kD2);
  }
  return stale;
}

// Record bytes of step s (ring slot cur, diagonals [lo, hi]) when an input
// is a wavefront the cut-off emptied: each byte becomes probed_record's.
template <int NC, int BLOCK, bool CUT, typename RingPtr>
__device__ __forceinline__ void probed_records(RingPtr ring, const MetaRegs<NC, CUT>& meta, int s, const RingPos& rp,
                                               const Penalties& pen, int wcap, int lo, int hi, uint8_t* rec) {
  Src c[7], b[7];
  if (!probe_sources<NC, CUT>(meta, s, rp, pen, wcap, c, b)) return;
  const auto* outM = ring + ring_base(pen, wcap, kM, rp, 0) - lo;
  for (int k = lo + (int)threadIdx.x; k <= hi; k += BLOCK)
    rec[k - lo] = (uint8_t)probed_record<NC>(ring, b[0], b[1], b[2], b[3], b[4], b[5], b[6], c[0], c[1], c[2], c[3],
                                             c[4], c[5], c[6], k, (int)outM[k]);
}

// Record byte of the end cell (s, k) of a z-drop, whose diagonal may lie
// outside the computed range of score s: the choices over the sources the
// reference's backtrace probes from it, or 0xFF when none is valid (the
// backtrace then takes no step, wavefront_backtrace.c:265-266, 470-471).
template <int NC, bool CUT, typename RingPtr>
__device__ __forceinline__ int end_record(RingPtr ring, const MetaRegs<NC, CUT>& meta, int s, const RingPos& rp,
                                          const Penalties& pen, int wcap, int k) {
  Src c[7], b[7];
  probe_sources<NC, CUT>(meta, s, rp, pen, wcap, c, b);
  const unsigned long long key = choice_key<NC>(ring, b[0], b[1], b[2], b[3], b[4], b[5], b[6], k);
  return (int)(unsigned)(key >> 32) >= 0 ? (int)(key & 0x7Fu) : 0xFF;
}

// Ring element stores.  The 16-bit LDS ring saturates to [-16384, 16383]:
// values below 0 are all NULL-like and values above the text length are all
// out of bounds, so every comparison the recurrence makes keeps its outcome
// (only used when both sequences are shorter than 8192).
__device__ __forceinline__ int16_t to_ring(int v, int16_t*) {
  return (int16_t)(v < -16384 ? -16384 : (v > 16383 ? 16383 : v));
}
__device__ __forceinline__ int to_ring(int v, int*) { return v; }

__device__ __forceinline__ bool in_bounds(int off, int k, int plen, int tlen) {
  return ((unsigned)off <= (unsigned)tlen) & ((unsigned)(off - k) <= (unsigned)plen);  // no branch
}

// 64-bit window of symbols starting at symbol `pos` (BITS bits per symbol,
// little-endian inside each word).  Reads words w and w+1: every packed
// sequence carries at least one padding word.
template <int BITS, typename WordPtr>
__device__ __forceinline__ uint64_t window(WordPtr W, int pos) {
  constexpr unsigned SPW = 64 / BITS;
  const unsigned w = (unsigned)pos / SPW;
  const unsigned sh = ((unsigned)pos % SPW) * BITS;
  const uint64_t lo = W[w];
  const uint64_t hi = W[w + 1];
  // (hi << 1) << (63 - sh) is hi << (64 - sh) for sh > 0 and 0 for sh == 0.
  return (lo >> sh) | ((hi << 1) << (63 - sh));
}

// Longest common extension of pattern[v..] and text[h..], bounded by both ends.
// The pattern starts at symbol pb of P and the text at symbol tb of T.
template <int BITS, typename WordPtr>
__device__ __forceinline__ int lcp_extend(WordPtr P, int pb, WordPtr T, int tb, int plen, int tlen, int k, int h) {
  constexpr int SPW = 64 / BITS;
  int v = h - k;
  for (;;) {
    const int rem = min(plen - v, tlen - h);
    if (rem <= 0) return h;
    const uint64_t x = window<BITS>(P, pb + v) ^ window<BITS>(T, tb + h);
    const int n = x ? (int)(__builtin_ctzll(x) / BITS) : SPW;
    if (n >= rem) return h + rem;
    h += n;
    v += n;
    if (n < SPW) return h;
  }
}

// wavefront_termination_endsfree (wavefront_termination.c:115-162) for the
// valid, extended cell at offset off of diagonal k.
__device__ __forceinline__ bool endsfree_done(int off, int k, int plen, int tlen, int pef, int tef) {
  const int v = off - k;
  return (off >= tlen && plen - v <= pef) || (v >= plen && tlen - off <= tef);
}

struct WaveRange {  // first/last valid diagonal seen by one wave (uniform)
  int first, last;
  __device__ __forceinline__ void init() { first = INT_MAX; last = INT_MIN; }
  __device__ __forceinline__ void add(unsigned long long ballot, int k_lane0) {
    if (ballot) {
      if (first == INT_MAX) first = k_lane0 + (int)__builtin_ctzll(ballot);
      last = k_lane0 + 63 - (int)__builtin_clzll(ballot);
    }
  }
  __device__ __forceinline__ void merge(int f, int l) { first = min(first, f); last = max(last, l); }
  __device__ __forceinline__ bool null() const { return first == INT_MAX; }
  // Trimmed range as the reference stores it: [first valid, last valid], or (lo, lo-1).
  __device__ __forceinline__ int tlo(int lo) const { return null() ? lo : first; }
  __device__ __forceinline__ int thi(int lo) const { return null() ? lo - 1 : last; }
};

// wavefront_compute_edit_exact_prune (wavefront_compute_edit.c:219-278) on
// the trimmed M range r of an edit-distance wavefront.  The reference's best
// case |(plen - v) - (tlen - h)| of a diagonal does not depend on the offset:
// it is |k - kend|.  Its two scans therefore keep [max(lo, kend - W),
// min(hi, kend + W)], W being the least worst case max(plen - v, tlen - h)
// over the valid pre-extension cells (wmin, uniform).  The prune is tried when
// the range holds at least kEditPruneWidth diagonals and the middle diagonal's
// pre-extension offset, recomputed here from the previous wavefront, is valid
// and bounds both ends' best cases from below.
template <typename RingPtr>
__device__ __forceinline__ void edit_prune(WaveRange& r, RingPtr ring, const Src& prev, int wmin, int plen,
                                           int tlen) {
  const int lo = r.first, hi = r.last;
  if (hi - lo + 1 < kEditPruneWidth) return;
  const int ks = lo + (hi - lo) / 2;
  int m = max(rd(ring, prev, ks + 1), max(rd(ring, prev, ks - 1), rd(ring, prev, ks)) + 1);
  if (!in_bounds(m, ks, plen, tlen)) m = kOffsetNull;
  m = uniform(m);
  if (m < 0) return;
  const int kend = tlen - plen;
  const int smax = max(plen - (m - ks), tlen - m);
  if (abs(lo - kend) <= smax && abs(hi - kend) <= smax) return;
  r.first = max(lo, kend - wmin);
  r.last = min(hi, kend + wmin);
}

template <int BLOCK>
struct BlockRed {
  static constexpr int NW = BLOCK / 64;
  int* sh;                 // NW ints of LDS
  // Block-wide min of per-thread values (uniform result).
  __device__ __forceinline__ int min_all(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    v = uniform(v);  // every lane holds the minimum
    if constexpr (NW == 1) return v;
    const int lane = threadIdx.x & 63, wid = uniform(threadIdx.x >> 6);
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    int r = sh[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) r = min(r, sh[w]);
    return uniform(r);
  }
  __device__ __forceinline__ int max_all(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    v = uniform(v);  // every lane holds the maximum
    if constexpr (NW == 1) return v;
    const int lane = threadIdx.x & 63, wid = uniform(threadIdx.x >> 6);
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    int r = sh[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) r = max(r, sh[w]);
    return uniform(r);
  }
};

// Distance of a NULL cell (-WAVEFRONT_OFFSET_NULL, wavefront_heuristic.c:125-133).
constexpr int kNullDistance = 1 << 30;

// Float to int conversion as x86's cvttss2si does it (the reference's
// host): truncation, and INT_MIN for NaN and out-of-range values.
__device__ __forceinline__ int trunc_to_int(float f) {
  return (f >= -2147483648.0f && f < 2147483648.0f) ? (int)f : INT_MIN;
}

// Distance of cell (k, off) to the end: max(plen - v, tlen - h)
// (wf_distance_end2end, wavefront_heuristic.c:125-133) or, weighted (wfmash,
// :134-145), what is left of each sequence as a fraction of its length
// scaled to the mean length mfactor, in IEEE single precision and truncated;
// a NULL cell is kNullDistance.
__device__ __forceinline__ int end_distance(int off, int k, int plen, int tlen, bool weighted, int mfactor) {
  if (off < 0) return kNullDistance;
  const int v = off - k, h = off;
  if (!weighted) return max(plen - v, tlen - h);
  const int left_v = trunc_to_int((float)(plen - v) / (float)plen * (float)mfactor);
  const int left_h = trunc_to_int((float)(tlen - h) / (float)tlen * (float)mfactor);
  return max(left_v, left_h);
}

// wf-adaptive cut-off (wavefront_heuristic_wfadaptive and its reduce step,
// wavefront_heuristic.c:190-290) of the extended M wavefront [lo, hi], whose
// diagonal k holds m[k]: keeps the diagonals whose distance to the end (plain,
// or weighted for wfmash) is within thr of the least one, scanning in from
// both ends but never past the end diagonal.  Two block passes: the least
// distance, then the first and last diagonals within thr of it.  The kept
// range (uniform) goes to *clo, *chi.
template <int BLOCK, typename RingPtr>
__device__ void wfadaptive_cut(RingPtr m, int lo, int hi, int plen, int tlen, int thr, bool weighted,
                               BlockRed<BLOCK>& red, int* clo, int* chi) {
  const int tid = threadIdx.x;
  const int mfactor = weighted ? trunc_to_int((float)(plen + tlen) / 2.0f) : 0;
  int dmin = max(plen, tlen);
  for (int k = lo + tid; k <= hi; k += BLOCK) dmin = min(dmin, end_distance((int)m[k], k, plen, tlen, weighted, mfactor));
  dmin = red.min_all(dmin);
  int first = INT_MAX, last = INT_MIN;
  for (int k = lo + tid; k <= hi; k += BLOCK) {
    const int d = end_distance((int)m[k], k, plen, tlen, weighted, mfactor);
    if (d - dmin <= thr) { first = min(first, k); last = max(last, k); }
  }
  first = red.min_all(first);
  last = red.max_all(last);
  const int kend = tlen - plen;
  const int top = min(kend, hi);
  const int lo_r = lo >= top ? lo : min(first, top);
  const int bottom = max(kend, lo_r);
  *clo = lo_r;
  *chi = hi <= bottom ? hi : max(last, bottom);
}

// x-drop / z-drop state of one alignment (wavefront_heuristic_clear,
// wavefront_heuristic.c:114-121): the best Smith-Waterman score seen, once
// there is one, with its cell (z-drop).
struct DropState {
  int best, k, off;
  bool has;
  __device__ __forceinline__ void clear() { best = 0; k = 0; off = kOffsetNull; has = false; }
};

// x-drop cut-off (wavefront_heuristic_xdrop, wavefront_heuristic.c:294-370)
// of the extended M wavefront [*clo, *chi] of score s: Smith-Waterman scores
// (swg_match * (v + h) - s) / 2 of the valid cells; once a best score has
// been seen, the range keeps from the first to the last valid diagonal
// scoring less than xdrop below it (none: an empty range [hi + 1, hi]), and
// the best rises to this wavefront's best.  Differences wrap around in 32
// bits as the reference's int arithmetic does.
template <int BLOCK, typename RingPtr>
__device__ void xdrop_cut(RingPtr m, int s, int swg_match, int xdrop, BlockRed<BLOCK>& red, DropState& xs,
                          int* clo, int* chi) {
  const int tid = threadIdx.x;
  const int lo = *clo, hi = *chi;
  auto sw_of = [&](int off, int k) { return (swg_match * ((off - k) + off) - s) / 2; };
  int cmax = INT_MIN;
  for (int k = lo + tid; k <= hi; k += BLOCK) {
    const int off = (int)m[k];
    if (off >= 0) cmax = max(cmax, sw_of(off, k));
  }
  cmax = red.max_all(cmax);
  if (xs.has) {
    const int best = xs.best;
    int first = INT_MAX, last = INT_MIN;
    for (int k = lo + tid; k <= hi; k += BLOCK) {
      const int off = (int)m[k];
      if (off >= 0 && (int)((unsigned)best - (unsigned)sw_of(off, k)) < xdrop) {
        first = min(first, k);
        last = max(last, k);
      }
    }
    first = red.min_all(first);
    last = red.max_all(last);
    if (first == INT_MAX) {
      *clo = hi + 1;  // both scans pass every diagonal
    } else {
      *clo = first;
      *chi = last;
    }
    if (cmax > xs.best) xs.best = cmax;
  } else {
    xs.best = cmax;
    xs.has = true;
  }
}

// z-drop test (wavefront_heuristic_zdrop, wavefront_heuristic.c:371-431) of
// the extended M wavefront [lo, hi] of score s: its best Smith-Waterman score
// (first diagonal in ascending order) replaces a lower best seen, with its
// cell; one more than zdrop below the best drops the alignment (true), its
// end the best cell.  Differences wrap around in 32 bits.
template <int BLOCK, typename RingPtr>
__device__ bool zdrop_check(RingPtr m, int lo, int hi, int s, int swg_match, int zdrop, BlockRed<BLOCK>& red,
                            DropState& ds) {
  const int tid = threadIdx.x;
  auto sw_of = [&](int off, int k) { return (swg_match * ((off - k) + off) - s) / 2; };
  int cmax = INT_MIN;
  for (int k = lo + tid; k <= hi; k += BLOCK) {
    const int off = (int)m[k];
    if (off >= 0) cmax = max(cmax, sw_of(off, k));
  }
  cmax = red.max_all(cmax);
  int cmax_k = INT_MAX;
  for (int k = lo + tid; k <= hi; k += BLOCK) {
    const int off = (int)m[k];
    if (off >= 0 && sw_of(off, k) == cmax) cmax_k = min(cmax_k, k);
  }
  cmax_k = red.min_all(cmax_k);
  const int cmax_off = cmax_k == INT_MAX ? 0 : uniform((int)m[cmax_k]);
  if (cmax_k == INT_MAX) cmax_k = 0;  // no valid cell: the reference's (0, 0)
  if (ds.has) {
    if (cmax > ds.best) {
      ds.best = cmax;
      ds.k = cmax_k;
      ds.off = cmax_off;
    } else if ((int)((unsigned)ds.best - (unsigned)cmax) > zdrop) {
      return true;
    }
  } else {
    ds.best = cmax;
    ds.k = cmax_k;
    ds.off = cmax_off;
    ds.has = true;
  }
  return false;
}

// Banded adaptive cut-off (wavefront_heuristic_banded_adaptive,
// wavefront_heuristic.c:460-500) of the extended M wavefront [*clo, *chi],
// whose diagonal k holds m[k]: once the step countdown has run out, a
// wavefront of at least 4 diagonals that is wider than the band keeps a
// band-wide window, moved up by half the excess when the low end is farther
// from the end than the high end, and again when the first quarter is
// farther than the middle.  Every lane reads the same four cells.
template <typename RingPtr>
__device__ __forceinline__ void banded_adaptive_cut(RingPtr m, int band_lo, int band_hi, int plen, int tlen,
                                                    int& steps_wait, int steps, int* clo, int* chi) {
  if (steps_wait > 0) return;
  const int lo = *clo, hi = *chi;
  const int len = hi - lo + 1;
  if (len < 4) return;
  const int max_len = band_hi - band_lo + 1;
  if (len > max_len) {
    auto dist = [&](int k) {
      const int off = (int)m[k];
      return off >= 0 ? max(plen - (off - k), tlen - off) : kNullDistance;
    };
    const int leeway = (len - max_len) / 2, quarter = len / 4;
    const int d0 = dist(lo), d1 = dist(lo + quarter), d2 = dist(lo + 2 * quarter), d3 = dist(hi);
    int new_lo = lo;
    if (d0 > d3) new_lo += leeway;
    if (d1 > d2) new_lo += leeway;
    *clo = max(new_lo, lo);
    *chi = min(new_lo + max_len - 1, hi);
  }
  steps_wait = steps;
}

// Orders LDS traffic between the lanes of the (single) wave of a workgroup:
// a wave executes its LDS instructions in program order, so only the
// compiler has to be kept from moving accesses across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace wfa
