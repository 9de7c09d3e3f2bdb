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
template <int NC>
struct MetaRegs {
  int tag;             // score held by the slot, -1 if none
  int alo;             // lo of the computed range: array origin of all components
  unsigned flags;      // bit c: component present; bit 8+c: component null (trimmed empty)
  int lo[NC];
  int hi[NC];
};

// A compute input.  lim_lo/lim_hi are the limits it contributes to the output
// range (the shared null wavefront of the reference has lo = 1, hi = -1).
// Reads at k in [lo, lo+span] return ring[base + k - lo], else NULL; an
// invalid source gets an empty read range far from any diagonal.
struct Src {
  int lim_lo, lim_hi;
  int lo, span, base;
  __device__ __forceinline__ bool valid() const { return lim_lo <= lim_hi; }
};
constexpr int kFarDiagonal = 0x40000000;

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int lane_of(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

template <int NC>
__device__ __forceinline__ Src get_src(const MetaRegs<NC>& m, int s_src, int slot, int R, int c, int wcap) {
  Src r{1, -1, kFarDiagonal, 0, 0};
  if (s_src < 0) return r;
  if (slot < 0) slot += R;
  slot = uniform(slot);
  const int tag = lane_of(m.tag, slot);
  const unsigned fl = (unsigned)lane_of((int)m.flags, slot);
  if (tag != s_src) return r;
  if (!(fl & (1u << c)) || (fl & (0x100u << c))) return r;
  r.lim_lo = r.lo = lane_of(m.lo[c], slot);
  r.lim_hi = lane_of(m.hi[c], slot);
  r.span = r.lim_hi - r.lim_lo;
  r.base = (slot * NC + c) * wcap + (r.lo - lane_of(m.alo, slot));
  return r;
}

// Branch-free predicated read: the load address is always inside the ring.
template <typename RingPtr>
__device__ __forceinline__ int rd(RingPtr ring, const Src& s, int k) {
  const int i = k - s.lo;
  const bool in = (unsigned)i <= (unsigned)s.span;
  const int v = (int)ring[s.base + (in ? i : 0)];
  return in ? v : kOffsetNull;
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
  return (unsigned)off <= (unsigned)tlen && (unsigned)(off - k) <= (unsigned)plen;
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

// Orders LDS traffic between the lanes of the (single) wave of a workgroup:
// a wave executes its LDS instructions in program order, so only the
// compiler has to be kept from moving accesses across this point.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace wfa
