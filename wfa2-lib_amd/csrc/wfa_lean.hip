// Lean forward kernel for gap-affine alignment of short pairs on CDNA4.
//
// Same recurrence, trimming, termination and backtrace records as
// wfa_forward_kernel (wavefront_compute_affine.c:44-86,
// wavefront_compute.c:40-86 and 579-613, wavefront_extend_kernels.c:68-123,
// wavefront_termination.c:37-114), specialised on the slot structure of the
// penalties: with g = gcd(x, o+e, e) the sources of score s are the slots
// (s - x)/g, (s - o - e)/g and (s - e)/g steps back in a ring of
// R = (max_score_scope - 1)/g + 1 slots.  When R and these three distances are
// template constants the step loop is unrolled R times, every slot index is a
// compile-time constant and the ring metadata (origin and trimmed [lo, hi] of
// M, I1, D1 per slot) lives in scalar registers: no lane shuffles or memory
// round trips sit between two score steps.
//
// One 64-lane wavefront per pair, persistent over the work queue; the ring
// (16-bit, saturated like the LDS family of wfa_forward_kernel) and both 2-bit
// packed sequences sit in LDS.  After each step the two diagonals just outside
// the trimmed range of every component are overwritten with NULL, so a read
// of a source at diagonal k is one clamp of k to [lo-1, hi+1]: everything
// outside the trimmed range reads NULL, as wavefront_compute_init_ends
// (wavefront_compute.c:498-575) arranges in the reference.  The extension
// compares 16 bases per step with one v_alignbit per sequence.
#include "wfa_common.h"

namespace wfa {
namespace {

constexpr int16_t kRingNull = -16384;  // to_ring(kOffsetNull)

// A compute input: its contribution to the output limits (1, -1 for the null
// wavefront, wavefront_compute.c:40-86) and its read window.  M inputs clamp
// k into [lo-1, hi+1] of their computed range: M cells outside the trimmed
// range hold NULL, and the two cells just outside the computed range are
// written NULL before each step.  I1/D1 inputs hold raw values outside their
// trimmed range, so a read checks the range (wavefront_compute_init_ends,
// wavefront_compute.c:498-575, gives the reference the same NULLs).
//
// Every empty range is kept as (1, -1), the null wavefront's limits, so a
// range is its own limits contribution whether or not it is empty.
struct LeanSrc {
  int llo, lhi;   // limits contribution; valid iff llo <= lhi
  int clo, chi;   // reads clamp k into [clo, chi]
  int sb;         // ring element of diagonal 0
};

// M input of a slot whose ring origin is org, with limits [lo, hi], read
// through a clamp to [cl, ch] (cells holding NULL at both bounds).
__device__ __forceinline__ LeanSrc lean_src_m(int cl, int ch, int lo, int hi, int base, int org) {
  LeanSrc r;
  const bool v = lo <= hi;
  r.llo = lo;
  r.lhi = hi;
  r.clo = v ? cl : 0;
  r.chi = v ? ch : 0;
  r.sb = v ? base - org : 0;  // element 0 of the ring is a NULL cell
  return r;
}

// Without a cut-off the ring element of diagonal 0 is kept per slot, chosen
// when the slot is written (mb: M; gb: the I1 or D1 cells of the last
// step): an empty range (1, -1) then reads element 0, the NULL cell, with
// no selects per read.  M reads clamp to [lo - 1, hi + 1], i.e. [0, 0] when
// empty (element mb = 0).  Gap reads clamp to [lo, hi] and take NULL unless
// the clamp kept k; empty, the clamp gives -1 and gb = 1 reads element 0.
__device__ __forceinline__ LeanSrc lean_src_m_kept(int lo, int hi, int mb) {
  return LeanSrc{lo, hi, lo - 1, hi + 1, mb};
}
__device__ __forceinline__ LeanSrc lean_src_g_kept(int lo, int hi, int gb) { return LeanSrc{lo, hi, lo, hi, gb}; }

__device__ __forceinline__ int lean_rd(const int16_t* ring, const LeanSrc& s, int k) {
  return (int)ring[s.sb + min(max(k, s.clo), s.chi)];
}

// A gap input read: NULL outside its trimmed range.
__device__ __forceinline__ int lean_rd_g(const int16_t* ring, const LeanSrc& s, int k) {
  const int c = min(max(k, s.clo), s.chi);
  const int v = (int)ring[s.sb + c];
  return c == k ? v : (int)kRingNull;
}

// Minimum over the 64 lanes, uniform: prefix minima within each row of 16
// lanes by row shifts of 1, 2, 4 and 8 (a lane without a source keeps
// INT_MAX), then the rows joined by the row broadcasts of lanes 15 and 31;
// lane 63 ends with the minimum of all.
__device__ __forceinline__ int wave_min(int v) {
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x142, 0xa, 0xf, false));  // row_bcast:15 into rows 1, 3
  v = min(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x143, 0xc, 0xf, false));  // row_bcast:31 into rows 2, 3
  return __builtin_amdgcn_readlane(v, 63);
}

// The choices of cell k over clamp-read sources (winning M offset, its
// source, the gap-extension flags), ties as the compute step breaks them;
// reads below 0 count as the one NULL offset (the generic kernel's
// choice_key over the lean ring).
__device__ __forceinline__ unsigned long long lean_choice_key(const int16_t* ring, const LeanSrc& mx, const LeanSrc& mo,
                                                              const LeanSrc& i1, const LeanSrc& d1, int k) {
  auto rdn = [&](const LeanSrc& sr, int kk) {
    const int v = lean_rd(ring, sr, kk);
    return v < 0 ? kOffsetNull : v;
  };
  int best = rdn(mx, k) + 1;
  uint32_t src = kSrcX;
  const int oml = rdn(mo, k - 1), omr = rdn(mo, k + 1), eil = rdn(i1, k - 1), edr = rdn(d1, k + 1);
  const uint32_t bits = (eil >= oml ? kRecI1Ext : 0u) | (edr >= omr ? kRecD1Ext : 0u);
  const int del1 = max(omr, edr), ins1 = max(oml, eil) + 1;
  if (del1 > best) { best = del1; src = kSrcD1; }
  if (ins1 > best) { best = ins1; src = kSrcI1; }
  return ((unsigned long long)(unsigned)best << 32) | (src | bits);
}

// wf-adaptive cut-off (wavefront_heuristic_wfadaptive, wavefront_heuristic.c:
// 190-290; weighted: wfmash, :134-145) of the extended M wavefront [lo, hi]
// of one wave (m[k] its ring cells): the least distance to the end over the
// wave, then the first and last diagonals within thr of it by ballot, never
// scanning past the end diagonal; the same result as the generic kernel's
// wfadaptive_cut.
__device__ __forceinline__ void lean_wfadaptive(const int16_t* m, int lo, int hi, int plen, int tlen, int thr,
                                                bool weighted, int* clo, int* chi) {
  const int lane = threadIdx.x;
  const int mfactor = weighted ? trunc_to_int((float)(plen + tlen) / 2.0f) : 0;
  int dmin = max(plen, tlen);
  for (int kb = lo; kb <= hi; kb += 64) {
    const int k = kb + lane;
    if (k <= hi) dmin = min(dmin, end_distance((int)m[k], k, plen, tlen, weighted, mfactor));
  }
  dmin = wave_min(dmin);
  int first = INT_MAX, last = INT_MIN;
  for (int kb = lo; kb <= hi; kb += 64) {
    const int k = kb + lane;
    const bool keep = k <= hi && end_distance((int)m[k], k, plen, tlen, weighted, mfactor) - dmin <= thr;
    const unsigned long long bal = __ballot(keep);
    if (bal) {
      if (first == INT_MAX) first = kb + (int)__builtin_ctzll(bal);
      last = kb + 63 - (int)__builtin_clzll(bal);
    }
  }
  const int kend = tlen - plen;
  const int top = min(kend, hi);
  const int lo_r = lo >= top ? lo : min(first, top);
  const int bottom = max(kend, lo_r);
  *clo = lo_r;
  *chi = hi <= bottom ? hi : max(last, bottom);
}

// lean_wfadaptive for a step of at most 64 diagonals from registers: lane l
// holds d, the distance of diagonal lo + l computed in the pass (NULL and
// absent cells: kNullDistance); only [mlo, mhi] takes part, as there.
__device__ __forceinline__ void lean_wfadaptive_regs(int d, int lo, int mlo, int mhi, int plen, int tlen, int thr,
                                                     int* clo, int* chi) {
  const int k = lo + (int)threadIdx.x;
  const bool in = k >= mlo && k <= mhi;
  const int dmin = min(max(plen, tlen), wave_min(in ? d : INT_MAX));
  const unsigned long long bal = __ballot(in && d - dmin <= thr);
  const int first = bal ? lo + (int)__builtin_ctzll(bal) : INT_MAX;
  const int last = bal ? lo + 63 - (int)__builtin_clzll(bal) : INT_MIN;
  const int kend = tlen - plen;
  const int top = min(kend, mhi);
  const int lo_r = mlo >= top ? mlo : min(first, top);
  const int bottom = max(kend, lo_r);
  *clo = lo_r;
  *chi = mhi <= bottom ? mhi : max(last, bottom);
}

// Matching symbols at the start of a 16-symbol window whose xor is x, capped
// at cap: v_ffbl gives -1 for x == 0, which the shift makes 0x7fffffff, so
// one three-way minimum takes both the full window (16) and the cap.  (As a
// count-trailing-zeros the compiler would add a compare and a select for
// x == 0; the instruction is written out so it keeps its own zero result.)
__device__ __forceinline__ int run16(uint32_t x, int cap) {
  uint32_t f;
  asm("v_ffbl_b32 %0, %1" : "=v"(f) : "v"(x));
  const int t = (int)(f >> 1);
  return min(min(t, 16), cap);
}

// 16 symbols of a 2-bit packed sequence starting at symbol pos (dword view).
__device__ __forceinline__ uint32_t win16(const uint32_t* W, int pos) {
  const unsigned w = (unsigned)pos >> 4;
  return __builtin_amdgcn_alignbit(W[w + 1], W[w], ((unsigned)pos & 15u) * 2u);
}

__device__ __forceinline__ int lcp16(const uint32_t* P, int pb, const uint32_t* T, int tb, int plen, int tlen,
                                     int k, int h) {
  int v = h - k;
  for (;;) {
    const int rem = min(plen - v, tlen - h);
    if (rem <= 0) return h;
    const uint32_t x = win16(P, pb + v) ^ win16(T, tb + h);
    const int n = run16(x, 16);
    if (n >= rem) return h + rem;
    h += n;
    v += n;
    if (n < 16) return h;
  }
}

// The kernel arguments, reread from the kernel-argument segment where a
// value is needed outside the step loop (the asm makes each pointer opaque,
// so the compiler reloads instead of keeping every field in scalar
// registers across the loop).
using KernArgs = const __attribute__((address_space(4))) ForwardArgs*;
__device__ __forceinline__ KernArgs args_now() {
#if defined(__HIP_DEVICE_COMPILE__)
  KernArgs p = (KernArgs)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return p;
#else
  return nullptr;  // host pass of the single-source compile: never executed
#endif
}

// Ring elements before the packed sequences in LDS (16-byte aligned): M in
// R slots, I1 and D1 in two slots each (they are read one step later only).
template <int R, int WCAP>
constexpr int lean_ring_elems() {
  return ((2 + (R + 4) * WCAP) + 7) / 8 * 8;
}

// EF: ends-free form (score-0 wavefront over [-pbf, tbf], the alignment ends
// at the first diagonal whose extended cell satisfies the free ends).
// CUT: the heuristic cut-offs other than the drops: wf-adaptive or wfmash,
// then a static or adaptive band (wavefront_heuristic.c:190-290, 451-561).
// The slot ranges then hold what the cut-off left (the limits of the
// wavefronts computed from them), while reads still see every cell of the
// trimmed range: gap cells outside it are overwritten with NULL after each
// step, so every input is read with one clamp, like M.  `emptied` marks the
// wavefronts a cut-off emptied that hold readable cells: the reference's
// backtrace probes those (wavefront_backtrace.c:64-219), so a step reading
// one rewrites its record bytes with the probed choices, as the generic
// kernel's probed_records does.
template <int R, int DQX, int DQO, int DQE, int WCAP, bool REC, bool EF, bool CUT>
__global__ void __launch_bounds__(64) wfa_forward_lean_kernel(ForwardArgs a) {
  // The ring is a fixed-size LDS array, whose address the compiler knows
  // (an address in dynamic LDS is resolved at load time: every ring and
  // sequence address then costs an add of it); the sequences, sized by the
  // batch's longest pair, follow in dynamic LDS.
  __shared__ __align__(16) int16_t ring_lds[lean_ring_elems<R, WCAP>()];
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NC = 3;
  static_assert(R * NC <= 32, "emptied mask");
  static_assert(DQX > 0 && DQO > 0 && DQE > 0 && DQX < R && DQO < R && DQE < R, "slot distances");
  const int lane = threadIdx.x;
  const int g = a.pen.g;
  constexpr int wcap = WCAP;  // ring elements per slot and component (>= width + 2)

  // LDS: the ring (element 0: NULL cell read by absent sources; slots from
  // element 2), then both sequences.
  int16_t* ring = ring_lds;
  uint64_t* seqP64 = reinterpret_cast<uint64_t*>(smem);
  uint64_t* seqT64 = seqP64 + a.lmax_words;
  const uint32_t* seqP = reinterpret_cast<const uint32_t*>(seqP64);
  const uint32_t* seqT = reinterpret_cast<const uint32_t*>(seqT64);
  static_assert(DQE == 1, "the gap rings hold two steps");
  auto mbase = [&](int slot) { return 2 + slot * wcap; };
  // I1/D1 ring slot of this step (cur) or of the previous one: the step
  // index's parity, kept in par.
  int par = 0;
  auto gbase = [&](int c, bool cur) { return 2 + (R + (c == kD1 ? 2 : 0) + (cur ? par : par ^ 1)) * wcap; };
  if (lane == 0) ring[0] = kRingNull;

  uint32_t chunk_cur = 0, chunk_end = 0;  // record chunk of this wave (16-byte units)
  unsigned long long wg_dsteps = 0;

  for (;;) {
    int qv = 0;
    KernArgs A = args_now();
    if (lane == 0) qv = (int)atomicAdd(A->queue, 1u);
    const int qi = lane_of(qv, 0);
    if (qi >= A->n_pairs) break;
    const int p = A->order ? A->order[qi] : qi;
    const int plen = A->plen[p];
    const int tlen = A->tlen[p];
    const int kend = tlen - plen;
    const int s_cap = (A->pen.x + A->pen.o1 + A->pen.e1 + 2) * (plen + tlen + 2);
    // One bound per step: the step limit or the (never reached) score cap.
    const int s_stop = min(A->max_steps, s_cap);
    const long long ppos = A->pat_pos[p], tpos = A->txt_pos[p];
    const int pb = (int)(ppos & 31), tb = (int)(tpos & 31);
    {
      const uint64_t* gP = A->words + (ppos >> 5);
      const uint64_t* gT = A->words + (tpos >> 5);
      const int nwp = (pb + plen) / 32 + 2, nwt = (tb + tlen) / 32 + 2;
      wave_sync();  // the previous pair is done reading the sequences
      for (int i = lane; i < nwp; i += 64) seqP64[i] = gP[i];
      for (int i = lane; i < nwt; i += 64) seqT64[i] = gT[i];
      wave_sync();
    }

    // Ring metadata (uniform): origin and trimmed ranges per slot; without a
    // cut-off, the ring elements of diagonal 0 (s_mb per slot, s_gb for the
    // I1 and D1 cells of the last step; see lean_src_m_kept).
    int s_org[R], s_top[R], s_lo[R][NC], s_hi[R][NC];
    int s_mb[R], s_gb[2] = {1, 1};
#pragma unroll
    for (int r = 0; r < R; ++r) {
      s_org[r] = 0;
      s_top[r] = 0;
      s_mb[r] = 0;
#pragma unroll
      for (int c = 0; c < NC; ++c) { s_lo[r][c] = 1; s_hi[r][c] = -1; }
    }
    int s = 0;
    par = 0;
    int status = 1;
    uint32_t prev_blk = kNoBlock;
    int dsteps = 0;  // below 2^31 for pairs of at most 1,024 bases
    int end_k = kend, end_off = tlen;
    int steps_wait = a.heur.steps;  // wavefront_heuristic_clear (wavefront_heuristic.c:114-121)
    const bool wmash = CUT && a.heur.wfadaptive == 2;  // weighted distances (wfmash)
    const int mfactor = wmash ? trunc_to_int((float)(plen + tlen) / 2.0f) : 0;
    int nulls = 0;                  // null steps in a row, every score counted (align_status.num_null_steps)
    uint32_t emptied = 0;           // bit slot*3+c: emptied by a cut-off, readable by the backtrace
    // wavefront_heuristic_cufoff (wavefront_heuristic.c:506-561) on the
    // trimmed M range [mlo, mhi] (m[k]: its cells), one step countdown for
    // every strategy: wf-adaptive / wfmash cut every steps-th wavefront of at
    // least min_len diagonals; then the band: static clamps every time,
    // adaptive moves a band-wide window.  The kept range goes to clo, chi.
    // A step of at most 64 diagonals (one pass, first diagonal plo) passes
    // the distances its lanes computed (dl); otherwise the ring is read.
    auto cut_range = [&](const int16_t* m, int mlo, int mhi, int& clo, int& chi, bool one_pass, int plo, int dl) {
      KernArgs H = args_now();
      --steps_wait;
      clo = mlo;
      chi = mhi;
      const int wa = H->heur.wfadaptive;
      if (wa && steps_wait <= 0 && mhi - mlo + 1 >= H->heur.min_len) {
        if (one_pass) lean_wfadaptive_regs(dl, plo, mlo, mhi, plen, tlen, H->heur.thr, &clo, &chi);
        else lean_wfadaptive(m, mlo, mhi, plen, tlen, H->heur.thr, wa == 2, &clo, &chi);
        steps_wait = H->heur.steps;
      }
      const int band = H->heur.banded;
      if (band == 1) {
        clo = max(clo, H->heur.band_lo);
        chi = min(chi, H->heur.band_hi);
      } else if (band == 2) {
        banded_adaptive_cut(m, H->heur.band_lo, H->heur.band_hi, plen, tlen, steps_wait, H->heur.steps, &clo, &chi);
      }
    };
    if constexpr (EF) {
      // Ends-free score 0: (v=0, h=k) for 0 < k <= tbf and (v=-k, h=0) for
      // -pbf <= k < 0 (wavefront_aligner.c:252-307), each extended, the range
      // untrimmed; the first diagonal that ends the alignment wins
      // (wavefront_extend_kernels.c:154-196).
      const int lo0 = -a.pbf, hi0 = a.tbf;
      if (hi0 - lo0 + 3 > wcap) {
        status = kDevRingOverflow;
      } else {
        const int b = mbase(0) - (lo0 - 1);  // ring element of diagonal 0
        int hitk = INT_MAX, hitoff = 0;
        for (int kb = lo0; kb <= hi0; kb += 64) {
          const int k = kb + lane;
          const bool act = k <= hi0;
          int m = 0;
          if (act) {
            m = lcp16(seqP, pb, seqT, tb, plen, tlen, k, k > 0 ? k : 0);
            ring[b + k] = to_ring(m, ring);
          }
          const unsigned long long bal = __ballot(act && endsfree_done(m, k, plen, tlen, a.pef, a.tef));
          if (bal && hitk == INT_MAX) {
            const int l = (int)__builtin_ctzll(bal);
            hitk = kb + l;
            hitoff = __builtin_amdgcn_readlane(m, l);
          }
        }
        if (lane < 2) ring[b + (lane ? hi0 + 1 : lo0 - 1)] = kRingNull;
        s_org[0] = lo0 - 1;
        s_top[0] = hi0 + 1;
        s_lo[0][kM] = lo0;
        s_hi[0][kM] = hi0;
        s_mb[0] = b;
        if (hitk != INT_MAX) {
          status = kDevCompleted;
          end_k = hitk;
          end_off = hitoff;
        }
        wave_sync();
        // the heuristic also runs after the score-0 extension
        // (wavefront_extend.c:112-118), on the untrimmed range
        if (CUT && status == 1) {
          int clo, chi;
          cut_range(ring + b, lo0, hi0, clo, chi, false, 0, 0);
          if (clo != lo0 || chi != hi0) {
            if (clo > chi) {
              s_lo[0][kM] = 1; s_hi[0][kM] = -1;
              emptied |= 1u << kM;
            } else {
              s_lo[0][kM] = clo; s_hi[0][kM] = chi;
            }
          }
        }
      }
      wave_sync();
    } else {
      // Score 0: M holds offset 0 at k = 0 (wavefront_aligner.c:252-380), extended.
      const int v0 = uniform(lcp16(seqP, pb, seqT, tb, plen, tlen, 0, 0));
      s_org[0] = -1;
      s_top[0] = 1;
      s_lo[0][kM] = 0;
      s_hi[0][kM] = 0;
      const int b = mbase(0);
      s_mb[0] = b + 1;
      if (lane < 3) ring[b + lane] = (lane == 1) ? to_ring(v0, ring) : kRingNull;
      if (kend == 0 && v0 >= tlen) status = kDevCompleted;
      wave_sync();
      // the heuristic also runs after the score-0 extension
      // (wavefront_extend.c:112-118)
      if (CUT && status == 1) {
        int clo, chi;
        cut_range(ring + b + 1, 0, 0, clo, chi, false, 0, 0);
        if (clo != 0 || chi != 0) {
          if (clo > chi) {
            s_lo[0][kM] = 1; s_hi[0][kM] = -1;
            emptied |= 1u << kM;
          } else {
            s_lo[0][kM] = clo; s_hi[0][kM] = chi;
          }
        }
      }
    }

    if (status != 1) goto pair_done;
    for (;;) {
#pragma unroll
      for (int j = 1; j <= R; ++j) {
        const int ph = j % R;  // slot of this step (compile-time after unrolling)
        s += g;
        par ^= 1;
        if constexpr (CUT) {
          // The scores between two multiples of g hold no wavefront: each is
          // a null step of the reference, whose extension ends the alignment
          // as unreachable once the run exceeds the score scope
          // (wavefront_extend.c:94-101) unless the step limit comes first.
          KernArgs H = args_now();
          const int scope = H->pen.scope;
          if (nulls + g - 1 > scope) {
            s = s - g + scope + 1 - nulls;
            status = (s >= H->max_steps) ? kDevMaxSteps : kDevUnreachable;
            goto pair_done;
          }
          nulls += g - 1;
        }
        if (s >= s_stop) { status = s_stop == s_cap ? kDevBadTrace : kDevMaxSteps; goto pair_done; }
        const int rx = (ph - DQX + R) % R, ro = (ph - DQO + R) % R, re = (ph - DQE + R) % R;
        // Without a cut-off the M ranges are the trimmed ranges, whose
        // neighbouring cells hold NULL (a pad or a cell trimmed away), so the
        // clamp may use them in place of the computed range.
        // With a cut-off every input is clamp-read over its computed range
        // [org, top] (cells outside the trimmed range hold NULL).
        const LeanSrc mx = CUT ? lean_src_m(s_org[rx], s_top[rx], s_lo[rx][kM], s_hi[rx][kM], mbase(rx), s_org[rx])
                               : lean_src_m_kept(s_lo[rx][kM], s_hi[rx][kM], s_mb[rx]);
        const LeanSrc mo = CUT ? lean_src_m(s_org[ro], s_top[ro], s_lo[ro][kM], s_hi[ro][kM], mbase(ro), s_org[ro])
                               : lean_src_m_kept(s_lo[ro][kM], s_hi[ro][kM], s_mb[ro]);
        const LeanSrc i1 = CUT ? lean_src_m(s_org[re], s_top[re], s_lo[re][kI1], s_hi[re][kI1], gbase(kI1, false),
                                            s_org[re])
                               : lean_src_g_kept(s_lo[re][kI1], s_hi[re][kI1], s_gb[0]);
        const LeanSrc d1 = CUT ? lean_src_m(s_org[re], s_top[re], s_lo[re][kD1], s_hi[re][kD1], gbase(kD1, false),
                                            s_org[re])
                               : lean_src_g_kept(s_lo[re][kD1], s_hi[re][kD1], s_gb[1]);
        const bool vx = mx.llo <= mx.lhi, vo = mo.llo <= mo.lhi, vi = i1.llo <= i1.lhi, vd = d1.llo <= d1.lhi;
        if (!(vx || vo || vi || vd)) {  // null step (wavefront_compute_affine.c:236-243)
          if constexpr (CUT) {
            if (++nulls > args_now()->pen.scope) { status = kDevUnreachable; goto pair_done; }
            emptied &= ~(7u << (3 * ph));
          }
#pragma unroll
          for (int c = 0; c < NC; ++c) { s_lo[ph][c] = 1; s_hi[ph][c] = -1; }
          s_mb[ph] = 0;
          s_gb[0] = s_gb[1] = 1;
          continue;
        }
        if constexpr (CUT) nulls = 0;
        // wavefront_compute_limits_input (wavefront_compute.c:40-86)
        const int lo = min(min(mx.llo, mo.llo - 1), min(i1.llo + 1, d1.llo - 1));
        const int hi = max(max(mx.lhi, mo.lhi + 1), max(i1.lhi + 1, d1.lhi - 1));
        const int width = hi - lo + 1;
        if (width + 2 > wcap) { status = kDevRingOverflow; goto pair_done; }
        dsteps += width;
        const bool has_i1 = vo || vi;
        const bool has_d1 = vo || vd;

        uint8_t* rec = nullptr;
        if constexpr (REC) {
          const uint32_t units = (uint32_t)rec_units_for(width);
          if (chunk_cur + units > chunk_end) {
            const uint32_t want = units > kChunkUnits ? units : kChunkUnits;
            unsigned long long b0 = 0;
            KernArgs C = args_now();  // the arena's fields only on this rare path
            if (lane == 0) b0 = atomicAdd(C->rec_top, (unsigned long long)want);
            const unsigned blo = (unsigned)lane_of((int)(unsigned)(b0 & 0xFFFFFFFFull), 0);
            const unsigned bhi = (unsigned)lane_of((int)(unsigned)(b0 >> 32), 0);
            if (bhi != 0 || (unsigned long long)blo + want > C->rec_units) {
              chunk_cur = chunk_end = 0;
              status = kDevArenaFull;
              goto pair_done;
            }
            chunk_cur = blo;
            chunk_end = blo + want;
          }
          const uint32_t blk = chunk_cur;
          chunk_cur += units;
          uint4* const hdr = args_now()->rec_arena + blk;
          if (lane == 0) *hdr = make_uint4((unsigned)s, (unsigned)lo, (unsigned)hi, prev_blk);
          prev_blk = blk;
          rec = reinterpret_cast<uint8_t*>(hdr + 1);
        }

        // Outputs of this step go to slot ph, origin lo - 1.
        const int org = lo - 1;
        int16_t* oM = ring + mbase(ph) - org;
        int16_t* oI = ring + gbase(kI1, true) - org;
        int16_t* oD = ring + gbase(kD1, true) - org;
        // NULL in the two M cells just outside the computed range (and in
        // the gap cells there when every input is clamp-read).
        // (lane < 2 compared here from an opaque copy of the lane: a
        // compare hoisted out of the step loop is kept as a spilled mask)
        int pad_lane = lane;
        asm volatile("" : "+v"(pad_lane));
        if (pad_lane < 2) {
          const int kk = pad_lane ? hi + 1 : lo - 1;
          oM[kk] = kRingNull;
          if constexpr (CUT) {
            oI[kk] = kRingNull;
            oD[kk] = kRingNull;
          }
        }
        int dl = kNullDistance;  // CUT: this lane's distance to the end in the last pass
        // One pass over 64 diagonals starting at kb: compute, record, extend,
        // store; returns the validity ballots and the ballot of the lanes
        // that end the alignment: k_end reaching the end of the text
        // (wavefront_termination.c:37-114), or, ends-free, the free ends
        // (fin_off: the offset of the first such lane).
        auto pass = [&](int kb, unsigned long long& bM, unsigned long long& bI, unsigned long long& bD,
                        int& fin_off) -> unsigned long long {
          const int k = kb + lane;
          const bool act = k <= hi;
          // Reads are clamped, so every lane may read; only active lanes store.
          const int om_l = lean_rd(ring, mo, k - 1);
          const int om_r = lean_rd(ring, mo, k + 1);
          const int ei1 = CUT ? lean_rd(ring, i1, k - 1) : lean_rd_g(ring, i1, k - 1);
          const int ed1 = CUT ? lean_rd(ring, d1, k + 1) : lean_rd_g(ring, d1, k + 1);
          const int ins1 = max(om_l, ei1) + 1;
          const int del1 = max(om_r, ed1);
          int best = lean_rd(ring, mx, k) + 1;
          uint32_t src = kSrcX;
          // argmax with ties to the higher backtrace type: M > D1 > I1
          if (del1 > best) { best = del1; src = kSrcD1; }
          if (ins1 > best) { best = ins1; src = kSrcI1; }
          const uint32_t bits = (ei1 >= om_l ? kRecI1Ext : 0u) | (ed1 >= om_r ? kRecD1Ext : 0u);
          // The record byte is made before the branch of its store (an
          // opaque copy keeps it there): inside, its compare-select chain
          // has nothing to interleave with and waits on every select.
          uint32_t rb = src | bits;
          asm volatile("" : "+v"(rb));
          if (REC && act) rec[(unsigned)(k - lo)] = (uint8_t)rb;
          // null-after-max, then the extension of the valid M cells: one
          // 16-symbol window for every lane, more only for lanes whose run
          // fills it (uniform branch).
          const bool okM = act & in_bounds(best, k, plen, tlen);
          int h = okM ? best : 0;
          int rem = okM ? min(plen - (best - k), tlen - best) : 0;
          int pp = okM ? pb + best - k : 0, pt = okM ? tb + best : 0;
          {
            const uint32_t x = win16(seqP, pp) ^ win16(seqT, pt);
            const int n = run16(x, rem);
            h += n;
            bool more = (n == 16) & (rem > 16);
            // (the AND of two single-compare ballots: a ballot of the combined
            // condition is taken from a 0/1 vector value compared back)
            if (__ballot(n == 16) & __ballot(rem > 16)) {
              while (more) {
                rem -= 16;
                pp += 16;
                pt += 16;
                const uint32_t y = win16(seqP, pp) ^ win16(seqT, pt);
                const int n2 = run16(y, rem);
                h += n2;
                more = n2 == 16 && rem > 16;
              }
            }
          }
          if constexpr (CUT) dl = okM ? end_distance(h, k, plen, tlen, wmash, mfactor) : kNullDistance;
          const bool vI = act & has_i1 & in_bounds(ins1, k, plen, tlen);
          const bool vD = act & has_d1 & in_bounds(del1, k, plen, tlen);
          if (act) {
            oM[k] = to_ring(okM ? h : kOffsetNull, ring);
            if (has_i1) oI[k] = to_ring(ins1, ring);
            if (has_d1) oD[k] = to_ring(del1, ring);
          }
          bM = __ballot(okM);
          bI = __ballot(vI);
          bD = __ballot(vD);
          if constexpr (EF) {
            const unsigned long long fb = __ballot(okM && endsfree_done(h, k, plen, tlen, a.pef, a.tef));
            if (fb) fin_off = __builtin_amdgcn_readlane(h, (int)__builtin_ctzll(fb));
            return fb;
          } else {
            return __ballot(okM & (k == kend) & (h >= tlen));
          }
        };
        // Trimmed ranges (wavefront_compute.c:579-613) from the ballots;
        // absent or empty = (1, -1).  Most steps are at most 64 wide: the
        // first pass is straight-line code, further passes loop.
        unsigned long long bM, bI, bD;
        int fin_off = 0;
        unsigned long long fin = pass(lo, bM, bI, bD, fin_off);
        int fin_k = (EF && fin) ? lo + (int)__builtin_ctzll(fin) : INT_MAX;
        int mlo = bM ? lo + (int)__builtin_ctzll(bM) : 1, mhi = bM ? lo + 63 - (int)__builtin_clzll(bM) : -1;
        int ilo = bI ? lo + (int)__builtin_ctzll(bI) : 1, ihi = bI ? lo + 63 - (int)__builtin_clzll(bI) : -1;
        int dlo = bD ? lo + (int)__builtin_ctzll(bD) : 1, dhi = bD ? lo + 63 - (int)__builtin_clzll(bD) : -1;
        for (int kb = lo + 64; kb <= hi; kb += 64) {
          int off2 = 0;
          const unsigned long long f2 = pass(kb, bM, bI, bD, off2);
          if constexpr (EF) {
            if (f2 && fin_k == INT_MAX) {
              fin_k = kb + (int)__builtin_ctzll(f2);
              fin_off = off2;
            }
          }
          fin |= f2;
          if (bM) { if (mlo > mhi) mlo = kb + (int)__builtin_ctzll(bM); mhi = kb + 63 - (int)__builtin_clzll(bM); }
          if (bI) { if (ilo > ihi) ilo = kb + (int)__builtin_ctzll(bI); ihi = kb + 63 - (int)__builtin_clzll(bI); }
          if (bD) { if (dlo > dhi) dlo = kb + (int)__builtin_ctzll(bD); dhi = kb + 63 - (int)__builtin_clzll(bD); }
        }
        if constexpr (CUT) {
          // gap cells outside their trimmed range read NULL from now on
          for (int kb = lo; kb <= hi; kb += 64) {
            const int k = kb + lane;
            if (k <= hi) {
              if (has_i1 && (k < ilo || k > ihi)) oI[k] = kRingNull;
              if (has_d1 && (k < dlo || k > dhi)) oD[k] = kRingNull;
            }
          }
          if constexpr (REC) {
            // An input a cut-off emptied still has cells the reference's
            // backtrace probes: the record bytes take the probed choices and
            // flag the cells whose trace would take operations back
            // (probed_record of the generic kernel).
            const uint32_t smask = (1u << (3 * rx + kM)) | (1u << (3 * ro + kM)) | (1u << (3 * re + kI1)) |
                                   (1u << (3 * re + kD1));
            if (emptied & smask) {
              wave_sync();  // this step's M cells
              auto bview = [&](const LeanSrc& c, int slot, int comp) {
                LeanSrc b = c;
                if ((emptied >> (3 * slot + comp)) & 1u) {
                  b.clo = s_org[slot];
                  b.chi = s_top[slot];
                  b.sb = (comp == kM ? mbase(slot) : gbase(comp, false)) - s_org[slot];
                }
                return b;
              };
              const LeanSrc bx = bview(mx, rx, kM), bo = bview(mo, ro, kM), bi = bview(i1, re, kI1),
                            bd = bview(d1, re, kD1);
              for (int kb = lo; kb <= hi; kb += 64) {
                const int k = kb + lane;
                if (k <= hi) {
                  const unsigned long long kbk = lean_choice_key(ring, bx, bo, bi, bd, k);
                  const unsigned long long kck = lean_choice_key(ring, mx, mo, i1, d1, k);
                  const int q = (int)(unsigned)(kbk >> 32), best = (int)(unsigned)(kck >> 32);
                  uint32_t r = (uint32_t)kbk & 0x7Fu;
                  if (q >= 0 && q > max((int)oM[k], best)) r |= kRecStale;
                  rec[(unsigned)(k - lo)] = (uint8_t)r;
                }
              }
            }
          }
          emptied &= ~(7u << (3 * ph));
        }
        s_org[ph] = org;
        s_top[ph] = hi + 1;
        s_lo[ph][kM] = mlo; s_hi[ph][kM] = mhi;
        s_lo[ph][kI1] = ilo; s_hi[ph][kI1] = ihi;
        s_lo[ph][kD1] = dlo; s_hi[ph][kD1] = dhi;
        if constexpr (!CUT) {
          s_mb[ph] = mlo <= mhi ? mbase(ph) - org : 0;
          s_gb[0] = ilo <= ihi ? gbase(kI1, true) - org : 1;
          s_gb[1] = dlo <= dhi ? gbase(kD1, true) - org : 1;
        }
        wave_sync();  // ring writes of this step before the next step's reads
        // M[k_end] reached the end of the text: a valid cell at k_end lies in
        // the trimmed range, so the lane that computed it decides.
        if (fin) {
          status = kDevCompleted;
          if constexpr (EF) {
            end_k = fin_k;
            end_off = fin_off;
          }
          goto pair_done;
        }
        if constexpr (CUT) {
          // The cut-off of a non-empty M wavefront; when M changed, I1/D1
          // are narrowed to M's range (wavefront_heuristic.c:163-171).
          if (mlo <= mhi) {
            {
              int clo, chi;
              cut_range(oM, mlo, mhi, clo, chi, hi - lo < 64, lo, dl);
              if (clo != mlo || chi != mhi) {
                if (clo > chi) {
                  s_lo[ph][kM] = 1; s_hi[ph][kM] = -1;
                  emptied |= 1u << (3 * ph + kM);
                } else {
                  s_lo[ph][kM] = clo; s_hi[ph][kM] = chi;
                }
#pragma unroll
                for (int c = 1; c < NC; ++c) {
                  if (s_lo[ph][c] <= s_hi[ph][c]) {
                    const int nlo = max(s_lo[ph][c], clo), nhi = min(s_hi[ph][c], chi);
                    if (nlo > nhi) {
                      s_lo[ph][c] = 1; s_hi[ph][c] = -1;
                      emptied |= 1u << (3 * ph + c);
                    } else {
                      s_lo[ph][c] = nlo; s_hi[ph][c] = nhi;
                    }
                  }
                }
              }
            }
          }
        }
      }
    }
  pair_done:
    {
      KernArgs B = args_now();
      if (lane == 0) {
        B->out_status[p] = status;
        B->out_score[p] = s;
        B->out_last_block[p] = prev_blk;
        B->out_dsteps[p] = (long long)dsteps;
        if (EF && B->out_end_k) {
          B->out_end_k[p] = end_k;
          B->out_end_off[p] = end_off;
        }
        if (status == kDevRingOverflow || status == kDevArenaFull) atomicAdd(B->retry_count, 1u);
        else if (status != kDevCompleted) atomicAdd(B->fail_count, 1u);
      }
    }
    if (status == kDevCompleted || status == kDevMaxSteps || status == kDevUnreachable)
      wg_dsteps += (unsigned long long)dsteps;
  }
  if (lane == 0 && wg_dsteps) atomicAdd(args_now()->dsteps_total, wg_dsteps);
}

// Ring widths with an instance (diagonals per slot and component).
constexpr int kLeanWidths[] = {224, 384, 768};

template <int R, int DQX, int DQO, int DQE, bool EF, bool CUT>
const void* lean_fn(int wcap, bool rec) {
#define WFA_LEAN_W(W)                                                                                      \
  if (wcap == W)                                                                                           \
    return rec ? reinterpret_cast<const void*>(&wfa_forward_lean_kernel<R, DQX, DQO, DQE, W, true, EF, CUT>) \
               : reinterpret_cast<const void*>(&wfa_forward_lean_kernel<R, DQX, DQO, DQE, W, false, EF, CUT>);
  WFA_LEAN_W(224) WFA_LEAN_W(384) WFA_LEAN_W(768)
#undef WFA_LEAN_W
  return nullptr;
}

// Slot structures with an instance: (R, dq_x, dq_o, dq_e).
bool lean_structure(const Penalties& pen) {
  if (pen.g <= 0) return false;
  const int dx = pen.x / pen.g, dq = (pen.o1 + pen.e1) / pen.g, de = pen.e1 / pen.g, R = pen.nslots;
  // gap-affine (4, 6, 2) and every multiple of (2, 3, 1)
  return R == 5 && dx == 2 && dq == 4 && de == 1;
}

const void* select_lean(const Penalties& pen, int wcap, bool rec, bool ef, bool cut) {
  if (!lean_structure(pen)) return nullptr;
  if (cut) return ef ? lean_fn<5, 2, 4, 1, true, true>(wcap, rec) : lean_fn<5, 2, 4, 1, false, true>(wcap, rec);
  return ef ? lean_fn<5, 2, 4, 1, true, false>(wcap, rec) : lean_fn<5, 2, 4, 1, false, false>(wcap, rec);
}

}  // namespace

bool lean_supported(const Penalties& pen) { return lean_structure(pen); }

int lean_width(int wcap) {
  for (int w : kLeanWidths)
    if (w >= wcap) return w;
  return 0;
}

size_t lean_lds_bytes(const Penalties& pen, int wcap, int lmax_words) {
  (void)wcap;  // the ring is static LDS of each instance
  if (!lean_structure(pen)) return 0;
  return (size_t)2 * lmax_words * sizeof(uint64_t);
}

int lean_max_blocks_per_cu(const Penalties& pen, int wcap, size_t lds_bytes) {
  const void* f = select_lean(pen, wcap, true, false, false);
  if (!f) return 0;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 64, lds_bytes) != hipSuccess) return 0;
  return n;
}

hipError_t launch_forward_lean(const ForwardArgs& a, const LaunchShape& shape, hipStream_t stream) {
  const void* f = select_lean(a.pen, a.wcap, a.records != 0, a.endsfree != 0, a.heur.cut() != 0);
  if (!f || shape.block != 64) return hipErrorInvalidValue;
  if (shape.lds_bytes > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shape.lds_bytes);
    if (e != hipSuccess) return e;
  }
  void* args[] = {const_cast<ForwardArgs*>(&a)};
  return hipLaunchKernel(f, dim3(shape.grid), dim3(64), args, shape.lds_bytes, stream);
}

}  // namespace wfa
