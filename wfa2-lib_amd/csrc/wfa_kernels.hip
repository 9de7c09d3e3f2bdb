// HIP kernels for the exact WFA hot path on CDNA4 (gfx950).
//
// Forward kernel: one pair per workgroup, persistent over a device work queue.
// Each non-null score step s is one fused pass over the diagonals [lo,hi]:
// read the s-x / s-o-e / s-e wavefronts from a ring of slots, compute I1, D1,
// (I2, D2) and M (wavefront_compute_affine.c:44-86 and
// wavefront_compute_affine2p.c:45-106), extend M along exact matches over the
// 2-bit packed sequences (wavefront_extend_kernels.c:68-123), write the step's
// backtrace records, then trim the ends with wave ballots
// (wavefront_compute.c:579-613) and test termination
// (wavefront_termination.c:37-114).
//
// Records replace the reference's stored wavefronts (memory mode high).  The
// reference backtrace (wavefront_backtrace.c:320-529) takes, at every cell it
// visits, the maximum of (offset << 4 | type) over the candidate sources; the
// candidate offsets it probes are exactly the values the forward pass read to
// compute that cell, so the winner (and its tie-break M > D2 > D1 > I2 > I1,
// extension before opening) is decided here once, and stored as 7 bits next
// to the cell's pre-extension M offset.  The backtrace kernel replays them.
#include <cstdio>
#include <cstdlib>

#include "wfa_common.h"

namespace wfa {
namespace {

// CUT: 0 exact; 1 wf-adaptive, wfmash and the bands; 2 also x-drop / z-drop.
// The 1024-thread instance asks for at least 4 waves per SIMD, which leaves
// it up to 128 VGPRs (it uses about 80): one workgroup, one pair, per CU,
// and registers enough for a chunk's ring reads to go out together.
template <int NC, int BLOCK, bool SMALL, int BITS, int CUT>
__global__ void __launch_bounds__(BLOCK, BLOCK == 1024 ? 4 : 1) wfa_forward_kernel(ForwardArgs a) {
  constexpr bool HC = CUT != 0;
  // The LDS family stores the ring as int16, the HBM family as int32.
  using RT = typename std::conditional<SMALL, int16_t, int>::type;
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NW = BLOCK / 64;
  constexpr int SPW = 64 / BITS;
  // One wave and the ring in LDS: no workgroup barrier is ever needed.
  constexpr bool WAVE_ONLY = SMALL && NW == 1;
  __shared__ int sh_bcast[2];
  __shared__ int sh_end[2];  // the end cell's value, by computed-step parity (see the termination test)
  // per-wave ranges of a step, by the parity of the computed steps: a wave
  // already in the next computed step writes the other half, so one barrier
  // per step separates the writes from every wave's merge (sh_end likewise)
  __shared__ int sh_red[2][NW > 1 ? NW : 1][2 * NC + 1];
  __shared__ int sh_hit[NW > 1 ? NW : 1];
  BlockRed<BLOCK> red{sh_hit};

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = uniform(tid >> 6);  // wave index: the same in every lane of a wave
  const Penalties pen = a.pen;
  const int R = pen.nslots;  // <= 64 (checked by the host)
  const int wcap = a.wcap;
  const int g = pen.g;
  const int dq_x = pen.x / g, dq_o1 = (pen.o1 + pen.e1) / g, dq_e1 = pen.e1 / g;
  const int dq_o2 = (NC == 5) ? (pen.o2 + pen.e2) / g : 0;
  const int dq_e2 = (NC == 5) ? pen.e2 / g : 0;

  uint64_t* seqP = reinterpret_cast<uint64_t*>(smem);
  uint64_t* seqT = seqP + a.lmax_words;
  RT* ring;
  if constexpr (SMALL) {
    ring = reinterpret_cast<RT*>(seqT + a.lmax_words);
  } else {
    ring = a.ring_global + (long long)blockIdx.x * a.ring_stride;
  }
  auto block_sync = [&]() {
    if constexpr (WAVE_ONLY) wave_sync(); else __syncthreads();
  };
  // Lane 0 computes v; every thread of the workgroup receives it.
  auto broadcast0 = [&](int v, int slot) -> int {
    if constexpr (NW == 1) {
      return lane_of(v, 0);
    } else {
      if (tid == 0) sh_bcast[slot] = v;
      __syncthreads();
      const int r = uniform(sh_bcast[slot]);
      __syncthreads();
      return r;
    }
  };

  // Heuristic cut-off of the M wavefront of score s in ring slot `slot`,
  // trimmed to [lo, hi] (wavefront_heuristic_cufoff,
  // wavefront_heuristic.c:506-561), with one step countdown for every
  // strategy, in the reference's order: wf-adaptive / wfmash cuts every
  // steps-th non-empty M wavefront of at least min_len diagonals; x-drop cuts
  // on Smith-Waterman scores; then the band: static clamps to [band_lo,
  // band_hi] every time, adaptive moves a band-wide window.  When M changed,
  // I1/D1/I2/D2 are narrowed to M's range (:163-171).  Indel and edit then
  // read M only inside the cut range: their compute step NULLs the two cells
  // beyond each end (wavefront_compute_edit.c:355-358).
  auto cutoff = [&](MetaRegs<NC, HC>& mt, int& steps_wait, DropState& xs, int slot, int alo, int lo, int hi,
                    int s_, int plen_, int tlen_) -> bool {
    if constexpr (CUT) {
      const HeurArgs& hr = a.heur;
      --steps_wait;
      const RT* m = ring + slot * wcap - alo;  // the M ring shares the metadata slots
      int clo = lo, chi = hi;
      if (hr.wfadaptive && steps_wait <= 0 && hi - lo + 1 >= hr.min_len) {
        wfadaptive_cut<BLOCK>(m, lo, hi, plen_, tlen_, hr.thr, hr.wfadaptive == 2, red, &clo, &chi);
        steps_wait = hr.steps;
      }
      if (CUT == 2 && hr.xdrop && steps_wait <= 0) {
        xdrop_cut<BLOCK>(m, s_, hr.swg_match, hr.xdrop_thr, red, xs, &clo, &chi);
        steps_wait = hr.steps;
      } else if (CUT == 2 && hr.zdrop && steps_wait <= 0) {
        // a z-drop ends the alignment here, its countdown untouched
        if (zdrop_check<BLOCK>(m, clo, chi, s_, hr.swg_match, hr.zdrop_thr, red, xs)) return true;
        steps_wait = hr.steps;
      }
      if (hr.banded == 1) {
        clo = max(clo, hr.band_lo);
        chi = min(chi, hr.band_hi);
      } else if (hr.banded == 2) {
        banded_adaptive_cut(m, hr.band_lo, hr.band_hi, plen_, tlen_, steps_wait, hr.steps, &clo, &chi);
      }
      // unchanged M: nothing else is narrowed either (wavefront_heuristic.c:540-541)
      if (clo == lo && chi == hi) return false;
      if (lane == slot) {
        unsigned fl = mt.flags;
        mt.cut.lo[kM] = clo;
        mt.cut.hi[kM] = chi;
        if (clo > chi) fl |= 0x100u;
        if (pen.metric <= kEdit) { mt.lo[kM] = clo; mt.hi[kM] = chi; }
#pragma unroll
        for (int c = 1; c < NC; ++c) {
          if ((fl & (1u << c)) && !(fl & (0x100u << c))) {
            mt.cut.lo[c] = max(mt.cut.lo[c], clo);
            mt.cut.hi[c] = min(mt.cut.hi[c], chi);
            if (mt.cut.lo[c] > mt.cut.hi[c]) fl |= 0x100u << c;
          }
        }
        mt.flags = fl;
      }
    }
    return false;
  };

  unsigned long long chunk_cur = 0, chunk_end = 0;  // record chunk of this workgroup (16-byte units)
  unsigned long long wg_dsteps = 0;                  // diagonal-steps of this workgroup
  MetaRegs<NC, HC> meta;

  for (;;) {
    int qv = 0;
    if (tid == 0) qv = (int)atomicAdd(a.queue, 1u);
    const int qi = broadcast0(qv, 0);
    if (qi >= a.n_pairs) break;
    const int p = a.order ? a.order[qi] : qi;
    const int plen = a.plen[p];
    const int tlen = a.tlen[p];
    const int kend = tlen - plen;
    const int cbeg = a.comp_begin_p ? (int)a.comp_begin_p[p] : a.comp_begin;
    const int cend = a.comp_end_p ? (int)a.comp_end_p[p] : a.comp_end;
    // Generous bound on the end score of any end-to-end alignment; guarantees
    // that every workgroup leaves the loop even on inconsistent input.
    const int s_cap = (pen.x + pen.o1 + pen.e1 + (NC == 5 ? pen.o2 + pen.e2 : 0) + 2) * (plen + tlen + 2);

    // Sequences: word-aligned base pointer plus a symbol offset below SPW.
    const long long ppos = a.pat_pos[p], tpos = a.txt_pos[p];
    const int pb = (int)(ppos % SPW), tb = (int)(tpos % SPW);
    // The LDS family, and the HBM-ring family when the host gave room
    // (lmax_words > 0), read the sequences from LDS: the extension's
    // windows then cost an LDS round trip instead of an L2 one.  The two
    // cases keep separate pointers (seqP/seqT into LDS, gP/gT in global
    // memory) so that each extension is compiled with its memory's loads; a
    // pointer that may be either would make every window a flat load.
    const uint64_t* gP = a.words + ppos / SPW;
    const uint64_t* gT = a.words + tpos / SPW;
    const bool lds_seq = SMALL || a.lmax_words > 0;
    if (lds_seq) {
      const int nwp = (pb + plen) / SPW + 2, nwt = (tb + tlen) / SPW + 2;
      block_sync();  // the previous pair is done reading the sequences
      for (int i = tid; i < nwp; i += BLOCK) seqP[i] = gP[i];
      for (int i = tid; i < nwt; i += BLOCK) seqT[i] = gT[i];
    }
    auto extend = [&](int k, int h) {
      return lds_seq ? lcp_extend<BITS>(seqP, pb, seqT, tb, plen, tlen, k, h)
                     : lcp_extend<BITS>(gP, pb, gT, tb, plen, tlen, k, h);
    };
    // Score 0: the begin component holds offset 0 at k = 0 (wavefront_aligner.c:252-380).
    meta.tag = (lane == 0) ? 0 : -1;
    meta.alo = 0;
    meta.flags = (lane == 0) ? (1u << cbeg) : 0u;
#pragma unroll
    for (int c = 0; c < NC; ++c) { meta.lo[c] = 0; meta.hi[c] = 0; }
    if constexpr (CUT) {
#pragma unroll
      for (int c = 0; c < NC; ++c) { meta.cut.lo[c] = 0; meta.cut.hi[c] = 0; }
    }
    int steps_wait = a.heur.steps;  // wavefront_heuristic_clear (wavefront_heuristic.c:114-121)
    DropState xs;
    xs.clear();
    int cur = 0;
    RingPos rp{0, 0, 0};  // rp.cur == cur: the M ring shares the metadata slots
    int s = 0;
    int status = 1;  // running
    int nb = 0;      // computed (non-null) steps: the parity of sh_red and sh_end
    uint32_t prev_blk = kNoBlock;
    long long dsteps = 0;
    if constexpr (!SMALL) {
      // A long pair starts only if the record arena can still hold about
      // L^2 >> defer_shift units (or is empty); otherwise it is handed back
      // for a later launch before any work (a pair stopped half way wastes
      // the records it wrote).
      if (a.records && a.defer_shift > 0) {
        int defer = 0;
        if (tid == 0) {
          const unsigned long long top = __hip_atomic_load(a.rec_top, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          const unsigned long long L = (unsigned long long)max(plen, tlen);
          defer = top > 0 && top + ((L * L) >> a.defer_shift) > a.rec_units;  // an empty arena always takes it
        }
        if (broadcast0(defer, 1)) status = kDevArenaFull;
      }
    }
    block_sync();  // sequences visible to every lane
    int end_k = kend, end_off = tlen;
    if (status == 1 && a.endsfree) {
      // Ends-free score 0: (v=0, h=k) for 0 < k <= tbf, (v=-k, h=0) for
      // -pbf <= k < 0 (wavefront_aligner.c:252-307), untrimmed; each cell is
      // extended and the first diagonal that ends the alignment wins
      // (wavefront_extend_kernels.c:154-196).
      const int lo0 = -a.pbf, hi0 = a.tbf;
      if (hi0 - lo0 + 1 > wcap) {
        status = kDevRingOverflow;
      } else {
        int hitk = INT_MAX;
        for (int kb = lo0; kb <= hi0; kb += BLOCK) {
          const int k = kb + tid;
          bool hit = false;
          if (k <= hi0) {
            const int m = extend(k, k > 0 ? k : 0);
            ring[(0 * NC + kM) * wcap + (k - lo0)] = to_ring(m, ring);
            hit = endsfree_done(m, k, plen, tlen, a.pef, a.tef);
          }
          const unsigned long long bal = __ballot(hit);
          if (bal && hitk == INT_MAX) hitk = kb + wid * 64 + (int)__builtin_ctzll(bal);
        }
        if (lane == 0) {
          meta.alo = lo0;
          meta.lo[kM] = lo0;
          meta.hi[kM] = hi0;
          if constexpr (CUT) { meta.cut.lo[kM] = lo0; meta.cut.hi[kM] = hi0; }
        }
        hitk = red.min_all(hitk);
        block_sync();
        if (hitk != INT_MAX) {
          status = kDevCompleted;
          end_k = hitk;
          end_off = uniform((int)ring[(0 * NC + kM) * wcap + (hitk - lo0)]);
        }
      }
    } else if (status == 1) {
      int v0 = 0;
      if (tid == 0 && cbeg == kM) v0 = extend(0, 0);
      v0 = broadcast0(v0, 1);
      if (tid == 0) ring[ring_base(pen, wcap, cbeg, rp, 0)] = to_ring(v0, ring);
      if (cbeg == kM && cend == kM && kend == 0 && v0 >= tlen) status = kDevCompleted;
    }
    block_sync();
    // the heuristic also runs after the score-0 extension (wavefront_extend.c:112-118)
    if (CUT && status == 1 && cbeg == kM) {
      const int lo0 = a.endsfree ? -a.pbf : 0, hi0 = a.endsfree ? a.tbf : 0;
      if (cutoff(meta, steps_wait, xs, 0, lo0, lo0, hi0, 0, plen, tlen)) status = kDevZdropped;
    }

    int nulls = 0;  // null steps in a row, every score counted (align_status.num_null_steps)
    while (status == 1) {
      s += g;
      if constexpr (CUT) {
        // The scores between two multiples of g hold no wavefront: each is a
        // null step of the reference, whose extension ends the alignment as
        // unreachable once the run exceeds the score scope
        // (wavefront_extend.c:94-101) unless the step limit comes first.
        if (nulls + g - 1 > pen.scope) {
          s = s - g + pen.scope + 1 - nulls;
          status = (s >= a.max_steps) ? kDevMaxSteps : kDevUnreachable;
          break;
        }
        nulls += g - 1;
      }
      if (s >= a.max_steps) { status = kDevMaxSteps; break; }
      cur = (cur + 1 == R) ? 0 : cur + 1;
      rp.advance(pen);
      auto src = [&](int s_src, int back, int c) {
        return get_src_at(meta, s_src, cur - back, R, c, ring_base(pen, wcap, c, rp, back));
      };
      Src mx{1, -1, kFarDiagonal, 0, 0}, i1 = mx, d1 = mx, mo2 = mx, i2 = mx, d2 = mx;
      if (NC > 1 || pen.metric != kIndel) mx = src(s - pen.x, dq_x, kM);
      const Src mo1 = src(s - pen.o1 - pen.e1, dq_o1, kM);
      if constexpr (NC > 1) {
        i1 = src(s - pen.e1, dq_e1, kI1);
        d1 = src(s - pen.e1, dq_e1, kD1);
      }
      if constexpr (NC == 5) {
        mo2 = src(s - pen.o2 - pen.e2, dq_o2, kM);
        i2  = src(s - pen.e2, dq_e2, kI2);
        d2  = src(s - pen.e2, dq_e2, kD2);
      }
      const bool any = mx.valid() || mo1.valid() || i1.valid() || d1.valid() ||
                       mo2.valid() || i2.valid() || d2.valid();
      // Indel and edit always extend the previous wavefront, which only a
      // cut-off can empty: every later one is then empty as well
      // (wavefront_compute_edit.c:344-352 spans the NULLed ends) and the
      // reference computes on to the step limit.
      if (NC == 1 && pen.metric <= kEdit && !mo1.valid()) { status = CUT ? kDevMaxSteps : kDevBadTrace; break; }
      if (!any) {  // null step (wavefront_compute_affine.c:236-243, wavefront_compute_linear.c:170-176)
        if constexpr (CUT) {
          if (++nulls > pen.scope) { status = kDevUnreachable; break; }
        }
        if (s >= s_cap) { status = kDevBadTrace; break; }
        if (lane == cur) { meta.tag = s; meta.flags = 0u; }
        continue;
      }
      if constexpr (CUT) nulls = 0;
      // wavefront_compute_limits_input (wavefront_compute.c:40-86)
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
      const int width = hi - lo + 1;
      if (width > wcap) { status = kDevRingOverflow; break; }
      dsteps += width;
      // Edit distance may prune this wavefront: collect the least worst-case
      // score over its valid pre-extension cells.
      const bool prune_try = NC == 1 && pen.metric == kEdit && !a.endsfree && width >= kEditPruneWidth;
      int wmin = INT_MAX;
      // outputs present (wavefront_compute.c:446-493)
      const bool has_i1 = (NC > 1) && (mo1.valid() || i1.valid());
      const bool has_d1 = (NC > 1) && (mo1.valid() || d1.valid());
      const bool has_i2 = (NC == 5) && (mo2.valid() || i2.valid());
      const bool has_d2 = (NC == 5) && (mo2.valid() || d2.valid());

      uint8_t* rec = nullptr;
      if (a.records) {
        const unsigned long long units = rec_units_for(width);
        if (chunk_cur + units > chunk_end) {
          const unsigned long long want = units > kChunkUnits ? units : kChunkUnits;
          int base_lo = 0, base_hi = 0;
          if (tid == 0) {
            const unsigned long long b0 = atomicAdd(a.rec_top, want);
            base_lo = (int)(unsigned)(b0 & 0xFFFFFFFFull);
            base_hi = (int)(unsigned)(b0 >> 32);
          }
          base_lo = broadcast0(base_lo, 0);
          base_hi = broadcast0(base_hi, 1);
          chunk_cur = ((unsigned long long)(unsigned)base_hi << 32) | (unsigned)base_lo;
          chunk_end = chunk_cur + want;
          if (chunk_end > a.rec_units) {
            chunk_cur = chunk_end = 0;
            status = kDevArenaFull;
            break;
          }
        }
        const uint32_t blk = (uint32_t)chunk_cur;
        chunk_cur += units;
        if (tid == 0) a.rec_arena[blk] = make_uint4((unsigned)s, (unsigned)lo, (unsigned)hi, prev_blk);
        prev_blk = blk;
        rec = reinterpret_cast<uint8_t*>(a.rec_arena + blk + 1);
      }

      RT* outM  = ring + ring_base(pen, wcap, kM, rp, 0) - lo;
      RT* outI1 = ring + ring_base(pen, wcap, NC > 1 ? kI1 : kM, rp, 0) - lo;
      RT* outD1 = ring + ring_base(pen, wcap, NC > 1 ? kD1 : kM, rp, 0) - lo;
      RT* outI2 = ring + ring_base(pen, wcap, NC == 5 ? kI2 : kM, rp, 0) - lo;
      RT* outD2 = ring + ring_base(pen, wcap, NC == 5 ? kD2 : kM, rp, 0) - lo;

      WaveRange rM, rI1, rD1, rI2, rD2;
      rM.init(); rI1.init(); rD1.init(); rI2.init(); rD2.init();
      int hitk = INT_MAX;  // ends-free: first diagonal that ends the alignment
      // The ring reads of a chunk of diagonals (NC > 1), issued one chunk
      // ahead: they read the slots of earlier steps, never the one this step
      // writes, so their latency overlaps the current chunk's compute,
      // extension and stores.
      struct Reads {
        int om_l, om_r, ei1, ed1, mx, o2l, o2r, ei2, ed2;
      };
      auto reads = [&](int k) {
        Reads r{kOffsetNull, kOffsetNull, kOffsetNull, kOffsetNull, kOffsetNull,
                kOffsetNull, kOffsetNull, kOffsetNull, kOffsetNull};
        if (NC > 1 && k <= hi) {
          r.om_l = rd(ring, mo1, k - 1);
          r.om_r = rd(ring, mo1, k + 1);
          r.ei1 = rd(ring, i1, k - 1);
          r.ed1 = rd(ring, d1, k + 1);
          r.mx = rd(ring, mx, k);
          if constexpr (NC == 5) {
            r.o2l = rd(ring, mo2, k - 1);
            r.o2r = rd(ring, mo2, k + 1);
            r.ei2 = rd(ring, i2, k - 1);
            r.ed2 = rd(ring, d2, k + 1);
          }
        }
        return r;
      };
      Reads next = reads(lo + tid);
      for (int kb = lo; kb <= hi; kb += BLOCK) {
        const int k = kb + tid;
        const Reads cur_r = next;
        if constexpr (NC > 1) next = reads(k + BLOCK);
        bool vM = false, vI1 = false, vD1 = false, vI2 = false, vD2 = false;
        bool hit = false;
        if constexpr (NC == 1) {
          if (k <= hi) {
            // wavefront_compute_{indel,edit}_idm (wavefront_compute_edit.c:44-102)
            // and wavefront_compute_linear_idm (wavefront_compute_linear.c:44-76);
            // argmax with ties to the higher backtrace type: X > D > I
            // (wavefront_backtrace.c:223-319)
            const int del = rd(ring, mo1, k + 1);
            const int ins = rd(ring, mo1, k - 1) + 1;
            int m = rd(ring, mx, k) + 1;  // NULL + 1 for indel (mx absent)
            uint32_t src = kSrcX;
            if (del > m) { m = del; src = kSrcD1; }
            if (ins > m) { m = ins; src = kSrcI1; }
            if (!in_bounds(m, k, plen, tlen)) m = kOffsetNull;
            if (m != kOffsetNull) {
              if (prune_try) wmin = min(wmin, max(plen - (m - k), tlen - m));
              m = extend(k, m);
              hit = a.endsfree && endsfree_done(m, k, plen, tlen, a.pef, a.tef);
            }
            if (rec) rec[k - lo] = (uint8_t)src;
            outM[k] = to_ring(m, ring);
            vM = (m != kOffsetNull);
            if (k == kend) sh_end[nb & 1] = (int)to_ring(m, ring);
          }
        } else if (k <= hi) {
          const int om_l = cur_r.om_l;
          const int om_r = cur_r.om_r;
          const int ei1 = cur_r.ei1;
          const int ed1 = cur_r.ed1;
          const int ins1 = max(om_l, ei1) + 1;
          const int del1 = max(om_r, ed1);
          int best = cur_r.mx + 1;
          uint32_t src = kSrcX;
          uint32_t bits = (ei1 >= om_l ? kRecI1Ext : 0u) | (ed1 >= om_r ? kRecD1Ext : 0u);
          int ins2 = kOffsetNull, del2 = kOffsetNull;
          if constexpr (NC == 5) {
            const int o2l = cur_r.o2l;
            const int o2r = cur_r.o2r;
            const int ei2 = cur_r.ei2;
            const int ed2 = cur_r.ed2;
            ins2 = max(o2l, ei2) + 1;
            del2 = max(o2r, ed2);
            bits |= (ei2 >= o2l ? kRecI2Ext : 0u) | (ed2 >= o2r ? kRecD2Ext : 0u);
            if (del2 > best) { best = del2; src = kSrcD2; }
          }
          // argmax with ties to the higher backtrace type: M > D2 > D1 > I2 > I1
          if (del1 > best) { best = del1; src = kSrcD1; }
          if constexpr (NC == 5) { if (ins2 > best) { best = ins2; src = kSrcI2; } }
          if (ins1 > best) { best = ins1; src = kSrcI1; }
          int m = best;
          if (!in_bounds(m, k, plen, tlen)) m = kOffsetNull;  // null-after-max
          if (m != kOffsetNull) {
            m = extend(k, m);
            hit = a.endsfree && endsfree_done(m, k, plen, tlen, a.pef, a.tef);
          }
          if (rec) rec[k - lo] = (uint8_t)(src | bits);
          outM[k] = to_ring(m, ring);
          vM = (m != kOffsetNull);
          if (has_i1) { outI1[k] = to_ring(ins1, ring); vI1 = in_bounds(ins1, k, plen, tlen); }
          if (has_d1) { outD1[k] = to_ring(del1, ring); vD1 = in_bounds(del1, k, plen, tlen); }
          if constexpr (NC == 5) {
            if (has_i2) { outI2[k] = to_ring(ins2, ring); vI2 = in_bounds(ins2, k, plen, tlen); }
            if (has_d2) { outD2[k] = to_ring(del2, ring); vD2 = in_bounds(del2, k, plen, tlen); }
          }
          if (k == kend) {
            // the end component's value, for the termination test after the step
            const int ve = cend == kM ? m : cend == kI1 ? ins1 : cend == kD1 ? del1 : cend == kI2 ? ins2 : del2;
            sh_end[nb & 1] = (int)to_ring(ve, ring);
          }
        }
        const int kw = kb + wid * 64;
        rM.add(__ballot(vM), kw);
        if (a.endsfree) {
          const unsigned long long bal = __ballot(hit);
          if (bal && hitk == INT_MAX) hitk = kw + (int)__builtin_ctzll(bal);
        }
        if constexpr (NC > 1) {
          rI1.add(__ballot(vI1), kw);
          rD1.add(__ballot(vD1), kw);
        }
        if constexpr (NC == 5) {
          rI2.add(__ballot(vI2), kw);
          rD2.add(__ballot(vD2), kw);
        }
      }
      if (prune_try) {
        for (int o = 32; o > 0; o >>= 1) wmin = min(wmin, __shfl_xor(wmin, o, 64));
        wmin = uniform(wmin);
      }
      if constexpr (NW > 1) {
        if (lane == 0) {
          sh_red[nb & 1][wid][0] = rM.first;  sh_red[nb & 1][wid][1] = rM.last;
          if constexpr (NC == 1) sh_red[nb & 1][wid][2] = wmin;
          if constexpr (NC > 1) {
            sh_red[nb & 1][wid][2] = rI1.first; sh_red[nb & 1][wid][3] = rI1.last;
            sh_red[nb & 1][wid][4] = rD1.first; sh_red[nb & 1][wid][5] = rD1.last;
          }
          if constexpr (NC == 5) {
            sh_red[nb & 1][wid][6] = rI2.first; sh_red[nb & 1][wid][7] = rI2.last;
            sh_red[nb & 1][wid][8] = rD2.first; sh_red[nb & 1][wid][9] = rD2.last;
          }
        }
        __syncthreads();
        rM.init(); rI1.init(); rD1.init(); rI2.init(); rD2.init();
#pragma unroll
        for (int w = 0; w < NW; ++w) {
          rM.merge(sh_red[nb & 1][w][0], sh_red[nb & 1][w][1]);
          if constexpr (NC == 1) wmin = min(wmin, sh_red[nb & 1][w][2]);
          if constexpr (NC > 1) {
            rI1.merge(sh_red[nb & 1][w][2], sh_red[nb & 1][w][3]);
            rD1.merge(sh_red[nb & 1][w][4], sh_red[nb & 1][w][5]);
          }
          if constexpr (NC == 5) {
            rI2.merge(sh_red[nb & 1][w][6], sh_red[nb & 1][w][7]);
            rD2.merge(sh_red[nb & 1][w][8], sh_red[nb & 1][w][9]);
          }
        }
      }
      // (no second barrier: the one above already orders this step's ring
      // writes before every read of the next step, and the next step's
      // range writes go to the other half of sh_red)
      if constexpr (NW == 1) block_sync();
      if constexpr (CUT) {
        // The reference's backtrace also probes the readable cells of an
        // input the cut-off emptied (a null wavefront to compute,
        // wavefront_backtrace.c:64-219).  Where one exists, the record bytes
        // take the choices over the probed sources (probed_record); each lane
        // rewrites the bytes it wrote above.
        if (rec) probed_records<NC, BLOCK>(ring, meta, s, rp, pen, wcap, lo, hi, rec);
      }
      if constexpr (NC == 1) {
        if (prune_try && !rM.null()) edit_prune(rM, ring, mo1, wmin, plen, tlen);
      }
      // Record the step in the slot registers (every wave keeps a copy).
      if (lane == cur) {
        meta.tag = s;
        meta.alo = lo;
        unsigned fl = 1u | (rM.null() ? 0x100u : 0u);
        meta.lo[kM] = rM.tlo(lo);
        meta.hi[kM] = rM.thi(lo);
        if constexpr (NC > 1) {
          if (has_i1) { fl |= 2u | (rI1.null() ? 0x200u : 0u); meta.lo[kI1] = rI1.tlo(lo); meta.hi[kI1] = rI1.thi(lo); }
          if (has_d1) { fl |= 4u | (rD1.null() ? 0x400u : 0u); meta.lo[kD1] = rD1.tlo(lo); meta.hi[kD1] = rD1.thi(lo); }
        }
        if constexpr (NC == 5) {
          if (has_i2) { fl |= 8u | (rI2.null() ? 0x800u : 0u); meta.lo[kI2] = rI2.tlo(lo); meta.hi[kI2] = rI2.thi(lo); }
          if (has_d2) { fl |= 16u | (rD2.null() ? 0x1000u : 0u); meta.lo[kD2] = rD2.tlo(lo); meta.hi[kD2] = rD2.thi(lo); }
        }
        if constexpr (CUT) {
#pragma unroll
          for (int c = 0; c < NC; ++c) { meta.cut.lo[c] = meta.lo[c]; meta.cut.hi[c] = meta.hi[c]; }
        }
        meta.flags = fl;
      }
      if (a.endsfree) {
        hitk = red.min_all(hitk);
        if (hitk != INT_MAX) {
          status = kDevCompleted;
          end_k = hitk;
          end_off = uniform((int)ring[ring_base(pen, wcap, kM, rp, 0) + (hitk - lo)]);
        }
      } else {
        // Termination (wavefront_termination.c:37-114): k_end inside the trimmed
        // range of the end component and its offset reaching the text end.
        const int ce = NC == 1 ? (int)kM : cend;
        bool pce = true;
        WaveRange rce = rM;
        if (ce == kI1) { pce = has_i1; rce = rI1; }
        if (ce == kD1) { pce = has_d1; rce = rD1; }
        if constexpr (NC == 5) {
          if (ce == kI2) { pce = has_i2; rce = rI2; }
          if (ce == kD2) { pce = has_d2; rce = rD2; }
        }
        if (pce && !rce.null() && kend >= rce.first && kend <= rce.last) {
          // The lane of k_end left the value in sh_end (by the parity of
          // the computed steps, so a wave already in the next one cannot
          // overwrite it before every wave read it: one block barrier per
          // step): no ring round trip.  The same LDS word for every lane:
          // keep the loop control scalar.
          const int vend = uniform(sh_end[nb & 1]);
          if (vend >= tlen) status = kDevCompleted;
        }
      }
      if (CUT && status == 1 && !rM.null() && cutoff(meta, steps_wait, xs, cur, lo, rM.first, rM.last, s, plen, tlen)) {
        status = kDevZdropped;
        const int er = end_record<NC, HC>(ring, meta, s, rp, pen, wcap, xs.k);
        if (tid == 0 && a.out_end_rec) a.out_end_rec[p] = er;
      }
      if (status == 1 && s >= s_cap) status = kDevBadTrace;  // unreachable for valid input
      ++nb;
    }
    if (CUT == 2 && status == kDevZdropped) {  // alignment_end_pos of the z-drop (wavefront_heuristic.c:424-431)
      end_k = xs.k;
      end_off = xs.off;
    }
    if (tid == 0) {
      a.out_status[p] = (status == 1) ? kDevCompleted : status;
      a.out_score[p] = s;
      if (a.out_end_k) { a.out_end_k[p] = end_k; a.out_end_off[p] = end_off; }
      a.out_last_block[p] = prev_blk;
      a.out_dsteps[p] = dsteps;
      if (status == kDevRingOverflow || status == kDevArenaFull) atomicAdd(a.retry_count, 1u);
      else if (status != 1 && status != kDevCompleted) atomicAdd(a.fail_count, 1u);
    }
    if (status == 1 || status == kDevCompleted || status == kDevMaxSteps || status == kDevUnreachable ||
        status == kDevZdropped)
      wg_dsteps += (unsigned long long)dsteps;
  }
  if (tid == 0 && wg_dsteps) atomicAdd(a.dsteps_total, wg_dsteps);
}

// Sixteen bytes of a CIGAR slot gathered in registers before one aligned
// 16-byte store (the backtrace is bound by memory transactions: one lane per
// pair, so every store of a wave touches 64 different lines).
struct OpChunk {
  uint32_t a0 = 0u, a1 = 0u, a2 = 0u, a3 = 0u;
  // Byte b (0..15).  The halves are combined before the select: a select
  // between two members would become a load from a selected address, which
  // keeps the chunk in memory instead of registers.
  __device__ __forceinline__ unsigned byte_at(unsigned b) const {
    const uint64_t lo = ((uint64_t)a1 << 32) | a0, hi = ((uint64_t)a3 << 32) | a2;
    return (unsigned)(((b & 8u) ? hi : lo) >> (8u * (b & 7u))) & 0xFFu;
  }
  __device__ __forceinline__ static uint32_t span(unsigned from, unsigned to, unsigned j) {
    const unsigned lo = from < 4 * j ? 0u : from > 4 * j + 4 ? 4u : from - 4 * j;
    const unsigned hi = to < 4 * j ? 0u : to > 4 * j + 4 ? 4u : to - 4 * j;
    return (uint32_t)((1ull << (8 * hi)) - (1ull << (8 * lo)));
  }
  // Bytes [from, to) (0 <= from <= to <= 16), zero before, set to c4's byte
  // (c4 holds it four times).
  __device__ __forceinline__ void fill(unsigned from, unsigned to, uint32_t c4) {
    a0 |= c4 & span(from, to, 0);
    a1 |= c4 & span(from, to, 1);
    a2 |= c4 & span(from, to, 2);
    a3 |= c4 & span(from, to, 3);
  }
  __device__ __forceinline__ void store(unsigned char* at) const {
    *reinterpret_cast<uint4*>(at) = make_uint4(a0, a1, a2, a3);
  }
  __device__ __forceinline__ void clear() { a0 = a1 = a2 = a3 = 0u; }
};

__device__ __forceinline__ unsigned chunk_phase(const unsigned char* at) {
  return (unsigned)(reinterpret_cast<uintptr_t>(at) & 15u);
}

// Forward writer of the replay: consecutive bytes from position w0 of a
// slot are combined into aligned 16-byte stores (a run of matches fills
// whole chunks directly).  A chunk that reaches below w0 is written byte by
// byte, and so is the partial chunk at the end (finish), so no byte outside
// [w0, w) is touched: the neighbouring slots belong to other lanes.
// Positions below w may sit in the chunk until it fills; the replay only
// reads positions at or above w.
struct OpWriter {
  unsigned char* out;
  int w, w0;
  OpChunk c;  // the chunk holding position w: its bytes below w
  __device__ __forceinline__ unsigned phase() const { return chunk_phase(out + w); }
  // The chunk [w - 16, w) is complete (w is 16-aligned).
  __device__ __forceinline__ void flush() {
    if (w - 16 >= w0) {
      c.store(out + w - 16);
    } else {
      for (int i = w0; i < w; ++i) out[i] = (unsigned char)c.byte_at((unsigned)(i - (w - 16)));
    }
    c.clear();
  }
  __device__ __forceinline__ void put(unsigned op) {
    const unsigned d = phase();
    c.fill(d, d + 1, (op & 0xFFu) * 0x01010101u);
    ++w;
    if (d == 15u) flush();
  }
  // m copies of op: the rest of the current chunk, then whole chunks, then
  // the beginning of the last one.
  __device__ __forceinline__ void run(unsigned op, int m) {
    if (m <= 0) return;
    const uint32_t c4 = (op & 0xFFu) * 0x01010101u;
    const unsigned d = phase();
    if (d != 0u) {
      const unsigned take = (unsigned)m < 16u - d ? (unsigned)m : 16u - d;
      c.fill(d, d + take, c4);
      w += (int)take;
      m -= (int)take;
      if (d + take < 16u) return;
      flush();
    }
    for (; m >= 16; m -= 16, w += 16) *reinterpret_cast<uint4*>(out + w) = make_uint4(c4, c4, c4, c4);
    if (m > 0) {
      c.fill(0u, (unsigned)m, c4);
      w += m;
    }
  }
  __device__ __forceinline__ void finish() {
    const int d = (int)phase();
    for (int i = max(w - d, w0); i < w; ++i) out[i] = (unsigned char)c.byte_at((unsigned)(i - (w - d)));
    c.clear();
  }
};

// Backward writer of the walk: bytes written at descending positions below
// top are combined into aligned 16-byte stores.  A chunk that reaches top
// (the bytes from top up belong to the trailing gaps or to another slot) is
// written byte by byte, and so is the partial chunk at the end (finish).
struct OpWriterBack {
  unsigned char* out;
  int p, top;
  OpChunk c;  // the chunk holding position p: its bytes from p up
  __device__ __forceinline__ void put(unsigned op) {
    --p;
    const unsigned d = chunk_phase(out + p);
    c.fill(d, d + 1, (op & 0xFFu) * 0x01010101u);
    if (d == 0u) {
      if (p + 16 <= top) {
        c.store(out + p);
      } else {
        for (int i = p; i < top; ++i) out[i] = (unsigned char)c.byte_at((unsigned)(i - p));
      }
      c.clear();
    }
  }
  __device__ __forceinline__ void finish() {
    const int d = (int)chunk_phase(out + p);
    if (d != 0) {
      const int c0 = p - d;
      for (int i = p; i < min(c0 + 16, top); ++i) out[i] = (unsigned char)c.byte_at((unsigned)(i - c0));
    }
    c.clear();
  }
};

// One lane per pair, in two passes over the pair's CIGAR slot
// ops[0 .. plen+tlen):
//  1. Walk back from the end cell to the origin over (score, diagonal,
//     component) with the record bytes (wavefront_backtrace.c:320-529: the
//     recorded source is the maximum of (offset << 4 | type) the reference
//     takes).  Each mismatch, insertion or deletion is written backwards at
//     the end of the slot, its top bit set when the cell it leads into is an
//     M cell (an X, or the last operation of a gap that closes into M).
//  2. Replay the operations forward from the origin, extending every M cell
//     maximally (wavefront_extend_kernels.c:68-123, the same extension the
//     forward pass made there), and write the full CIGAR into the end of the
//     slot.  The count of matches is known before the replay (2M + 2X + I + D
//     = plen + tlen), so the CIGAR starts at plen + tlen - (M + X + I + D) and
//     its writes never pass the operations still to be read.
// Near the matrix borders only gaps can reach a cell (offsets of valid M cells
// are never negative), so walking to score 0 emits the same characters as the
// reference's shortcut at v == 0 or h == 0.
template <int BITS>
__global__ void __launch_bounds__(256) wfa_backtrace_kernel(BacktraceArgs a) {
  constexpr int SPW = 64 / BITS;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= a.n_pairs) return;
  const int p = a.order ? a.order[q] : q;
  const int plen = a.plen[p], tlen = a.tlen[p];
  const int cap = plen + tlen;
  unsigned char* out = reinterpret_cast<unsigned char*>(a.ops + a.ops_off[p]);
  // A z-dropped pair is traced from the best cell it saw, at the score it
  // dropped at (wavefront_unialign.c:175-200), like an ends-free end cell.
  const bool zd = a.status[p] == kDevZdropped;
  if (a.status[p] != kDevCompleted && !zd) { a.ops_begin[p] = cap; return; }
  const Penalties pen = a.pen;
  const int cbeg = a.comp_begin_p ? (int)a.comp_begin_p[p] : a.comp_begin;
  int comp = a.comp_end_p ? (int)a.comp_end_p[p] : a.comp_end;
  int s = a.score[p];
  int k = tlen - plen;
  int pos = cap;
  int nx = 0, nid = 0;
  bool bad = false, stale = false;
  // A z-drop's first step uses the end cell's own record byte (its diagonal
  // may lie outside the computed range of its score); with no valid source,
  // or on a border, the backtrace takes no step (wavefront_backtrace.c:
  // 265-266, 296-316): matches along the diagonal, then gaps to the origin.
  const int endrec = zd && a.end_rec ? a.end_rec[p] : 0;
  const int vE = zd ? a.end_off[p] - a.end_k[p] : 0, hE = zd ? a.end_off[p] : 0;
  const bool direct = zd && (endrec == 0xFF || vE <= 0 || hE <= 0 || s <= 0);
  int ntrail_all = 0;
  if (a.endsfree || zd) {
    // Trailing gaps after the end cell (wavefront_backtrace.c:345-355): the
    // pattern's rest as deletions, then the text's rest as insertions.  A
    // z-dropped CIGAR that takes steps keeps none of them: max-trim never
    // keeps a gap at the end, and the taken-back ones are among them.
    k = a.end_k[p];
    const int off = a.end_off[p];
    const int nd = plen - (off - k), ni = tlen - off;
    ntrail_all = nd + ni;
    if (nd < 0 || ni < 0 || nd + ni > cap) {
      bad = true;
    } else if (!zd || direct) {
      for (int i = 0; i < nd; ++i) out[--pos] = 'D';
      for (int i = 0; i < ni; ++i) out[--pos] = 'I';
    }
  }
  const int ntrail = cap - pos;
  bool first = zd;
  if (direct && !bad) {
    const int m = (vE > 0 && hE > 0) ? min(vE, hE) : 0;
    if (vE < 0 || hE < 0 || vE + hE - m > pos) {
      bad = true;
    } else {
      for (int i = 0; i < m; ++i) out[--pos] = 'M';
      for (int i = 0; i < vE - m; ++i) out[--pos] = 'D';
      for (int i = 0; i < hE - m; ++i) out[--pos] = 'I';
    }
  }
  bool mark = comp == kM;  // the next operation written leads into an M cell
  uint32_t blk = a.last_block[p];
  uint4 hdr = make_uint4(0, 0, 0, kNoBlock);
  if (blk != kNoBlock && blk < a.rec_units) hdr = a.rec_arena[blk];
  OpWriterBack bw{out, pos, pos, OpChunk{}};
  while (s > 0 && !bad && !direct) {
    if (bw.p <= 0) {
      bad = true;
      break;
    }
    uint32_t rec;
    if (first) {
      first = false;
      rec = (uint32_t)endrec;
    } else {
      while (blk != kNoBlock && blk < a.rec_units && (int)hdr.x > s) {
        blk = hdr.w;
        if (blk != kNoBlock && blk < a.rec_units) hdr = a.rec_arena[blk];
      }
      if (blk == kNoBlock || blk >= a.rec_units || (int)hdr.x != s || k < (int)hdr.y || k > (int)hdr.z) {
        bad = true;
        break;
      }
      rec = reinterpret_cast<const uint8_t*>(a.rec_arena + blk + 1)[k - (int)hdr.y];
    }
    if (rec & kRecStale) {
      stale = bad = true;
      break;
    }
    if (comp == kM) {
      switch (rec & kRecSrcMask) {
        case kSrcX:
          bw.put('X' | (mark ? 0x80u : 0u));
          ++nx;
          s -= pen.x;
          mark = true;
          continue;
        case kSrcD2: comp = kD2; break;  // closes a gap: same cell of D2
        case kSrcD1: comp = kD1; break;
        case kSrcI2: comp = kI2; break;
        default: comp = kI1; break;
      }
    }
    // gap cell (s, k) of comp: extension from the same gap or opening from M
    const bool ins = comp == kI1 || comp == kI2;
    const bool one = comp == kI1 || comp == kD1;
    const uint32_t ext_bit = comp == kI1 ? kRecI1Ext : comp == kD1 ? kRecD1Ext : comp == kI2 ? kRecI2Ext : kRecD2Ext;
    bw.put((ins ? 'I' : 'D') | (mark ? 0x80u : 0u));
    ++nid;
    k += ins ? -1 : 1;
    if (rec & ext_bit) {
      s -= one ? pen.e1 : pen.e2;
      mark = false;
    } else {
      s -= one ? pen.o1 + pen.e1 : pen.o2 + pen.e2;
      comp = kM;
      mark = true;
    }
  }
  bw.finish();
  pos = bw.p;
  // The walk ends at score 0: at the origin, or (ends-free) at a free
  // beginning cell (v=0, h=k) or (v=-k, h=0) of the score-0 wavefront.
  if (!bad && !direct && (s != 0 || comp != cbeg || (k != 0 && !a.endsfree))) bad = true;
  const int nlead = k < 0 ? -k : k;
  int start = direct && !bad ? pos : cap;
  const long long ppos = a.pat_pos[p], tpos = a.txt_pos[p];
  const uint64_t* P = a.words + ppos / SPW;
  const uint64_t* T = a.words + tpos / SPW;
  const int pb = (int)(ppos % SPW), tb = (int)(tpos % SPW);
  const int last_walk = cap - ntrail - 1;  // the operation leading into the end cell
  // A z-drop's end cell: the reference writes the matches from the cell's
  // computed offset up to the best cell's offset whatever the bases
  // (wavefront_backtrace.c:77-98); a best offset behind the computed one
  // takes back that many of the trailing gaps written before it.  A dry run
  // of the replay finds the computed offset and the matches before it.
  int zrun = 0, zback = 0, zmatches = 0;
  if (!bad && zd && !direct) {
    int v = k < 0 ? -k : 0, h = k > 0 ? k : 0, q = h;
    auto dry = [&]() {
      const int m = lcp_extend<BITS>(P, pb, T, tb, plen, tlen, h - v, h) - h;
      zmatches += m;
      v += m;
      h += m;
    };
    if (cbeg == kM) {
      if (pos == cap - ntrail) q = h; else dry();
    }
    for (int i = pos; i < cap - ntrail; ++i) {
      const unsigned char b = out[i];
      const unsigned char op = b & 0x7F;
      if (op == 'X') { ++v; ++h; }
      else if (op == 'I') ++h;
      else ++v;
      if (b & 0x80) {
        if (i == last_walk) q = h; else dry();
      }
    }
    const int n = a.end_off[p] - q;
    zrun = n > 0 ? n : 0;
    zback = n < 0 ? -n : 0;
    if (zback > ntrail_all) stale = bad = true;  // would take back past the trailing gaps
  }
  if (!bad && !direct) {
    // 2M + 2X + I + D = plen + tlen: the match count follows, and with it
    // the CIGAR's length and start (rest >= 0 keeps start in [0, pos - nlead]).
    // A z-drop's CIGAR is counted from the dry run instead.
    int len;
    if (zd) {
      len = zmatches + zrun + (cap - pos) + nlead;
      if (len > cap) bad = true;
    } else {
      const int rest = cap - 2 * nx - nid - ntrail - nlead;
      if (rest < 0 || (rest & 1)) bad = true;
      len = bad ? 0 : rest / 2 + (cap - pos) + nlead;
    }
    start = bad ? cap : cap - len;
    OpWriter ow{out, start, start, OpChunk{}};
    int v = 0, h = 0;
    if (!bad) {
      ow.run(k > 0 ? 'I' : 'D', nlead);  // free leading gaps
      if (k > 0) h = k; else v = -k;
    }
    auto extend = [&]() {
      const int h2 = lcp_extend<BITS>(P, pb, T, tb, plen, tlen, h - v, h);
      const int m = h2 - h;
      if (ow.w + m > cap) { bad = true; return; }
      ow.run('M', m);
      v += m;
      h += m;
    };
    auto run_to_end = [&]() {
      if (ow.w + zrun > cap) { bad = true; return; }
      ow.run('M', zrun);
      v += zrun;
      h += zrun;
    };
    if (!bad && cbeg == kM) {
      if (zd && pos == cap - ntrail) run_to_end(); else extend();
    }
    // The operations are read 16 bytes at a time: an aligned 16-byte chunk
    // that holds a byte of the slot lies in the same page as that byte.
    // Bytes of the cached chunk below the position read may since have been
    // overwritten by the writer (which stays below it); they are not read
    // again.
    uintptr_t cached = ~(uintptr_t)0;
    uint4 chunk = make_uint4(0u, 0u, 0u, 0u);
    for (int i = pos; i < cap && !bad; ++i) {
      const uintptr_t ad = reinterpret_cast<uintptr_t>(out + i);
      const unsigned o = (unsigned)(ad & 15u);
      if ((ad & ~(uintptr_t)15) != cached) {
        cached = ad & ~(uintptr_t)15;
        chunk = *reinterpret_cast<const uint4*>(out + (i - (int)o));
      }
      const uint32_t lo = (o & 4u) ? chunk.y : chunk.x, hi = (o & 4u) ? chunk.w : chunk.z;
      const unsigned char b = (unsigned char)(((o & 8u) ? hi : lo) >> (8u * (o & 3u)));
      const unsigned char op = b & 0x7F;
      if (ow.w > i) { bad = true; break; }  // inconsistent records: never overrun unread operations
      ow.put(op);
      if (op == 'X') { ++v; ++h; }
      else if (op == 'I') ++h;
      else ++v;
      if (b & 0x80) {
        if (zd && i == last_walk) run_to_end(); else extend();
      }
    }
    ow.finish();
    // the end: (plen, tlen), or a z-drop's end cell moved by what it takes back
    if (!bad && (v != (zd ? vE + zback : plen) || h != (zd ? hE + zback : tlen) || ow.w != cap)) bad = true;
  }
  if (stale) {
    a.status[p] = kDevStaleTrace;
    start = cap;
    atomicAdd(a.fail_count, 1u);
  } else if (bad) {
    a.status[p] = kDevBadTrace;
    start = cap;
    atomicAdd(a.fail_count, 1u);
  }
  a.ops_begin[p] = start;
}

__global__ void wfa_reset_kernel(Counters* c, int all) {
  if (threadIdx.x == 0) {
    c->queue = 0u;
    c->rec_top = 0ull;
    if (all >= 1) c->retry = 0u;
    if (all >= 2) { c->dsteps = 0ull; c->fail = 0u; }
  }
}

template <int NC, int BLOCK, bool SMALL, int BITS, int CUT>
const void* forward_fn() {
  return reinterpret_cast<const void*>(&wfa_forward_kernel<NC, BLOCK, SMALL, BITS, CUT>);
}

// Every instance the host can select.
const void* select_forward(int family, int ncomp, int bits, int block, int cut) {
#define WFA_SEL(NC, BL, SM, BI, CU) \
  if (ncomp == NC && block == BL && (family == kSmall) == SM && bits == BI && cut == CU) \
    return forward_fn<NC, BL, SM, BI, CU>();
#define WFA_SEL_NC(BL, SM, CU) \
  WFA_SEL(1, BL, SM, 2, CU) WFA_SEL(1, BL, SM, 8, CU) WFA_SEL(3, BL, SM, 2, CU) WFA_SEL(3, BL, SM, 8, CU) \
  WFA_SEL(5, BL, SM, 2, CU) WFA_SEL(5, BL, SM, 8, CU)
  WFA_SEL_NC(64, true, 0)
  WFA_SEL_NC(64, false, 0)
  WFA_SEL_NC(256, false, 0)
  WFA_SEL_NC(1024, false, 0)
  WFA_SEL_NC(64, true, 1)
  WFA_SEL_NC(64, false, 1)
  WFA_SEL_NC(256, false, 1)
  WFA_SEL_NC(1024, false, 1)
  WFA_SEL_NC(64, true, 2)
  WFA_SEL_NC(64, false, 2)
  WFA_SEL_NC(256, false, 2)
  WFA_SEL_NC(1024, false, 2)
#undef WFA_SEL_NC
#undef WFA_SEL
  return nullptr;
}

}  // namespace

size_t forward_lds_bytes(int family, int ring_slots, int wcap, int lmax_words) {
  size_t b = (size_t)2 * lmax_words * sizeof(uint64_t);  // sequences (HBM-ring family: 0 when they do not fit)
  if (family == kSmall) b += (size_t)ring_slots * wcap * sizeof(int16_t);
  return b;
}

// Threads per pair of the HBM-ring family by n + m: 64 up to t64, 256 up to
// t256, else 1024.  WFA_LARGE_BLOCKS="t64,t256" overrides the limits (tuning).
int pick_block(int family, int max_width) {
  static const int* lim = [] {
    static int t[2] = {256, 4096};
    const char* v = getenv("WFA_LARGE_BLOCKS");
    int a = 0, b = 0;
    if (v && sscanf(v, "%d,%d", &a, &b) == 2 && a >= 0 && b >= a) { t[0] = a; t[1] = b; }
    return t;
  }();
  if (family == kSmall) return 64;
  if (max_width <= lim[0]) return 64;
  if (max_width <= lim[1]) return 256;
  return 1024;
}

int forward_max_blocks_per_cu(int family, int block, int ncomp, int bits, int cut, size_t lds_bytes) {
  const void* f = select_forward(family, ncomp, bits, block, cut);
  if (!f) return 0;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, block, lds_bytes) != hipSuccess) return 0;
  return n;
}

hipError_t launch_forward(const ForwardArgs& a, int ncomp, int bits, const LaunchShape& shape,
                          hipStream_t stream) {
  const void* f = select_forward(shape.family, ncomp, bits, shape.block, a.heur.cut());
  if (!f) return hipErrorInvalidValue;
  if (shape.lds_bytes > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shape.lds_bytes);
    if (e != hipSuccess) return e;
  }
  void* args[] = {const_cast<ForwardArgs*>(&a)};
  return hipLaunchKernel(f, dim3(shape.grid), dim3(shape.block), args, shape.lds_bytes, stream);
}

hipError_t launch_backtrace(const BacktraceArgs& a, hipStream_t stream) {
  if (a.n_pairs <= 0) return hipSuccess;
  const int block = 256;
  const int grid = (a.n_pairs + block - 1) / block;
  if (a.bits == 2) hipLaunchKernelGGL(wfa_backtrace_kernel<2>, dim3(grid), dim3(block), 0, stream, a);
  else if (a.bits == 8) hipLaunchKernelGGL(wfa_backtrace_kernel<8>, dim3(grid), dim3(block), 0, stream, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_reset_counters(Counters* c, int all, hipStream_t stream) {
  hipLaunchKernelGGL(wfa_reset_kernel, dim3(1), dim3(64), 0, stream, c, all);
  return hipGetLastError();
}

}  // namespace wfa
