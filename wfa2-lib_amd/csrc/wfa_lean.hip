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
// wavefront, wavefront_compute.c:40-86) and its read window.
struct LeanSrc {
  int llo, lhi;   // limits contribution; valid iff llo <= lhi
  int clo, chi;   // reads clamp k into [clo, chi]
  int sb;         // ring element of diagonal 0
};

__device__ __forceinline__ LeanSrc lean_src(int org, int lo, int hi, int base) {
  LeanSrc r;
  const bool v = lo <= hi;
  r.llo = v ? lo : 1;
  r.lhi = v ? hi : -1;
  r.clo = v ? lo - 1 : 0;
  r.chi = v ? hi + 1 : 0;
  r.sb = v ? base - org : 0;  // element 0 of the ring is a NULL cell
  return r;
}

__device__ __forceinline__ int lean_rd(const int16_t* ring, const LeanSrc& s, int k) {
  return (int)ring[s.sb + min(max(k, s.clo), s.chi)];
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
    const int n = x ? (int)(__builtin_ctz(x) >> 1) : 16;
    if (n >= rem) return h + rem;
    h += n;
    v += n;
    if (n < 16) return h;
  }
}

template <int R, int DQX, int DQO, int DQE>
__global__ void __launch_bounds__(64) wfa_forward_lean_kernel(ForwardArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NC = 3;
  static_assert(DQX > 0 && DQO > 0 && DQE > 0 && DQX < R && DQO < R && DQE < R, "slot distances");
  const int lane = threadIdx.x;
  const Penalties pen = a.pen;
  const int g = pen.g;
  const int wcap = a.wcap;  // ring elements per slot and component (>= width + 2)

  const uint32_t* seqP = reinterpret_cast<const uint32_t*>(smem);
  const uint32_t* seqT = seqP + 2 * a.lmax_words;
  uint64_t* seqP64 = reinterpret_cast<uint64_t*>(smem);
  uint64_t* seqT64 = seqP64 + a.lmax_words;
  int16_t* ring = reinterpret_cast<int16_t*>(seqT64 + a.lmax_words);
  // element 0: NULL cell read by absent sources; slots start at element 2
  auto base_of = [&](int slot, int c) { return 2 + (slot * NC + c) * wcap; };
  if (lane == 0) ring[0] = kRingNull;

  unsigned long long chunk_cur = 0, chunk_end = 0;
  unsigned long long wg_dsteps = 0;

  for (;;) {
    int qv = 0;
    if (lane == 0) qv = (int)atomicAdd(a.queue, 1u);
    const int qi = lane_of(qv, 0);
    if (qi >= a.n_pairs) break;
    const int p = a.order ? a.order[qi] : qi;
    const int plen = a.plen[p];
    const int tlen = a.tlen[p];
    const int kend = tlen - plen;
    const int s_cap = (pen.x + pen.o1 + pen.e1 + 2) * (plen + tlen + 2);
    const long long ppos = a.pat_pos[p], tpos = a.txt_pos[p];
    const int pb = (int)(ppos & 31), tb = (int)(tpos & 31);
    {
      const uint64_t* gP = a.words + (ppos >> 5);
      const uint64_t* gT = a.words + (tpos >> 5);
      const int nwp = (pb + plen) / 32 + 2, nwt = (tb + tlen) / 32 + 2;
      wave_sync();  // the previous pair is done reading the sequences
      for (int i = lane; i < nwp; i += 64) seqP64[i] = gP[i];
      for (int i = lane; i < nwt; i += 64) seqT64[i] = gT[i];
      wave_sync();
    }

    // Ring metadata (uniform): origin and trimmed ranges per slot.
    int s_org[R], s_lo[R][NC], s_hi[R][NC];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      s_org[r] = 0;
#pragma unroll
      for (int c = 0; c < NC; ++c) { s_lo[r][c] = 1; s_hi[r][c] = -1; }
    }
    int s = 0;
    int status = 1;
    uint32_t prev_blk = kNoBlock;
    long long dsteps = 0;
    {
      // Score 0: M holds offset 0 at k = 0 (wavefront_aligner.c:252-380), extended.
      const int v0 = uniform(lcp16(seqP, pb, seqT, tb, plen, tlen, 0, 0));
      s_org[0] = -1;
      s_lo[0][kM] = 0;
      s_hi[0][kM] = 0;
      const int b = base_of(0, kM);
      if (lane < 3) ring[b + lane] = (lane == 1) ? to_ring(v0, ring) : kRingNull;
      if (kend == 0 && v0 >= tlen) status = kDevCompleted;
      wave_sync();
    }

    while (status == 1) {
#pragma unroll
      for (int j = 1; j <= R; ++j) {
        const int ph = j % R;  // slot of this step (compile-time after unrolling)
        s += g;
        if (s >= a.max_steps) { status = kDevMaxSteps; break; }
        const int rx = (ph - DQX + R) % R, ro = (ph - DQO + R) % R, re = (ph - DQE + R) % R;
        const LeanSrc mx = lean_src(s_org[rx], s_lo[rx][kM], s_hi[rx][kM], base_of(rx, kM));
        const LeanSrc mo = lean_src(s_org[ro], s_lo[ro][kM], s_hi[ro][kM], base_of(ro, kM));
        const LeanSrc i1 = lean_src(s_org[re], s_lo[re][kI1], s_hi[re][kI1], base_of(re, kI1));
        const LeanSrc d1 = lean_src(s_org[re], s_lo[re][kD1], s_hi[re][kD1], base_of(re, kD1));
        const bool vx = mx.llo <= mx.lhi, vo = mo.llo <= mo.lhi, vi = i1.llo <= i1.lhi, vd = d1.llo <= d1.lhi;
        if (!(vx || vo || vi || vd)) {  // null step (wavefront_compute_affine.c:236-243)
#pragma unroll
          for (int c = 0; c < NC; ++c) { s_lo[ph][c] = 1; s_hi[ph][c] = -1; }
          if (s >= s_cap) { status = kDevBadTrace; break; }
          continue;
        }
        // wavefront_compute_limits_input (wavefront_compute.c:40-86)
        const int lo = min(min(mx.llo, mo.llo - 1), min(i1.llo + 1, d1.llo - 1));
        const int hi = max(max(mx.lhi, mo.lhi + 1), max(i1.lhi + 1, d1.lhi - 1));
        const int width = hi - lo + 1;
        if (width + 2 > wcap) { status = kDevRingOverflow; break; }
        dsteps += width;
        const bool has_i1 = vo || vi;
        const bool has_d1 = vo || vd;

        uint8_t* rec = nullptr;
        if (a.records) {
          const unsigned long long units = rec_units_for(width);
          if (chunk_cur + units > chunk_end) {
            const unsigned long long want = units > kChunkUnits ? units : kChunkUnits;
            unsigned long long b0 = 0;
            if (lane == 0) b0 = atomicAdd(a.rec_top, want);
            const int blo = lane_of((int)(unsigned)(b0 & 0xFFFFFFFFull), 0);
            const int bhi = lane_of((int)(unsigned)(b0 >> 32), 0);
            chunk_cur = ((unsigned long long)(unsigned)bhi << 32) | (unsigned)blo;
            chunk_end = chunk_cur + want;
            if (chunk_end > a.rec_units) {
              chunk_cur = chunk_end = 0;
              status = kDevArenaFull;
              break;
            }
          }
          const uint32_t blk = (uint32_t)chunk_cur;
          chunk_cur += units;
          if (lane == 0) a.rec_arena[blk] = make_uint4((unsigned)s, (unsigned)lo, (unsigned)hi, prev_blk);
          prev_blk = blk;
          rec = reinterpret_cast<uint8_t*>(a.rec_arena + blk + 1);
        }

        // Outputs of this step go to slot ph, origin lo - 1.
        const int org = lo - 1;
        int16_t* oM = ring + base_of(ph, kM) - org;
        int16_t* oI = ring + base_of(ph, kI1) - org;
        int16_t* oD = ring + base_of(ph, kD1) - org;
        WaveRange rM, rI, rD;
        rM.init(); rI.init(); rD.init();
        bool fin = false;  // termination (wavefront_termination.c:37-114)
        for (int kb = lo; kb <= hi; kb += 64) {
          const int k = kb + lane;
          const bool act = k <= hi;
          // Reads are clamped, so every lane may read; only active lanes store.
          const int om_l = lean_rd(ring, mo, k - 1);
          const int om_r = lean_rd(ring, mo, k + 1);
          const int ei1 = lean_rd(ring, i1, k - 1);
          const int ed1 = lean_rd(ring, d1, k + 1);
          const int ins1 = max(om_l, ei1) + 1;
          const int del1 = max(om_r, ed1);
          int best = lean_rd(ring, mx, k) + 1;
          uint32_t src = kSrcX;
          // argmax with ties to the higher backtrace type: M > D1 > I1
          if (del1 > best) { best = del1; src = kSrcD1; }
          if (ins1 > best) { best = ins1; src = kSrcI1; }
          const uint32_t bits = (ei1 >= om_l ? kRecI1Ext : 0u) | (ed1 >= om_r ? kRecD1Ext : 0u);
          if (rec && act) rec[k - lo] = (uint8_t)(src | bits);
          // null-after-max, then the extension of the valid M cells: one
          // 16-symbol window for every lane, more only for lanes whose run
          // fills it (uniform branch).
          const bool okM = act && in_bounds(best, k, plen, tlen);
          int h = okM ? best : 0;
          int rem = okM ? min(plen - (best - k), tlen - best) : 0;
          int pp = okM ? pb + best - k : 0, pt = okM ? tb + best : 0;
          {
            const uint32_t x = win16(seqP, pp) ^ win16(seqT, pt);
            const int n = min(x ? (int)(__builtin_ctz(x) >> 1) : 16, rem);
            h += n;
            bool more = n == 16 && rem > 16;
            if (__ballot(more)) {
              while (more) {
                rem -= 16;
                pp += 16;
                pt += 16;
                const uint32_t y = win16(seqP, pp) ^ win16(seqT, pt);
                const int n2 = min(y ? (int)(__builtin_ctz(y) >> 1) : 16, rem);
                h += n2;
                more = n2 == 16 && rem > 16;
              }
            }
          }
          const bool vM = okM;
          const bool vI = act && has_i1 && in_bounds(ins1, k, plen, tlen);
          const bool vD = act && has_d1 && in_bounds(del1, k, plen, tlen);
          if (act) {
            oM[k] = to_ring(okM ? h : kOffsetNull, ring);
            if (has_i1) oI[k] = to_ring(ins1, ring);
            if (has_d1) oD[k] = to_ring(del1, ring);
          }
          fin = fin || __ballot(okM && k == kend && h >= tlen);
          rM.add(__ballot(vM), kb);
          rI.add(__ballot(vI), kb);
          rD.add(__ballot(vD), kb);
        }
        // Trimmed ranges (wavefront_compute.c:579-613); absent or empty = (1, -1).
        const int mlo = rM.null() ? 1 : rM.first, mhi = rM.null() ? -1 : rM.last;
        const int ilo = (!has_i1 || rI.null()) ? 1 : rI.first, ihi = (!has_i1 || rI.null()) ? -1 : rI.last;
        const int dlo = (!has_d1 || rD.null()) ? 1 : rD.first, dhi = (!has_d1 || rD.null()) ? -1 : rD.last;
        // NULL just outside each trimmed range: lanes 0..5 write one cell each.
        if (lane < 6) {
          const int c = lane >> 1;
          const int clo = c == 0 ? mlo : (c == 1 ? ilo : dlo);
          const int chi = c == 0 ? mhi : (c == 1 ? ihi : dhi);
          if (clo <= chi) {
            int16_t* o = c == 0 ? oM : (c == 1 ? oI : oD);
            o[(lane & 1) ? chi + 1 : clo - 1] = kRingNull;
          }
        }
        s_org[ph] = org;
        s_lo[ph][kM] = mlo; s_hi[ph][kM] = mhi;
        s_lo[ph][kI1] = ilo; s_hi[ph][kI1] = ihi;
        s_lo[ph][kD1] = dlo; s_hi[ph][kD1] = dhi;
        wave_sync();  // ring writes of this step before the next step's reads
        // M[k_end] reached the end of the text: a valid cell at k_end lies in
        // the trimmed range, so the lane that computed it decides.
        if (fin) { status = kDevCompleted; break; }
        if (s >= s_cap) { status = kDevBadTrace; break; }
      }
    }
    if (lane == 0) {
      a.out_status[p] = status;
      a.out_score[p] = s;
      a.out_last_block[p] = prev_blk;
      a.out_dsteps[p] = dsteps;
      if (status == kDevRingOverflow || status == kDevArenaFull) atomicAdd(a.retry_count, 1u);
      else if (status != kDevCompleted) atomicAdd(a.fail_count, 1u);
    }
    if (status == kDevCompleted || status == kDevMaxSteps) wg_dsteps += (unsigned long long)dsteps;
  }
  if (lane == 0 && wg_dsteps) atomicAdd(a.dsteps_total, wg_dsteps);
}

template <int R, int DQX, int DQO, int DQE>
const void* lean_fn() {
  return reinterpret_cast<const void*>(&wfa_forward_lean_kernel<R, DQX, DQO, DQE>);
}

// Slot structures with an instance: (R, dq_x, dq_o, dq_e).
const void* select_lean(const Penalties& pen) {
  if (pen.g <= 0) return nullptr;
  const int dx = pen.x / pen.g, dq = (pen.o1 + pen.e1) / pen.g, de = pen.e1 / pen.g, R = pen.nslots;
  // gap-affine (4, 6, 2) and every multiple of (2, 3, 1)
  if (R == 5 && dx == 2 && dq == 4 && de == 1) return lean_fn<5, 2, 4, 1>();
  return nullptr;
}

}  // namespace

bool lean_supported(const Penalties& pen) { return select_lean(pen) != nullptr; }

size_t lean_lds_bytes(const Penalties& pen, int wcap, int lmax_words) {
  return (size_t)2 * lmax_words * sizeof(uint64_t) + (size_t)(2 + pen.nslots * 3 * wcap) * sizeof(int16_t);
}

int lean_max_blocks_per_cu(const Penalties& pen, size_t lds_bytes) {
  const void* f = select_lean(pen);
  if (!f) return 0;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, 64, lds_bytes) != hipSuccess) return 0;
  return n;
}

hipError_t launch_forward_lean(const ForwardArgs& a, const LaunchShape& shape, hipStream_t stream) {
  const void* f = select_lean(a.pen);
  if (!f || shape.block != 64) return hipErrorInvalidValue;
  if (shape.lds_bytes > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shape.lds_bytes);
    if (e != hipSuccess) return e;
  }
  void* args[] = {const_cast<ForwardArgs*>(&a)};
  return hipLaunchKernel(f, dim3(shape.grid), dim3(64), args, shape.lds_bytes, stream);
}

}  // namespace wfa
