/*
 * CPU oracle for the exact WFA hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of WFA2-lib v2.3 (smarco/WFA2-lib) wavefront_align for
 * the indel, edit, gap-linear, gap-affine and gap-affine-2p metrics; end-to-end,
 * ends-free and extension forms; the heuristics none, wf-adaptive, wfmash,
 * x-drop, z-drop, banded static and banded adaptive; memory modes
 * high/med/low (one algorithm, identical CIGARs) and ultralow (BiWFA).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * as the checker; the product library never links or calls it.
 *
 * Pinning (tests/test_oracle.py): golden .alg files of the reference's own
 * regression suite (tests/wfa.utest.check/), the reference itself compiled from
 * source (oracle/_ref/), and an independent O(nm) DP scorer + CIGAR validity.
 *
 * Model.  A wavefront is (present, null, lo, hi, values over the computed range
 * [alo,ahi]).  "present" is the reference's non-NULL component pointer; "null"
 * is wavefront->null after trimming.  Every read outside [lo,hi] of a present,
 * non-null wavefront returns NULL: that is what wavefront_compute_init_ends
 * (wavefront_compute.c:498-575) guarantees for the compute kernels, and the
 * backtrace probes only the init range, which holds NULLs beyond [lo,hi]
 * (wavefront_backtrace.c:64-219).  Absent or null inputs are the shared null
 * wavefront, whose limits are lo=1, hi=-1 (wavefront.c:105-133).
 *
 * Extension compares bytes and stops at either sequence end.  The reference
 * stops there because of the EOS sentinels '!' (pattern) and '?' (text)
 * (wavefront_sequences.c:37-39); the two agree unless the text holds '!' or
 * the pattern holds '?' at the very position after the other sequence ends.
 */
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include "wfa_oracle.h"

#define M_INDEL 0   /* distance_metric_t, wavefront_penalties.h:41-47 */
#define M_EDIT 1
#define M_LINEAR 2
#define ONULL (INT32_MIN/2)                 /* WAVEFRONT_OFFSET_NULL, wavefront_offset.h:46 */

/* Status codes, wfa.h:45-55 */
#define ST_ALG_COMPLETED      0
#define ST_MAX_STEPS        -100
#define ST_UNATTAINABLE     -300
#define ST_OK                 -1
#define ST_END_REACHED        -2
#define ST_END_UNREACHABLE    -3
#define ST_ALG_PARTIAL         1

/* BiWFA constants, wavefront_bialign.c:52-54 */
#define BI_FALLBACK_MIN_SCORE  250
#define BI_FALLBACK_MIN_LENGTH 100
#define BI_RECOVERY_MIN_SCORE  500

enum { C_M = 0, C_I1 = 1, C_D1 = 2, C_I2 = 3, C_D2 = 4, NCOMP = 5 };

/* Backtrace type tags, wavefront_backtrace.c:49-59 */
enum { BT_I1O = 1, BT_I1E = 2, BT_I2O = 3, BT_I2E = 4, BT_D1O = 5, BT_D1E = 6,
       BT_D2O = 7, BT_D2E = 8, BT_M = 9 };

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
/* MAX as (a)>=(b)?a:b, utils/commons.h:195 (value-identical for int). */
static inline int32_t vmax(int32_t a, int32_t b) { return a >= b ? a : b; }
static inline int64_t lmax(int64_t a, int64_t b) { return a >= b ? a : b; }

typedef struct {
  int score;          /* score held by this slot, -1 if none */
  bool present;
  bool null;
  int lo, hi;         /* trimmed range (narrowed further by a heuristic cut-off) */
  int rlo, rhi;       /* readable range: the trimmed range before any cut-off
                         (wf_elements_init_min/max, wavefront_compute.c:579-613) */
  int alo, ahi;       /* computed range */
  int32_t* v;         /* values, v[k-alo] */
  int cap;
} owf_t;

typedef struct {
  /* internal penalties (wavefront_penalties.c:95-173) */
  int metric, match, x, o1, e1, o2, e2;
  int scope;          /* max_score_scope, wavefront_components.c:81-124 */
  bool modular;       /* ring of `scope` slots (score-only / BiWFA sides) */
  const char* P; int plen;
  const char* T; int tlen;
  int comp_begin, comp_end;
  owf_t* slots[NCOMP];
  int nslots;
  int max_steps;
  int64_t dsteps;
  int status;         /* ST_OK while running, ST_END_REACHED at the end */
  int null_steps;     /* consecutive compute steps with every input null
                         (align_status.num_null_steps, wavefront_compute_affine.c:236-244) */
  int end_score;
  /* alignment form (wavefront_attributes.h:63-72); extension sets the ends
     free up to both sequence ends (wavefront_align.c:57-63) */
  bool endsfree;
  int pbf, pef, tbf, tef;
  int end_k, end_off; /* end cell (wf_aligner->alignment_end_pos) */
  /* heuristics (wavefront_heuristic.c:229-290, 467-500) */
  bool wfadaptive, wfmash, xdrop, zdrop, banded_static, banded_adaptive;
  int h_min_len, h_max_thr, h_steps, steps_wait, band_min_k, band_max_k, h_xdrop, h_zdrop;
  /* x-drop / z-drop: best Smith-Waterman score seen, its diagonal (INT_MAX:
     none), offset and score */
  int max_sw, max_sw_k, max_sw_off, max_wf_score;
  bool stale_hit;     /* the backtrace met a cell whose choices change when
                         cut-away wavefronts are read (see cell_stale) */
  /* CIGAR, written backwards by the backtrace into ops[begin..end) */
  char* ops; int ops_cap; int ops_begin; int ops_end;
  int cigar_score;
} oaln_t;

/* ---------------------------------------------------------------- storage */

static void aln_free(oaln_t* a) {
  int c, i;
  for (c = 0; c < NCOMP; ++c) {
    if (!a->slots[c]) continue;
    for (i = 0; i < a->nslots; ++i) free(a->slots[c][i].v);
    free(a->slots[c]);
    a->slots[c] = NULL;
  }
  a->nslots = 0;
  free(a->ops);
  a->ops = NULL;
}

static void aln_reset_slots(oaln_t* a) {
  int c, i;
  for (c = 0; c < NCOMP; ++c)
    for (i = 0; i < a->nslots; ++i) { a->slots[c][i].score = -1; a->slots[c][i].present = false; }
}

static void aln_grow(oaln_t* a, int need) {
  if (need <= a->nslots) return;
  int n = imax(need, a->nslots * 2 + 16);
  int c, i;
  for (c = 0; c < NCOMP; ++c) {
    a->slots[c] = realloc(a->slots[c], sizeof(owf_t) * n);
    for (i = a->nslots; i < n; ++i) {
      memset(&a->slots[c][i], 0, sizeof(owf_t));
      a->slots[c][i].score = -1;
    }
  }
  a->nslots = n;
}

static inline int slot_of(const oaln_t* a, int s) { return a->modular ? s % a->scope : s; }

/* The wavefront stored for score s (present), or NULL. */
static owf_t* wf_at(oaln_t* a, int c, int s) {
  if (s < 0) return NULL;
  const int i = slot_of(a, s);
  if (i >= a->nslots) return NULL;
  owf_t* w = &a->slots[c][i];
  return (w->score == s && w->present) ? w : NULL;
}

/* Compute input: a present, non-null wavefront, else NULL (the null wavefront). */
static owf_t* wf_src(oaln_t* a, int c, int s) {
  owf_t* w = wf_at(a, c, s);
  return (w && !w->null) ? w : NULL;
}

static owf_t* wf_claim(oaln_t* a, int c, int s) {
  const int i = slot_of(a, s);
  aln_grow(a, i + 1);
  owf_t* w = &a->slots[c][i];
  w->score = s;
  w->present = false;
  return w;
}

static void wf_alloc(owf_t* w, int lo, int hi) {
  const int n = hi - lo + 1;
  if (w->cap < n) {
    free(w->v);
    w->cap = n + n / 2 + 8;
    w->v = malloc(sizeof(int32_t) * w->cap);
  }
  w->present = true;
  w->null = false;
  w->lo = w->alo = w->rlo = lo;
  w->hi = w->ahi = w->rhi = hi;
}

static inline int32_t wf_raw(const owf_t* w, int k) { return w->v[k - w->alo]; }

/* Compute-kernel read: NULL outside the readable range or for the null
   wavefront.  A heuristic cut-off narrows [lo,hi] but leaves the cells it
   cut readable: wavefront_compute_init_ends (wavefront_compute.c:498-575)
   fills NULLs only beyond the initialised range. */
static inline int32_t wf_rd(const owf_t* w, int k) {
  if (!w || k < w->rlo || k > w->rhi) return ONULL;
  return wf_raw(w, k);
}

/* ------------------------------------------------------------ penalties */

static void aln_set_penalties(oaln_t* a, const oracle_conf_t* conf) {
  a->metric = conf->metric;
  /* the strategies as wavefront_aligner_init_heuristic selects them
     (wavefront_aligner.c:189-228): wf-adaptive before wfmash, static band
     before adaptive band */
  a->wfadaptive = (conf->heuristic & 4) != 0;
  a->wfmash = !a->wfadaptive && (conf->heuristic & 0x40) != 0;
  a->xdrop = (conf->heuristic & 0x10) != 0;
  a->zdrop = !a->xdrop && (conf->heuristic & 0x20) != 0;
  a->h_xdrop = conf->xdrop;
  a->h_zdrop = conf->zdrop;
  a->banded_static = (conf->heuristic & 1) != 0;
  a->banded_adaptive = !a->banded_static && (conf->heuristic & 2) != 0;
  a->band_min_k = conf->min_k;
  a->band_max_k = conf->max_k;
  a->h_min_len = conf->min_wavefront_length;
  a->h_max_thr = conf->max_distance_threshold;
  a->h_steps = conf->steps_between_cutoffs;
  a->max_steps = (conf->max_steps > 0) ? conf->max_steps : INT_MAX;
  if (a->metric <= M_LINEAR) {
    /* wavefront_penalties_set_{indel,edit,linear}, wavefront_penalties.c:39-93;
       scopes from wavefront_components.c:45-80 */
    a->match = 0;
    a->e1 = a->o2 = a->e2 = -1;
    if (a->metric == M_INDEL) { a->x = -1; a->o1 = 1; a->scope = 2; return; }
    if (a->metric == M_EDIT) { a->x = 1; a->o1 = 1; a->scope = 2; return; }
    if (conf->match < 0) {  /* Eizenga's transform */
      a->match = conf->match;
      a->x = 2 * conf->mismatch - 2 * conf->match;
      a->o1 = 2 * conf->gap_opening1 - conf->match;
    } else {
      a->x = conf->mismatch;
      a->o1 = conf->gap_opening1;
    }
    a->scope = imax(a->x, a->o1) + 1;
    return;
  }
  if (conf->match < 0) { /* Eizenga's transform, wavefront_penalties.c:114-119,152-159 */
    a->match = conf->match;
    a->x = 2 * conf->mismatch - 2 * conf->match;
    a->o1 = 2 * conf->gap_opening1;
    a->e1 = 2 * conf->gap_extension1 - conf->match;
    a->o2 = 2 * conf->gap_opening2;
    a->e2 = 2 * conf->gap_extension2 - conf->match;
  } else {
    a->match = 0;
    a->x = conf->mismatch; a->o1 = conf->gap_opening1; a->e1 = conf->gap_extension1;
    a->o2 = conf->gap_opening2; a->e2 = conf->gap_extension2;
  }
  if (a->metric != 4) { a->o2 = -1; a->e2 = -1; }
  /* max_score_scope, wavefront_components.c:81-124 */
  int indel = a->o1 + a->e1;
  if (a->metric == 4) indel = imax(indel, a->o2 + a->e2);
  a->scope = imax(indel, a->x) + 1;
  a->max_steps = (conf->max_steps > 0) ? conf->max_steps : INT_MAX;
}

/* wavefront_compute_classic_score, wavefront_compute.c:108-120 */
static int classic_score(const oaln_t* a, int plen, int tlen, int s) {
  if (a->metric <= M_EDIT) return s;
  const int swg_match = -a->match;
  if (swg_match == 0) return -s;
  return (swg_match * (plen + tlen) - s) / 2;
}

/* --------------------------------------------------------------- compute */

static void trim_ends(const oaln_t* a, owf_t* w) {
  /* wavefront_compute_trim_ends, wavefront_compute.c:579-613 */
  int k;
  for (k = w->hi; k >= w->lo; --k) {
    const int32_t off = wf_raw(w, k);
    if ((uint32_t)off <= (uint32_t)a->tlen && (uint32_t)(off - k) <= (uint32_t)a->plen) break;
  }
  w->hi = k;
  for (k = w->lo; k <= w->hi; ++k) {
    const int32_t off = wf_raw(w, k);
    if ((uint32_t)off <= (uint32_t)a->tlen && (uint32_t)(off - k) <= (uint32_t)a->plen) break;
  }
  w->lo = k;
  w->null = (w->lo > w->hi);
  w->rlo = w->lo;
  w->rhi = w->hi;
}

#define WLO(w) ((w) ? (w)->lo : 1)
#define WHI(w) ((w) ? (w)->hi : -1)

/* wf_compute_edit_{best,worst}_score, wavefront_compute_edit.c:199-218
   (plain int arithmetic, as there: a NULL offset cancels out of best) */
static int edit_best(const oaln_t* a, int k, int32_t off) {
  const int left_v = a->plen - (off - k), left_h = a->tlen - off;
  return (left_v >= left_h) ? left_v - left_h : left_h - left_v;
}
static int edit_worst(const oaln_t* a, int k, int32_t off) {
  return imax(a->plen - (off - k), a->tlen - off);
}

/* wavefront_compute_edit_exact_prune, wavefront_compute_edit.c:219-278 */
static void edit_exact_prune(const oaln_t* a, owf_t* w) {
  const int lo = w->lo, hi = w->hi;
  if (hi - lo + 1 < 1000) return;
  const int sample_k = lo + (hi - lo) / 2;
  const int32_t sample = wf_raw(w, sample_k);
  if (sample < 0) return;
  const int smax_sample = edit_worst(a, sample_k, sample);
  const int smin_lo = edit_best(a, lo, wf_raw(w, lo));
  const int smin_hi = edit_best(a, hi, wf_raw(w, hi));
  if (smin_lo <= smax_sample && smin_hi <= smax_sample) return;
  int min_worst = INT_MAX, k;
  for (k = lo; k <= hi; ++k) {
    const int32_t off = wf_raw(w, k);
    if (off < 0) continue;
    const int sw = edit_worst(a, k, off);
    if (sw < min_worst) min_worst = sw;
  }
  int lo_r = lo;
  for (k = lo; k <= hi; ++k) {
    if (edit_best(a, k, wf_raw(w, k)) <= min_worst) break;
    ++lo_r;
  }
  w->lo = lo_r;
  int hi_r = hi;
  for (k = hi; k > lo_r; --k) {
    if (edit_best(a, k, wf_raw(w, k)) <= min_worst) break;
    --hi_r;
  }
  w->hi = hi_r;
  w->rlo = w->lo;  /* wavefront_compute_edit NULLs the two cells beyond each end (:355-358) */
  w->rhi = w->hi;
}

/* wavefront_compute_edit (compute_edit.c:329-373: indel and edit, from the
   previous wavefront only) and wavefront_compute_linear (compute_linear.c:
   166-195); kernels compute_edit.c:44-102, compute_linear.c:44-76 */
static void compute_linear(oaln_t* a, int s) {
  aln_grow(a, slot_of(a, s) + 1);
  owf_t* out = wf_claim(a, C_M, s);
  int lo, hi, k;
  if (a->metric <= M_EDIT) {
    owf_t* prev = wf_at(a, C_M, s - 1);
    if (!prev) return;
    lo = prev->lo - 1;
    hi = prev->hi + 1;
    wf_alloc(out, lo, hi);
    for (k = lo; k <= hi; ++k) {
      const int32_t ins = wf_rd(prev, k - 1), del = wf_rd(prev, k + 1);
      int32_t m = (a->metric == M_INDEL) ? vmax(del, ins + 1) : vmax(del, vmax(ins, wf_rd(prev, k)) + 1);
      if ((uint32_t)m > (uint32_t)a->tlen) m = ONULL;
      if ((uint32_t)(m - k) > (uint32_t)a->plen) m = ONULL;
      out->v[k - lo] = m;
    }
    a->dsteps += hi - lo + 1;
    trim_ends(a, out);
    if (a->metric == M_EDIT && !a->endsfree) edit_exact_prune(a, out);  /* end-to-end only, :370-373 */
    return;
  }
  owf_t* mx = wf_src(a, C_M, s - a->x);
  owf_t* mo = wf_src(a, C_M, s - a->o1);
  if (!mx && !mo) { ++a->null_steps; return; } /* null step (compute_linear.c:170-177) */
  a->null_steps = 0;
  lo = imin(WLO(mx), WLO(mo) - 1);
  hi = imax(WHI(mx), WHI(mo) + 1);
  wf_alloc(out, lo, hi);
  for (k = lo; k <= hi; ++k) {
    int32_t m = vmax(wf_rd(mo, k + 1), vmax(wf_rd(mx, k), wf_rd(mo, k - 1)) + 1);
    if ((uint32_t)m > (uint32_t)a->tlen) m = ONULL;
    if ((uint32_t)(m - k) > (uint32_t)a->plen) m = ONULL;
    out->v[k - lo] = m;
  }
  a->dsteps += hi - lo + 1;
  trim_ends(a, out);
}

static void compute(oaln_t* a, int s) {
  /* wavefront_compute_affine (compute_affine.c:229-260) and
     wavefront_compute_affine2p (compute_affine2p.c:334-368) */
  if (a->metric <= M_LINEAR) { compute_linear(a, s); return; }
  const bool two = (a->metric == 4);
  aln_grow(a, slot_of(a, s) + 1);  /* before taking pointers into the slot arrays */
  owf_t* mx  = wf_src(a, C_M, s - a->x);
  owf_t* mo1 = wf_src(a, C_M, s - a->o1 - a->e1);
  owf_t* i1  = wf_src(a, C_I1, s - a->e1);
  owf_t* d1  = wf_src(a, C_D1, s - a->e1);
  owf_t *mo2 = NULL, *i2 = NULL, *d2 = NULL;
  if (two) {
    mo2 = wf_src(a, C_M, s - a->o2 - a->e2);
    i2  = wf_src(a, C_I2, s - a->e2);
    d2  = wf_src(a, C_D2, s - a->e2);
  }
  owf_t* out[NCOMP];
  int c;
  for (c = 0; c < NCOMP; ++c) out[c] = wf_claim(a, c, s);
  if (!mx && !mo1 && !i1 && !d1 && !mo2 && !i2 && !d2) { ++a->null_steps; return; } /* null step */
  a->null_steps = 0;
  /* wavefront_compute_limits_input, wavefront_compute.c:40-86 */
  int lo = WLO(mx), hi = WHI(mx);
  lo = imin(lo, WLO(mo1) - 1); hi = imax(hi, WHI(mo1) + 1);
  lo = imin(lo, WLO(i1) + 1);  hi = imax(hi, WHI(i1) + 1);
  lo = imin(lo, WLO(d1) - 1);  hi = imax(hi, WHI(d1) - 1);
  if (two) {
    lo = imin(lo, WLO(mo2) - 1); hi = imax(hi, WHI(mo2) + 1);
    lo = imin(lo, WLO(i2) + 1);  hi = imax(hi, WHI(i2) + 1);
    lo = imin(lo, WLO(d2) - 1);  hi = imax(hi, WHI(d2) - 1);
  }
  /* wavefront_compute_allocate_output, wavefront_compute.c:409-494 */
  wf_alloc(out[C_M], lo, hi);
  const bool has_i1 = (mo1 || i1), has_d1 = (mo1 || d1);
  const bool has_i2 = two && (mo2 || i2), has_d2 = two && (mo2 || d2);
  if (has_i1) wf_alloc(out[C_I1], lo, hi);
  if (has_d1) wf_alloc(out[C_D1], lo, hi);
  if (has_i2) wf_alloc(out[C_I2], lo, hi);
  if (has_d2) wf_alloc(out[C_D2], lo, hi);
  /* kernels, compute_affine.c:44-86 / compute_affine2p.c:45-106 */
  int k;
  for (k = lo; k <= hi; ++k) {
    const int32_t ins1 = vmax(wf_rd(mo1, k - 1), wf_rd(i1, k - 1)) + 1;
    const int32_t del1 = vmax(wf_rd(mo1, k + 1), wf_rd(d1, k + 1));
    const int32_t misms = wf_rd(mx, k) + 1;
    int32_t m;
    if (two) {
      const int32_t ins2 = vmax(wf_rd(mo2, k - 1), wf_rd(i2, k - 1)) + 1;
      const int32_t del2 = vmax(wf_rd(mo2, k + 1), wf_rd(d2, k + 1));
      if (has_i2) out[C_I2]->v[k - lo] = ins2;
      if (has_d2) out[C_D2]->v[k - lo] = del2;
      m = vmax(vmax(del1, del2), vmax(misms, vmax(ins1, ins2)));
    } else {
      m = vmax(del1, vmax(misms, ins1));
    }
    if (has_i1) out[C_I1]->v[k - lo] = ins1;
    if (has_d1) out[C_D1]->v[k - lo] = del1;
    /* null-after-max: the maximum itself must be inside the DP matrix */
    if ((uint32_t)m > (uint32_t)a->tlen) m = ONULL;
    if ((uint32_t)(m - k) > (uint32_t)a->plen) m = ONULL;
    out[C_M]->v[k - lo] = m;
  }
  a->dsteps += hi - lo + 1;
  /* wavefront_compute_process_ends, wavefront_compute.c:614-632 */
  for (c = 0; c < NCOMP; ++c) if (out[c]->present) trim_ends(a, out[c]);
}

/* ---------------------------------------------------------------- extend */

/* wavefront_extend_matches_kernel_*, wavefront_extend_kernels.c:49-92 */
static inline int32_t lcp_extend(const oaln_t* a, int k, int32_t off) {
  int h = off, v = off - k;
  while (v < a->plen && h < a->tlen && a->P[v] == a->T[h]) { ++v; ++h; }
  return h;
}

/* wavefront_termination_end2end, wavefront_termination.c:37-114 */
static bool termination(oaln_t* a, owf_t* m, int s) {
  const int k_end = a->tlen - a->plen;
  owf_t* w = (a->comp_end == C_M) ? m : wf_at(a, a->comp_end, s);
  if (!w || w->lo > k_end || k_end > w->hi) return false;
  if (wf_raw(w, k_end) < a->tlen) return false;
  a->status = ST_END_REACHED;
  a->end_score = s;
  return true;
}

/* wf_distance_end2end (wavefront_heuristic.c:125-133): the larger of what is
   left of the pattern and of the text, or -NULL for a NULL cell. */
static int distance_end2end(const oaln_t* a, int k, int32_t off) {
  return (off >= 0) ? imax(a->plen - (off - k), a->tlen - off) : -ONULL;
}

/* wf_distance_end2end_weighted (wavefront_heuristic.c:134-145): what is left
   of each sequence as a fraction of its length, scaled to the mean length
   mfactor, in single-precision float and truncated. */
static int distance_weighted(const oaln_t* a, int k, int32_t off, int mfactor) {
  const int v = off - k, h = off;
  const int left_v = (int)((float)(a->plen - v) / a->plen * mfactor);
  const int left_h = (int)((float)(a->tlen - h) / a->tlen * mfactor);
  return (off >= 0) ? imax(left_v, left_h) : -ONULL;
}

/* wf-adaptive and wfmash (wavefront_heuristic_wfadaptive and its reduce
   step, wavefront_heuristic.c:190-290): keep the diagonals whose distance to
   the end (plain, or weighted for wfmash) is within max_distance_threshold of
   the least one, scanning in from both ends but never past the end
   diagonal. */
static void wfadaptive_cut(oaln_t* a, owf_t* m, bool weighted) {
  if (a->steps_wait > 0) return;
  const int lo = m->lo, hi = m->hi;
  if (hi - lo + 1 < a->h_min_len) return;
  int32_t* d = malloc(sizeof(int32_t) * (hi - lo + 1));
  int min_d = imax(a->plen, a->tlen), k;
  const int mfactor = (int)((float)(a->plen + a->tlen) / 2);
  for (k = lo; k <= hi; ++k) {
    d[k - lo] = weighted ? distance_weighted(a, k, wf_raw(m, k), mfactor) : distance_end2end(a, k, wf_raw(m, k));
    min_d = imin(min_d, d[k - lo]);
  }
  const int k_end = a->tlen - a->plen;
  const int top = imin(k_end, hi);
  int lo_r = lo;
  for (k = lo; k < top; ++k) {
    if (d[k - lo] - min_d <= a->h_max_thr) break;
    ++lo_r;
  }
  m->lo = lo_r;
  const int bottom = imax(k_end, m->lo);
  int hi_r = hi;
  for (k = hi; k > bottom; --k) {
    if (d[k - lo] - min_d <= a->h_max_thr) break;
    --hi_r;
  }
  m->hi = hi_r;
  free(d);
  a->steps_wait = a->h_steps;
}

/* wavefront_heuristic_xdrop (wavefront_heuristic.c:294-370): Smith-Waterman
   scores (swg_match * (v + h) - s) / 2 of the valid cells, swg_match being
   -match, or -1 when match is 0; once a best score has been seen, the
   diagonals scoring xdrop or more below it are cut from both ends (cells
   with a negative offset are passed over), then the best is raised to this
   wavefront's best (first diagonal in ascending order). */
static void xdrop_cut(oaln_t* a, owf_t* m, int s) {
  if (a->steps_wait > 0) return;
  const int lo = m->lo, hi = m->hi;
  const int swg_match = (a->match != 0) ? -a->match : -1;
  int32_t* sw = malloc(sizeof(int32_t) * (hi - lo + 1));
  int cmax = INT_MIN, cmax_k = 0, k;
  for (k = lo; k <= hi; ++k) {
    const int32_t off = wf_raw(m, k);
    if (off < 0) continue;
    const int v = off - k, h = off;
    sw[k - lo] = (swg_match * (v + h) - s) / 2;
    if (cmax < sw[k - lo]) { cmax = sw[k - lo]; cmax_k = k; }
  }
  if (a->max_sw_k != INT_MAX) {
    const int best = a->max_sw;
    /* best - sw in 32-bit wrap-around arithmetic */
#define XDROP_KEEP(kk) ((int32_t)((uint32_t)best - (uint32_t)sw[(kk) - lo]) < a->h_xdrop)
    for (k = lo; k <= hi; ++k) {
      if (wf_raw(m, k) < 0) continue;
      if (XDROP_KEEP(k)) break;
    }
    m->lo = k;
    for (k = hi; k >= m->lo; --k) {
      if (wf_raw(m, k) < 0) continue;
      if (XDROP_KEEP(k)) break;
    }
    m->hi = k;
#undef XDROP_KEEP
    if (cmax > a->max_sw) { a->max_sw = cmax; a->max_sw_k = cmax_k; }
  } else {
    a->max_sw = cmax;
    a->max_sw_k = cmax_k;
  }
  free(sw);
  a->steps_wait = a->h_steps;
}

/* wavefront_heuristic_zdrop (wavefront_heuristic.c:371-431): the same
   Smith-Waterman scores; a wavefront whose best exceeds the best seen
   replaces it (with its diagonal, offset and score); one whose best falls
   more than zdrop below it drops the alignment, its end set to the best
   cell.  Returns whether the alignment dropped (the countdown then stays). */
static bool zdrop_check(oaln_t* a, owf_t* m, int s) {
  if (a->steps_wait > 0) return false;
  const int swg_match = (a->match != 0) ? -a->match : -1;
  int cmax = INT_MIN, cmax_k = 0, cmax_off = 0, k;
  for (k = m->lo; k <= m->hi; ++k) {
    const int32_t off = wf_raw(m, k);
    if (off < 0) continue;
    const int sw = (swg_match * ((off - k) + off) - s) / 2;
    if (cmax < sw) { cmax = sw; cmax_k = k; cmax_off = off; }
  }
  if (a->max_sw_k != INT_MAX) {
    if (cmax > a->max_sw) {
      a->max_sw = cmax; a->max_wf_score = s; a->max_sw_k = cmax_k; a->max_sw_off = cmax_off;
    } else if ((int32_t)((uint32_t)a->max_sw - (uint32_t)cmax) > a->h_zdrop) {
      a->end_k = a->max_sw_k;
      a->end_off = a->max_sw_off;
      return true;
    }
  } else {
    a->max_sw = cmax; a->max_wf_score = s; a->max_sw_k = cmax_k; a->max_sw_off = cmax_off;
  }
  a->steps_wait = a->h_steps;
  return false;
}

/* wavefront_heuristic_banded_adaptive (wavefront_heuristic.c:477-513): a
   wavefront wider than the band keeps a band-wide window placed by the
   distances to the end sampled at its ends and quarters. */
static void banded_adaptive(oaln_t* a, owf_t* m) {
  if (a->steps_wait > 0) return;
  const int lo = m->lo, hi = m->hi;
  const int len = hi - lo + 1;
  if (len < 4) return;
  const int max_len = a->band_max_k - a->band_min_k + 1;
  if (len > max_len) {
    const int leeway = (len - max_len) / 2, quarter = len / 4;
    const int d0 = distance_end2end(a, lo, wf_raw(m, lo));
    const int d1 = distance_end2end(a, lo + quarter, wf_raw(m, lo + quarter));
    const int d2 = distance_end2end(a, lo + 2 * quarter, wf_raw(m, lo + 2 * quarter));
    const int d3 = distance_end2end(a, hi, wf_raw(m, hi));
    int new_lo = lo;
    if (d0 > d3) new_lo += leeway;
    if (d1 > d2) new_lo += leeway;
    m->lo = imax(new_lo, lo);
    m->hi = imin(new_lo + max_len - 1, hi);
  }
  a->steps_wait = a->h_steps;
}

static bool has_cutoff(const oaln_t* a) {
  return a->wfadaptive || a->wfmash || a->xdrop || a->zdrop || a->banded_static || a->banded_adaptive;
}

/* wavefront_heuristic_cufoff (wavefront_heuristic.c:506-561) after the
   extension of score s: one countdown for all strategies, wf-adaptive then the
   band (static: clamp to [min_k, max_k] every time), and when M changed,
   I1/D1/I2/D2 narrowed to M's range.  The cut narrows [lo,hi] only: the cut
   cells stay readable (rlo/rhi) except for indel / edit, whose compute step
   NULLs the two cells beyond each end (wavefront_compute_edit.c:355-358). */
static bool cutoff(oaln_t* a, int s) {
  owf_t* m = wf_at(a, C_M, s);
  if (!m || m->lo > m->hi) return false;
  --a->steps_wait;
  const int lo = m->lo, hi = m->hi;
  if (a->wfadaptive) wfadaptive_cut(a, m, false);
  else if (a->wfmash) wfadaptive_cut(a, m, true);
  if (a->xdrop) xdrop_cut(a, m, s);
  else if (a->zdrop && zdrop_check(a, m, s)) return true;
  if (a->banded_static) {
    if (m->lo < a->band_min_k) m->lo = a->band_min_k;
    if (m->hi > a->band_max_k) m->hi = a->band_max_k;
  } else if (a->banded_adaptive) {
    banded_adaptive(a, m);
  }
  if (m->lo == lo && m->hi == hi) return false;
  if (m->lo > m->hi) m->null = true;
  if (a->metric <= M_EDIT) { m->rlo = m->lo; m->rhi = m->hi; }
  if (a->metric <= M_LINEAR) return false;
  int c;
  for (c = C_I1; c < NCOMP; ++c) {  /* wf_heuristic_equate (:163-171) */
    owf_t* w = wf_at(a, c, s);
    if (!w) continue;
    if (m->lo > w->lo) w->lo = m->lo;
    if (m->hi < w->hi) w->hi = m->hi;
    if (w->lo > w->hi) w->null = true;
  }
  return false;
}

/* A z-drop ends the alignment as unreachable at score s, its end cell the
   best one seen (wavefront_extend.c:112-118, 201-207, 284-290). */
static int dropped(oaln_t* a, int s) {
  a->status = ST_END_UNREACHABLE;
  a->end_score = s;
  return 1;
}

/* wavefront_termination_endsfree, wavefront_termination.c:115-162 */
static bool termination_endsfree(oaln_t* a, int s, int k, int32_t off) {
  const int h = off, v = off - k;
  const bool done = (h >= a->tlen && a->plen - v <= a->pef) || (v >= a->plen && a->tlen - h <= a->tef);
  if (!done) return false;
  a->status = ST_END_REACHED;
  a->end_score = s;
  a->end_k = k;
  a->end_off = off;
  return true;
}

/* A null step leaves no M wavefront; more of them in a row than the score
   scope spans means no later wavefront can have an input: the end is
   unreachable, which only a heuristic cut-off can bring about
   (wavefront_extend.c:94-101, 184-191, 268-275). */
static int unreachable(oaln_t* a, int s) {
  if (a->null_steps <= a->scope) return 0;
  a->status = ST_END_UNREACHABLE;
  a->end_score = s;
  return 1;
}

/* wavefront_extend_endsfree, wavefront_extend.c:212-330, with the scalar
   kernel wavefront_extend_matches_packed_endsfree (wavefront_extend_kernels.c:
   154-196): diagonals in ascending order, each extended and then tested; the
   first that ends the alignment stops the extension. */
static int extend_endsfree(oaln_t* a, int s) {
  owf_t* m = wf_at(a, C_M, s);
  if (!m) return unreachable(a, s);
  int k;
  for (k = m->lo; k <= m->hi; ++k) {
    const int32_t off = wf_raw(m, k);
    if (off == ONULL) continue;
    const int32_t ext = lcp_extend(a, k, off);
    m->v[k - m->alo] = ext;
    if (termination_endsfree(a, s, k, ext)) return 1;
  }
  if (has_cutoff(a) && cutoff(a, s)) return dropped(a, s);
  return 0;
}

/* wavefront_extend_end2end{,_max}, wavefront_extend.c:86-211.  Returns 1 when
   the end is reached.  max_ak (if given) receives the max antidiagonal. */
static int extend(oaln_t* a, int s, int* max_ak) {
  if (a->endsfree) return extend_endsfree(a, s);
  if (max_ak) *max_ak = 0;
  owf_t* m = wf_at(a, C_M, s);
  if (!m) return unreachable(a, s);
  int32_t best = 0;
  int k;
  for (k = m->lo; k <= m->hi; ++k) {
    const int32_t off = wf_raw(m, k);
    if (off == ONULL) continue;
    const int32_t ext = lcp_extend(a, k, off);
    m->v[k - m->alo] = ext;
    const int32_t ak = 2 * ext - k;
    if (best < ak) best = ak;
  }
  if (termination(a, m, s)) {
    a->end_k = a->tlen - a->plen;
    a->end_off = a->tlen;
    return 1;
  }
  if (has_cutoff(a) && cutoff(a, s)) return dropped(a, s);
  if (max_ak) *max_ak = best;
  return 0;
}

/* ---------------------------------------------------------------- align */

static void aln_init(oaln_t* a, int comp_begin, int comp_end) {
  /* wavefront_aligner_init + init_wf, wavefront_aligner.c:252-418 */
  aln_reset_slots(a);
  a->comp_begin = comp_begin;
  a->comp_end = comp_end;
  a->status = ST_OK;
  a->end_score = -1;
  a->null_steps = 0;  /* wavefront_aligner.c:411 */
  a->max_sw = 0;      /* wavefront_heuristic_clear (wavefront_heuristic.c:114-121) */
  a->max_sw_k = INT_MAX;
  a->max_sw_off = ONULL;
  a->max_wf_score = 0;
  a->end_k = INT_MAX;  /* alignment_end_pos cleared (wavefront_aligner.c:411-414) */
  a->end_off = ONULL;
  a->stale_hit = false;
  a->steps_wait = a->h_steps;  /* wavefront_heuristic_clear (wavefront_heuristic.c:114-121) */
  owf_t* w[NCOMP];
  int c;
  for (c = 0; c < NCOMP; ++c) w[c] = wf_claim(a, c, 0);
  if (a->endsfree && a->match == 0 && comp_begin == C_M) {
    /* free beginnings: the score-0 wavefront holds (v=0, h) for h <= text
       begin-free and (v, h=0) for v <= pattern begin-free, untrimmed
       (wavefront_aligner_init_wf_m, wavefront_aligner.c:252-307) */
    int k;
    wf_alloc(w[C_M], -a->pbf, a->tbf);
    for (k = -a->pbf; k <= a->tbf; ++k) w[C_M]->v[k + a->pbf] = (k > 0) ? k : 0;
    return;
  }
  wf_alloc(w[comp_begin], 0, 0);
  w[comp_begin]->v[0] = 0;
}

static void aln_ops_reserve(oaln_t* a, int n) {
  if (a->ops_cap < n) { free(a->ops); a->ops_cap = n + 16; a->ops = malloc(a->ops_cap); }
}

/* A source of cell (s, k) as the compute step reads it (a null wavefront is
   no input) or, with probe_view, as the backtrace probes it (a wavefront a
   cut-off emptied keeps its readable cells); below 0 reads as NULL. */
static int32_t rd_view(oaln_t* a, int c, int s, int k, bool probe_view) {
  if (s < 0) return ONULL;
  owf_t* w = wf_at(a, c, s);
  if (!w || (!probe_view && w->null) || k < w->rlo || k > w->rhi) return ONULL;
  const int32_t v = wf_raw(w, k);
  return v < 0 ? ONULL : v;
}

/* The choices at cell (s, k): best M offset, winning M source and the gap
   extension flags, ties broken M > D2 > D1 > I2 > I1 and extension before
   opening (as the backtrace's maximum of offset * 16 + type does). */
static int64_t cell_choices(oaln_t* a, int s, int k, bool pv) {
  const bool two = (a->metric == 4);
  int32_t best = ((a->metric == M_INDEL) ? ONULL : rd_view(a, C_M, s - a->x, k, pv)) + 1;
  int src = 0, bits = 0;
  if (a->metric <= M_LINEAR) {
    const int32_t del = rd_view(a, C_M, s - a->o1, k + 1, pv), ins = rd_view(a, C_M, s - a->o1, k - 1, pv) + 1;
    if (del > best) { best = del; src = 2; }
    if (ins > best) { best = ins; src = 4; }
  } else {
    const int so1 = s - a->o1 - a->e1, se1 = s - a->e1;
    const int32_t oml = rd_view(a, C_M, so1, k - 1, pv), omr = rd_view(a, C_M, so1, k + 1, pv);
    const int32_t eil = rd_view(a, C_I1, se1, k - 1, pv), edr = rd_view(a, C_D1, se1, k + 1, pv);
    bits = (eil >= oml ? 8 : 0) | (edr >= omr ? 16 : 0);
    int32_t ins2 = ONULL;
    if (two) {
      const int so2 = s - a->o2 - a->e2, se2 = s - a->e2;
      const int32_t o2l = rd_view(a, C_M, so2, k - 1, pv), o2r = rd_view(a, C_M, so2, k + 1, pv);
      const int32_t e2l = rd_view(a, C_I2, se2, k - 1, pv), e2r = rd_view(a, C_D2, se2, k + 1, pv);
      bits |= (e2l >= o2l ? 32 : 0) | (e2r >= o2r ? 64 : 0);
      ins2 = vmax(o2l, e2l) + 1;
      const int32_t del2 = vmax(o2r, e2r);
      if (del2 > best) { best = del2; src = 1; }
    }
    const int32_t del1 = vmax(omr, edr), ins1 = vmax(oml, eil) + 1;
    if (del1 > best) { best = del1; src = 2; }
    if (two && ins2 > best) { best = ins2; src = 3; }
    if (ins1 > best) { best = ins1; src = 4; }
  }
  return ((int64_t)best << 8) | src | bits;
}

/* A backtrace through cell (s, k) probes, where a cut-off emptied a source
   wavefront, cells the compute step did not read (wavefront_backtrace.c:
   64-219 probes any stored wavefront over its initialised range).  When the
   best probed offset passes the cell's extended offset, the match count
   before the cell is negative and operations are taken back (bt_matches).
   A NULL cell is compared by its computed best instead.  Recorded for the
   tests of a device path that declines such traces. */
static void cell_stale(oaln_t* a, int s, int k) {
  if (!(has_cutoff(a))) return;
  const int64_t q = cell_choices(a, s, k, true) >> 8;
  const int64_t p = cell_choices(a, s, k, false) >> 8;
  owf_t* w = wf_at(a, C_M, s);
  const int64_t post = (w && k >= w->alo && k <= w->ahi) ? wf_raw(w, k) : ONULL;
  if (q >= 0 && q > (post > p ? post : p)) a->stale_hit = true;
}

static int64_t probe(oaln_t* a, int c, int s, int k, int add, int type) {
  if (s < 0) return ONULL;
  owf_t* w = wf_at(a, c, s);
  /* the backtrace probes the initialised range (wavefront_backtrace.c:64-219) */
  if (!w || k < w->rlo || k > w->rhi) return ONULL;
  return (int64_t)(wf_raw(w, k) + add) * 16 + type;
}

/* wavefront_backtrace_matches (wavefront_backtrace.c:77-98): n matches go
   before the CIGAR written so far.  The write position moves by n whatever
   its sign: a source offset past the current one, which a heuristic cut-off
   can leave readable, takes back the last -n operations. */
static void bt_matches(oaln_t* a, int* pos, int n) {
  if (n < 0) { *pos = imin(*pos - n, a->ops_cap); return; }
  while (n-- > 0) a->ops[--*pos] = 'M';
}

/* wavefront_backtrace_linear, wavefront_backtrace.c:223-319 (indel, edit,
   gap-linear): M only, ties M > D > I. */
static void backtrace_linear(oaln_t* a, int s_end, int k_end, int off_end) {
  aln_ops_reserve(a, 2 * (a->plen + a->tlen) + 8);
  int pos = a->ops_cap;
  int s = s_end, k = k_end, off = off_end;
  int h = off, v = off - k, i;
  for (i = a->plen - v; i > 0; --i) a->ops[--pos] = 'D';
  for (i = a->tlen - h; i > 0; --i) a->ops[--pos] = 'I';
  while (v > 0 && h > 0 && s > 0) {
    cell_stale(a, s, k);
    const int sx = s - a->x, so = s - a->o1;
    const int64_t misms = (a->metric != M_INDEL) ? probe(a, C_M, sx, k, 1, BT_M) : ONULL;
    const int64_t ins = probe(a, C_M, so, k - 1, 1, BT_I1O);
    const int64_t del = probe(a, C_M, so, k + 1, 0, BT_D1O);
    const int64_t best = lmax(misms, lmax(ins, del));
    if (best < 0) break;  /* no source */
    const int max_off = (int)(best >> 4);
    bt_matches(a, &pos, off - max_off);
    off = max_off;
    v = off - k; h = off;
    if (v <= 0 || h <= 0) break;
    switch ((int)(best & 15)) {
      case BT_M:   s = sx; a->ops[--pos] = 'X'; --off; break;
      case BT_I1O: s = so; a->ops[--pos] = 'I'; --k; --off; break;
      default:     s = so; a->ops[--pos] = 'D'; ++k; break;
    }
    v = off - k; h = off;
  }
  if (v > 0 && h > 0) {
    int n = imin(v, h);
    v -= n; h -= n;
    while (n-- > 0) a->ops[--pos] = 'M';
  }
  while (v > 0) { a->ops[--pos] = 'D'; --v; }
  while (h > 0) { a->ops[--pos] = 'I'; --h; }
  a->ops_begin = pos;
  a->ops_end = a->ops_cap;
}

/* wavefront_backtrace_affine, wavefront_backtrace.c:320-529.  Ops are written
   backwards so that a->ops[ops_begin..ops_end) is the CIGAR. */
static void backtrace(oaln_t* a, int s_end, int k_end, int off_end) {
  if (a->metric <= M_LINEAR) { backtrace_linear(a, s_end, k_end, off_end); return; }
  const bool two = (a->metric == 4);
  aln_ops_reserve(a, 2 * (a->plen + a->tlen) + 8);
  int pos = a->ops_cap;               /* next op goes to ops[--pos] */
  int comp = a->comp_end;
  int s = s_end, k = k_end, off = off_end;
  int h = off, v = off - k;
  if (comp == C_M) {
    int i;
    for (i = a->plen - v; i > 0; --i) a->ops[--pos] = 'D';
    for (i = a->tlen - h; i > 0; --i) a->ops[--pos] = 'I';
  }
  while (v > 0 && h > 0 && s > 0) {
    cell_stale(a, s, k);
    const int sx = s - a->x, so1 = s - a->o1 - a->e1, se1 = s - a->e1;
    const int so2 = s - a->o2 - a->e2, se2 = s - a->e2;
    int64_t best;
    switch (comp) {
      case C_M: {
        const int64_t misms = probe(a, C_M, sx, k, 1, BT_M);
        const int64_t ins1 = lmax(probe(a, C_M, so1, k - 1, 1, BT_I1O), probe(a, C_I1, se1, k - 1, 1, BT_I1E));
        const int64_t del1 = lmax(probe(a, C_M, so1, k + 1, 0, BT_D1O), probe(a, C_D1, se1, k + 1, 0, BT_D1E));
        if (!two) { best = lmax(misms, lmax(ins1, del1)); break; }
        const int64_t ins2 = lmax(probe(a, C_M, so2, k - 1, 1, BT_I2O), probe(a, C_I2, se2, k - 1, 1, BT_I2E));
        const int64_t del2 = lmax(probe(a, C_M, so2, k + 1, 0, BT_D2O), probe(a, C_D2, se2, k + 1, 0, BT_D2E));
        best = lmax(misms, lmax(lmax(ins1, ins2), lmax(del1, del2)));
        break;
      }
      case C_I1: best = lmax(probe(a, C_M, so1, k - 1, 1, BT_I1O), probe(a, C_I1, se1, k - 1, 1, BT_I1E)); break;
      case C_I2: best = lmax(probe(a, C_M, so2, k - 1, 1, BT_I2O), probe(a, C_I2, se2, k - 1, 1, BT_I2E)); break;
      case C_D1: best = lmax(probe(a, C_M, so1, k + 1, 0, BT_D1O), probe(a, C_D1, se1, k + 1, 0, BT_D1E)); break;
      default:   best = lmax(probe(a, C_M, so2, k + 1, 0, BT_D2O), probe(a, C_D2, se2, k + 1, 0, BT_D2E)); break;
    }
    if (best < 0) break;  /* no source */
    if (comp == C_M) {
      const int max_off = (int)(best >> 4);
      bt_matches(a, &pos, off - max_off);
      off = max_off;
      v = off - k; h = off;
      if (v <= 0 || h <= 0) break;
    }
    const int type = (int)(best & 15);
    switch (type) {
      case BT_M:   s = sx;  comp = C_M;  break;
      case BT_I1O: s = so1; comp = C_M;  break;
      case BT_I1E: s = se1; comp = C_I1; break;
      case BT_I2O: s = so2; comp = C_M;  break;
      case BT_I2E: s = se2; comp = C_I2; break;
      case BT_D1O: s = so1; comp = C_M;  break;
      case BT_D1E: s = se1; comp = C_D1; break;
      case BT_D2O: s = so2; comp = C_M;  break;
      default:     s = se2; comp = C_D2; break;
    }
    if (type == BT_M) { a->ops[--pos] = 'X'; --off; }
    else if (type <= BT_I2E) { a->ops[--pos] = 'I'; --k; --off; }
    else { a->ops[--pos] = 'D'; ++k; }
    v = off - k; h = off;
  }
  if (comp == C_M) {
    if (v > 0 && h > 0) {
      int n = imin(v, h);
      v -= n; h -= n;
      while (n-- > 0) a->ops[--pos] = 'M';
    }
    while (v > 0) { a->ops[--pos] = 'D'; --v; }
    while (h > 0) { a->ops[--pos] = 'I'; --h; }
  }
  /* (the reference aborts on a gap-component backtrace that does not end at
     the origin, wavefront_backtrace.c:512-521; it does not happen for exact
     alignment) */
  a->ops_begin = pos;
  a->ops_end = a->ops_cap;
}

/* wavefront_unialign, wavefront_unialign.c:242-273 (+ terminate :147-237) */
static int unialign(oaln_t* a, bool score_only) {
  int s = 0;
  a->ops_begin = a->ops_end = 0;
  a->cigar_score = INT32_MIN;
  while (1) {
    if (extend(a, s, NULL)) {
      if (a->status == ST_END_UNREACHABLE) {
        /* wavefront_unialign_terminate (wavefront_unialign.c:149-236) at the
           end cell a z-drop recorded, or at none (alignment_end_pos keeps
           k = INT_MAX and offset = NULL, wavefront_aligner.c:411-414): a
           score-only run takes the classic score at that end, with the
           wrap-around of 32-bit int arithmetic; a backtrace runs from it
           at the current score when it is not NULL */
        if (score_only) {
          const int32_t v = (int32_t)((uint32_t)a->end_off - (uint32_t)a->end_k), h = a->end_off;
          if (a->metric <= M_EDIT || a->match == 0) {
            a->cigar_score = classic_score(a, 0, 0, s);
          } else {
            const uint32_t sum = (uint32_t)(-a->match) * ((uint32_t)v + (uint32_t)h) - (uint32_t)s;
            a->cigar_score = (int32_t)sum / 2;
          }
        } else if (a->end_off != ONULL) {
          backtrace(a, s, a->end_k, a->end_off);
        }
        return ST_ALG_PARTIAL;
      }
      if (score_only) {
        /* compute_score: the score is taken at the full lengths (:159-166) */
        a->cigar_score = classic_score(a, a->plen, a->tlen, s);
        return ST_ALG_COMPLETED;
      }
      backtrace(a, s, a->end_k, a->end_off);
      const int end_v = a->end_off - a->end_k, end_h = a->end_off;
      a->cigar_score = classic_score(a, end_v, end_h, s);
      return ST_ALG_COMPLETED;
    }
    ++s;
    compute(a, s);
    if (s >= a->max_steps) {
      a->cigar_score = -a->max_steps;
      return ST_MAX_STEPS;
    }
    if (a->metric <= M_EDIT) {
      /* A null indel / edit wavefront stays null (the next one spans its
         NULLed ends, wavefront_compute_edit.c:344-352) and the reference
         computes on to the step limit; its result is taken here at once. */
      owf_t* m = wf_at(a, C_M, s);
      if (m && m->null) {
        a->cigar_score = -a->max_steps;
        return ST_MAX_STEPS;
      }
    }
  }
}

/* ----------------------------------------------------------------- BiWFA */

typedef struct {
  int score, score_forward, score_reverse;
  int k_forward, k_reverse, offset_forward, offset_reverse;
  int component;
} obp_t;

typedef struct {
  oaln_t fwd, rev, base;
  const char *P, *T, *Prev, *Trev;
  int plen, tlen;
  /* output CIGAR, appended forwards */
  char* ops; int ops_len; int ops_cap;
  int cigar_score;
} obi_t;

static void bi_append(obi_t* b, const char* src, int n) {
  if (b->ops_len + n > b->ops_cap) {
    b->ops_cap = 2 * (b->ops_len + n) + 16;
    b->ops = realloc(b->ops, b->ops_cap);
  }
  memcpy(b->ops + b->ops_len, src, n);
  b->ops_len += n;
}
static void bi_append_run(obi_t* b, char op, int n) {
  if (b->ops_len + n > b->ops_cap) {
    b->ops_cap = 2 * (b->ops_len + n) + 16;
    b->ops = realloc(b->ops, b->ops_cap);
  }
  memset(b->ops + b->ops_len, op, n);
  b->ops_len += n;
}

/* wavefront_bialign_breakpoint_indel2indel (scalar), wavefront_bialign.c:477-573 */
static void bp_indel2indel(oaln_t* a0, bool forward, int s0, int s1,
                           owf_t* w0, owf_t* w1, int comp, int gap_open, obp_t* bp) {
  const int plen = a0->plen, tlen = a0->tlen;
  const int lo_0 = w0->lo, hi_0 = w0->hi;
  const int lo_1 = tlen - plen - w1->hi, hi_1 = tlen - plen - w1->lo;
  if (hi_1 < lo_0 || hi_0 < lo_1) return;
  const int min_hi = imin(hi_0, hi_1), max_lo = imax(lo_0, lo_1);
  if (s0 + s1 - gap_open >= bp->score) return;
  int k0;
  for (k0 = max_lo; k0 <= min_hi; ++k0) {
    const int k1 = tlen - plen - k0;
    const int32_t dh0 = wf_raw(w0, k0), dh1 = wf_raw(w1, k1);
    if (dh0 + dh1 >= tlen) {
      if (forward) {
        if (dh0 - k0 > plen || dh0 > tlen) continue;
        bp->score_forward = s0; bp->score_reverse = s1;
        bp->k_forward = k0; bp->k_reverse = k1;
        bp->offset_forward = dh0; bp->offset_reverse = dh1;
      } else {
        if (dh1 - k1 > plen || dh1 > tlen) continue;
        bp->score_forward = s1; bp->score_reverse = s0;
        bp->k_forward = k1; bp->k_reverse = k0;
        bp->offset_forward = dh1; bp->offset_reverse = dh0;
      }
      bp->score = s0 + s1 - gap_open;
      bp->component = comp;
      return;
    }
  }
}

/* wavefront_bialign_breakpoint_m2m (scalar), wavefront_bialign.c:798-874 */
static void bp_m2m(oaln_t* a0, bool forward, int s0, int s1, owf_t* w0, owf_t* w1, obp_t* bp) {
  if (s0 + s1 >= bp->score) return;
  const int plen = a0->plen, tlen = a0->tlen;
  const int lo_0 = w0->lo, hi_0 = w0->hi;
  const int lo_1 = tlen - plen - w1->hi, hi_1 = tlen - plen - w1->lo;
  if (hi_1 < lo_0 || hi_0 < lo_1) return;
  const int min_hi = imin(hi_0, hi_1), max_lo = imax(lo_0, lo_1);
  int k0;
  for (k0 = max_lo; k0 <= min_hi; ++k0) {
    const int k1 = tlen - plen - k0;
    const int32_t m0 = wf_raw(w0, k0), m1 = wf_raw(w1, k1);
    if (m0 + m1 >= tlen) {
      if (forward) {
        bp->score_forward = s0; bp->score_reverse = s1;
        bp->k_forward = k0; bp->k_reverse = k1;
        bp->offset_forward = m0; bp->offset_reverse = m1;
      } else {
        bp->score_forward = s1; bp->score_reverse = s0;
        bp->k_forward = k1; bp->k_reverse = k0;
        bp->offset_forward = m1; bp->offset_reverse = m0;
      }
      bp->score = s0 + s1;
      bp->component = C_M;
      return;
    }
  }
}

/* wavefront_bialign_overlap, wavefront_bialign.c:878-956 */
static void bi_overlap(oaln_t* a0, oaln_t* a1, int s0, int s1, bool forward, obp_t* bp) {
  const bool two = (a0->metric == 4);
  owf_t* m0 = wf_at(a0, C_M, s0);
  if (!m0) return;
  owf_t* d1w0 = wf_at(a0, C_D1, s0);
  owf_t* i1w0 = wf_at(a0, C_I1, s0);
  owf_t* d2w0 = two ? wf_at(a0, C_D2, s0) : NULL;
  owf_t* i2w0 = two ? wf_at(a0, C_I2, s0) : NULL;
  int i;
  for (i = 0; i < a0->scope; ++i) {
    const int si = s1 - i;
    if (si < 0) break;
    if (two) {
      if (s0 + si - a0->o2 >= bp->score) continue;
      owf_t* d2w1 = wf_at(a1, C_D2, si);
      if (d2w0 && d2w1) bp_indel2indel(a0, forward, s0, si, d2w0, d2w1, C_D2, a0->o2, bp);
      owf_t* i2w1 = wf_at(a1, C_I2, si);
      if (i2w0 && i2w1) bp_indel2indel(a0, forward, s0, si, i2w0, i2w1, C_I2, a0->o2, bp);
    }
    if (a0->metric >= 3) {  /* I1/D1 breakpoints: gap-affine metrics only */
      if (s0 + si - a0->o1 >= bp->score) continue;
      owf_t* d1w1 = wf_at(a1, C_D1, si);
      if (d1w0 && d1w1) bp_indel2indel(a0, forward, s0, si, d1w0, d1w1, C_D1, a0->o1, bp);
      owf_t* i1w1 = wf_at(a1, C_I1, si);
      if (i1w0 && i1w1) bp_indel2indel(a0, forward, s0, si, i1w0, i1w1, C_I1, a0->o1, bp);
    }
    if (s0 + si >= bp->score) continue;
    owf_t* mw1 = wf_at(a1, C_M, si);
    if (mw1) bp_m2m(a0, forward, s0, si, m0, mw1, bp);
  }
}

static void bi_bounds(obi_t* b, int pb, int pe, int tb, int te) {
  /* wavefront_sequences_set_bounds, wavefront_sequences.c:274-309 */
  b->fwd.P = b->P + pb; b->fwd.plen = pe - pb;
  b->fwd.T = b->T + tb; b->fwd.tlen = te - tb;
  b->rev.P = b->Prev + (b->plen - pe); b->rev.plen = pe - pb;
  b->rev.T = b->Trev + (b->tlen - te); b->rev.tlen = te - tb;
  b->base.P = b->fwd.P; b->base.plen = b->fwd.plen;
  b->base.T = b->fwd.T; b->base.tlen = b->fwd.tlen;
}

/* wavefront_bialign_find_breakpoint, wavefront_bialign.c:975-1083 */
static int bi_find_breakpoint(obi_t* b, int cb, int ce, obp_t* bp) {
  oaln_t *f = &b->fwd, *r = &b->rev;
  aln_init(f, cb, ce);
  aln_init(r, ce, cb);
  const int max_steps = f->max_steps;
  const int max_ad = f->plen + f->tlen - 1;
  int sf = 0, sr = 0, fmax = 0, rmax = 0, ak = 0;
  bp->score = INT_MAX;
  if (extend(f, sf, &fmax)) return f->status;
  if (extend(r, sr, &rmax)) return r->status;
  bool last_forward = false;
  while (1) {
    if (fmax + rmax >= max_ad) break;
    ++sf; compute(f, sf);
    int quit = extend(f, sf, &ak);
    if (fmax < ak) fmax = ak;
    last_forward = true;
    if (quit) return f->status;
    if (fmax + rmax >= max_ad) break;
    ++sr; compute(r, sr);
    quit = extend(r, sr, &ak);
    if (rmax < ak) rmax = ak;
    last_forward = false;
    if (quit) return r->status;
    if (sr + sf >= max_steps) return ST_MAX_STEPS;
  }
  const int scope = f->scope;
  /* wavefront_bialign_overlap_gopen_adjust, wavefront_bialign.c:960-973 */
  const int gap_open = (f->metric == 4) ? imax(f->o1, f->o2) : (f->metric == 3) ? f->o1 : 0;
  while (1) {
    if (last_forward) {
      const int min_sr = (sr > scope - 1) ? sr - (scope - 1) : 0;
      if (sf + min_sr - gap_open >= bp->score) break;
      bi_overlap(f, r, sf, sr, true, bp);
      ++sr; compute(r, sr);
      if (extend(r, sr, NULL)) return r->status;
    }
    const int min_sf = (sf > scope - 1) ? sf - (scope - 1) : 0;
    if (min_sf + sr - gap_open >= bp->score) break;
    bi_overlap(r, f, sr, sf, false, bp);
    ++sf; compute(f, sf);
    if (extend(f, sf, NULL)) return f->status;
    if (sr + sf >= max_steps) return ST_MAX_STEPS;
    last_forward = true;
  }
  return ST_OK;
}

/* wavefront_bialign_base, wavefront_bialign.c:159-190 */
static int bi_base(obi_t* b, int cb, int ce) {
  oaln_t* a = &b->base;
  aln_init(a, cb, ce);
  const int st = unialign(a, false);
  if (st == ST_ALG_COMPLETED || st == ST_MAX_STEPS) {
    if (a->ops_end > a->ops_begin) bi_append(b, a->ops + a->ops_begin, a->ops_end - a->ops_begin);
  }
  b->cigar_score = a->cigar_score;
  return (st == ST_ALG_COMPLETED) ? ST_OK : ST_UNATTAINABLE;
}

/* wavefront_bialign_alignment, wavefront_bialign.c:1145-1222 */
static int bi_alignment(obi_t* b, int pb, int pe, int tb, int te,
                        int cb, int ce, int score_remaining, int level) {
  const int plen = pe - pb, tlen = te - tb;
  if (tlen == 0) { bi_append_run(b, 'D', plen); return ST_OK; }
  if (plen == 0) { bi_append_run(b, 'I', tlen); return ST_OK; }
  bi_bounds(b, pb, pe, tb, te);
  if (score_remaining <= BI_FALLBACK_MIN_SCORE) return bi_base(b, cb, ce);
  obp_t bp;
  int st = bi_find_breakpoint(b, cb, ce, &bp);
  if (st != ST_OK) {
    /* wavefront_bialign_find_breakpoint_exception, :1084-1111 */
    if (st == ST_END_REACHED) {
      const int reached = (b->fwd.status == ST_END_REACHED) ? b->fwd.end_score : b->rev.end_score;
      if (reached <= BI_RECOVERY_MIN_SCORE) return bi_base(b, cb, ce);
      return ST_END_UNREACHABLE;
    }
    return st;
  }
  const int bp_h = bp.offset_forward;
  const int bp_v = bp.offset_forward - bp.k_forward;
  st = bi_alignment(b, pb, pb + bp_v, tb, tb + bp_h, cb, bp.component, bp.score_forward, level + 1);
  if (st != ST_OK) return st;
  st = bi_alignment(b, pb + bp_v, pe, tb + bp_h, te, bp.component, ce, bp.score_reverse, level + 1);
  if (st != ST_OK) return st;
  if (level == 0) b->cigar_score = classic_score(&b->fwd, plen, tlen, bp.score);
  return ST_OK;
}

static char* reversed(const char* s, int n) {
  char* r = malloc(n + 1);
  int i;
  for (i = 0; i < n; ++i) r[i] = s[n - 1 - i];
  r[n] = 0;
  return r;
}

/* wavefront_bialign, wavefront_bialign.c:1267-1294 */
static int bialign(const oracle_conf_t* conf, const char* P, int plen, const char* T, int tlen,
                   int* score, char* ops, int* ops_len, int64_t* dsteps, int* end_score) {
  obi_t b;
  memset(&b, 0, sizeof(b));
  aln_set_penalties(&b.fwd, conf); b.fwd.modular = true;
  aln_set_penalties(&b.rev, conf); b.rev.modular = true;
  aln_set_penalties(&b.base, conf); b.base.modular = false;
  /* the base aligner runs without heuristic (wavefront_bialigner.c:64-67) */
  b.base.wfadaptive = b.base.wfmash = b.base.xdrop = b.base.zdrop = false;
  b.base.banded_static = b.base.banded_adaptive = false;
  b.P = P; b.T = T; b.plen = plen; b.tlen = tlen;
  b.Prev = reversed(P, plen); b.Trev = reversed(T, tlen);
  b.cigar_score = INT32_MIN;
  int st;
  bi_bounds(&b, 0, plen, 0, tlen);
  if (conf->score_only) {
    /* wavefront_bialign_compute_score, :1226-1262 */
    obp_t bp;
    st = bi_find_breakpoint(&b, C_M, C_M, &bp);
    if (st == ST_OK || st == ST_END_REACHED) {
      if (st == ST_END_REACHED)
        bp.score = (b.fwd.status == ST_END_REACHED) ? b.fwd.end_score : b.rev.end_score;
      b.cigar_score = classic_score(&b.fwd, plen, tlen, bp.score);
      if (end_score) *end_score = bp.score;
      st = ST_OK;
    }
  } else {
    const int min_length = imax(plen, tlen) <= BI_FALLBACK_MIN_LENGTH;
    st = bi_alignment(&b, 0, plen, 0, tlen, C_M, C_M, min_length ? 0 : INT_MAX, 0);
  }
  int status;
  if (st == ST_OK) status = ST_ALG_COMPLETED;
  else if (st == ST_MAX_STEPS) status = ST_MAX_STEPS;
  else status = ST_UNATTAINABLE;
  *score = b.cigar_score;
  if (ops_len) *ops_len = 0;
  if (!conf->score_only && ops && ops_len) {
    const int n = imin(b.ops_len, plen + tlen);
    memcpy(ops, b.ops, n);
    *ops_len = n;
  }
  if (dsteps) *dsteps = b.fwd.dsteps + b.rev.dsteps + b.base.dsteps;
  aln_free(&b.fwd); aln_free(&b.rev); aln_free(&b.base);
  free(b.ops);
  free((void*)b.Prev); free((void*)b.Trev);
  return status;
}

/* ---------------------------------------------------------------- maxtrim */

/* cigar_maxtrim_gap_{linear,affine,affine2p} (alignment/cigar_utils.c:111-278):
   keeps the prefix of the CIGAR with the highest score, a match scoring
   -match (1 when match is 0).  Returns whether the CIGAR was shortened; *keep
   receives the kept length and *best the kept prefix's score (INT32_MIN when
   the best score is 0 and the CIGAR is cleared); both are left as they are
   when the CIGAR is not examined: indel and edit are never trimmed
   (wavefront_aligner.c:664-676), and affine2p skips an empty CIGAR. */
static bool maxtrim(const oracle_conf_t* c, const char* ops, int n, int* keep, int* best) {
  *keep = n;
  if (c->metric <= M_EDIT) return false;
  const int mscore = (c->match != 0) ? c->match : -1;
  int max_score = 0, max_i = 0, score = 0, i;
  if (c->metric == 4) {
    if (n <= 0) return false;
    char last = 0;
    int run = 0;
    for (i = 0; i <= n; ++i) {
      const char op = (i < n) ? ops[i] : 0;
      if (i == n || (op != last && last != 0)) {
        int cost;
        switch (last) {
          case 'M': cost = mscore * run; break;
          case 'X': cost = c->mismatch * run; break;
          default: cost = imin(c->gap_opening1 + c->gap_extension1 * run, c->gap_opening2 + c->gap_extension2 * run);
        }
        score -= cost;
        if (max_score < score) { max_score = score; max_i = i - 1; }
        run = 0;
      }
      last = op;
      ++run;
    }
  } else {
    char last = 0;
    for (i = 0; i < n; ++i) {
      switch (ops[i]) {
        case 'M': score -= mscore; break;
        case 'X': score -= c->mismatch; break;
        default:
          if (c->metric == M_LINEAR) score -= c->gap_opening1;
          else score -= c->gap_extension1 + ((last == ops[i]) ? 0 : c->gap_opening1);
      }
      last = ops[i];
      if (max_score < score) { max_score = score; max_i = i; }
    }
  }
  const bool trimmed = (max_i != n - 1);
  if (max_score == 0) {
    *keep = 0;
    *best = INT32_MIN;
  } else {
    *keep = max_i + 1;
    *best = max_score;
  }
  return trimmed;
}

/* ------------------------------------------------------------------ API */

int oracle_align(const oracle_conf_t* conf, const char* P, int plen, const char* T, int tlen,
                 int* score, char* ops, int* ops_len, int64_t* dsteps, int* end_score, int* stale) {
  if (end_score) *end_score = -1;
  if (stale) *stale = 0;
  if (ops_len) *ops_len = 0;
  const bool endsfree = conf->ends_free || conf->extension;
  if ((conf->heuristic & 0x30) && conf->metric <= M_EDIT) return ORACLE_UNSUPPORTED;
  /* z-drop alignments in memory modes med / low trace the piggybacked CIGAR
     of another cell (wavefront_unialign.c:177-189); not restated */
  if ((conf->heuristic & 0x30) == 0x20 && !conf->score_only && (conf->memory_mode == 1 || conf->memory_mode == 2))
    return ORACLE_UNSUPPORTED;
  if (conf->memory_mode == 3) {
    if (endsfree) return ORACLE_UNSUPPORTED;
    return bialign(conf, P, plen, T, tlen, score, ops, ops_len, dsteps, end_score);
  }
  oaln_t a;
  memset(&a, 0, sizeof(a));
  aln_set_penalties(&a, conf);
  a.modular = conf->score_only ? true : false;
  a.P = P; a.plen = plen; a.T = T; a.tlen = tlen;
  a.endsfree = endsfree;
  if (conf->extension) {
    a.pbf = 0; a.pef = plen; a.tbf = 0; a.tef = tlen;
  } else if (endsfree) {
    a.pbf = conf->pattern_begin_free; a.pef = conf->pattern_end_free;
    a.tbf = conf->text_begin_free; a.tef = conf->text_end_free;
    if (a.pbf > plen || a.pef > plen || a.tbf > tlen || a.tef > tlen) return ORACLE_UNSUPPORTED;
    if (a.match != 0 && (a.pbf > 0 || a.tbf > 0)) return ORACLE_UNSUPPORTED;
    /* indel / edit with a free beginning: the reference's backtrace sees the
       score-0 wavefront only at k = 0 (its initialised range stays [0,0]:
       wavefront.c:89-109, wavefront_aligner.c:252-307, and
       wavefront_compute_edit.c:329-373 never widens it), and its CIGARs then
       hold 'M' on mismatching bases; not restated */
    if (a.metric <= M_EDIT && (a.pbf > 0 || a.tbf > 0)) return ORACLE_UNSUPPORTED;
  }
  aln_init(&a, C_M, C_M);
  int st = unialign(&a, conf->score_only != 0);
  *score = a.cigar_score;
  const bool dropped = (a.status == ST_END_UNREACHABLE);
  if ((st == ST_ALG_COMPLETED || dropped) && !conf->score_only && ops && ops_len) {
    int n = a.ops_end - a.ops_begin;
    if (conf->extension || dropped) {
      /* wavefront_unialign_terminate, wavefront_unialign.c:188-236: the
         backtrace leaves the raw score; max-trim replaces it.  A dropped
         alignment has no backtrace and keeps the cleared CIGAR's score
         (INT32_MIN, cigar.c:79-90) unless max-trim replaces it. */
      int keep = n, best = (dropped && a.end_off == ONULL) ? INT32_MIN : a.end_score;
      const bool trimmed = maxtrim(conf, a.ops + a.ops_begin, n, &keep, &best);
      *score = best;
      n = keep;
      if (trimmed) st = ST_ALG_PARTIAL;
    }
    memcpy(ops, a.ops + a.ops_begin, n);
    *ops_len = n;
  }
  if (dsteps) *dsteps = a.dsteps;
  if (stale) *stale = a.stale_hit;
  if (end_score) *end_score = (st == ST_ALG_COMPLETED) ? a.end_score : -1;
  aln_free(&a);
  return st;
}

void oracle_align_batch(const oracle_conf_t* conf, int n, const char* buf,
                        const int64_t* p_off, const int* plen, const int64_t* t_off, const int* tlen,
                        int* status, int* score, char* ops_out, const int64_t* ops_off, int* ops_len,
                        int64_t* dsteps, int* end_score, int* stale) {
  int i;
  #pragma omp parallel for schedule(dynamic, 4)
  for (i = 0; i < n; ++i) {
    status[i] = oracle_align(conf, buf + p_off[i], plen[i], buf + t_off[i], tlen[i],
                             &score[i], ops_out ? ops_out + ops_off[i] : NULL,
                             ops_len ? &ops_len[i] : NULL,
                             dsteps ? &dsteps[i] : NULL, end_score ? &end_score[i] : NULL,
                             stale ? &stale[i] : NULL);
  }
}

/* cigar_score_gap_affine / _affine2p, alignment/cigar.c:277-342 */
int oracle_cigar_score(const oracle_conf_t* conf, const char* ops, int len) {
  int score = 0, i;
  if (conf->metric <= M_EDIT) {  /* cigar_score_edit, alignment/cigar.c:244-259 (also for indel) */
    for (i = 0; i < len; ++i) {
      if (ops[i] == 'X' || ops[i] == 'I' || ops[i] == 'D') ++score;
      else if (ops[i] != 'M') return INT_MIN;
    }
    return score;
  }
  if (conf->metric == M_LINEAR) {  /* cigar_score_gap_linear, alignment/cigar.c:260-276 */
    for (i = 0; i < len; ++i) {
      switch (ops[i]) {
        case 'M': score -= conf->match; break;
        case 'X': score -= conf->mismatch; break;
        case 'I': case 'D': score -= conf->gap_opening1; break;
        default: return INT_MIN;
      }
    }
    return score;
  }
  if (conf->metric != 4) {
    char last = 0;
    for (i = 0; i < len; ++i) {
      switch (ops[i]) {
        case 'M': score -= conf->match; break;
        case 'X': score -= conf->mismatch; break;
        case 'D': score -= conf->gap_extension1 + ((last == 'D') ? 0 : conf->gap_opening1); break;
        case 'I': score -= conf->gap_extension1 + ((last == 'I') ? 0 : conf->gap_opening1); break;
        default: return INT_MIN;
      }
      last = ops[i];
    }
    return score;
  }
  i = 0;
  while (i < len) {
    int j = i;
    while (j < len && ops[j] == ops[i]) ++j;
    const int run = j - i;
    switch (ops[i]) {
      case 'M': score -= conf->match * run; break;
      case 'X': score -= conf->mismatch * run; break;
      case 'D': case 'I': {
        const int g1 = conf->gap_opening1 + conf->gap_extension1 * run;
        const int g2 = conf->gap_opening2 + conf->gap_extension2 * run;
        score -= imin(g1, g2);
        break;
      }
      default: return INT_MIN;
    }
    i = j;
  }
  return score;
}
