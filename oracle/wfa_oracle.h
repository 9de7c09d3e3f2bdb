/*
 * CPU oracle for the exact WFA hot path (TEST INFRASTRUCTURE ONLY).
 *
 * A plain-C restatement of WFA2-lib v2.3's wavefront_align for indel, edit,
 * gap-linear, gap-affine and dual-cost gap-affine, end-to-end and ends-free
 * (with the extension form), in memory mode high (and med/low, whose
 * CIGARs the reference asserts equal memory-high: tests/wfa.utest.sh:67-78)
 * and ultralow (BiWFA, end-to-end), exact or with the heuristics (wf-adaptive,
 * wfmash, banded static / adaptive, x-drop, z-drop; wavefront_heuristic.c).
 * Ends-free termination takes the first
 * diagonal in ascending order, as the reference's scalar extend kernel
 * (wavefront_extend_kernels.c:154-196).  It is the checker for the HIP path; only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 *
 * Pinning: every function here is checked against the reference's golden
 * vectors (the .alg files under tests/golden/, copied from tests/wfa.utest.check/), against
 * the reference itself compiled from source into oracle/_ref/, and against
 * an independent O(nm) dynamic-programming scorer (tests/test_oracle.py).
 */
#ifndef WFA_ORACLE_H_
#define WFA_ORACLE_H_

#include <stdint.h>

typedef struct {
  int metric;          /* 0 indel, 1 edit, 2 gap-linear, 3 gap-affine, 4 gap-affine-2p
                          (wavefront_penalties.h:41-47) */
  int match;           /* user penalties, as in affine{,2p}_penalties_t; gap-linear:
                          match, mismatch, gap_opening1 = indel (linear_penalties_t) */
  int mismatch;
  int gap_opening1;
  int gap_extension1;
  int gap_opening2;
  int gap_extension2;
  int memory_mode;     /* 0 high, 1 med, 2 low, 3 ultralow (wavefront_attributes.h:104-109) */
  int score_only;      /* alignment_scope == compute_score */
  int max_steps;       /* <= 0 means unlimited (INT_MAX) */
  int ends_free;       /* alignment_form.span == alignment_endsfree */
  int pattern_begin_free, pattern_end_free, text_begin_free, text_end_free;
  int extension;       /* alignment_form.extension (wavefront_align.c:57-63) */
  int heuristic;       /* wavefront_heuristic_t.strategy bits (wavefront_heuristic.h:41-49):
                          1 banded static, 2 banded adaptive, 4 wf-adaptive,
                          0x10 x-drop, 0x40 wfmash */
  int min_wavefront_length, max_distance_threshold, steps_between_cutoffs;
  int min_k, max_k;    /* band of the banded strategies */
  int xdrop;           /* x-drop threshold */
  int zdrop;           /* z-drop threshold (0x20) */
} oracle_conf_t;

/* Status returned for a configuration this restatement does not cover:
   ends-free with a free beginning and match < 0 (wavefront_compute.c:124-262),
   indel / edit with a free beginning (see oracle_align), or what the
   reference rejects with exit(1): ends-free / extension with BiWFA
   (wavefront_align.c:66-80), x-drop / z-drop with indel / edit (:82-89);
   and a z-drop alignment in memory mode med / low, which traces a
   piggybacked CIGAR (wavefront_unialign.c:177-189). */
#define ORACLE_UNSUPPORTED (-400)

/*
 * Align one pair.  Returns the reference's status code (wfa.h:45-55).
 * score      <- cigar->score (classic score, wavefront_compute.c:108-120)
 * ops        <- CIGAR operations (M/X/I/D, one per column), capacity plen+tlen
 * ops_len    <- number of operations written
 * dsteps     <- diagonal-steps: sum of (hi-lo+1) over every non-null compute
 * stale      <- 1 when the backtrace passed a cell whose choices change once
 *               the cells of wavefronts a heuristic cut-off emptied are read
 *               (the reference's backtrace reads them; the compute does not)
 */
int oracle_align(
    const oracle_conf_t* conf,
    const char* pattern, int plen,
    const char* text, int tlen,
    int* score, char* ops, int* ops_len, int64_t* dsteps, int* end_score, int* stale);

/* Batch form over a shared byte buffer (see oracle/oracle.py). */
void oracle_align_batch(
    const oracle_conf_t* conf, int n, const char* buf,
    const int64_t* p_off, const int* plen, const int64_t* t_off, const int* tlen,
    int* status, int* score, char* ops_out, const int64_t* ops_off, int* ops_len,
    int64_t* dsteps, int* end_score, int* stale);

/* cigar_score_edit / _gap_linear / _gap_affine / _gap_affine2p (alignment/cigar.c:244-342). */
int oracle_cigar_score(const oracle_conf_t* conf, const char* ops, int len);

#endif
