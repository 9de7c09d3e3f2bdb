/*
 * Test-infrastructure driver around the reference library (WFA2-lib v2.3),
 * compiled together with the reference's own sources by oracle/Makefile into
 * oracle/_ref/.  It is never linked into the product library.
 *
 * It exposes one flat entry point so that Python (ctypes) can run the
 * reference's wavefront_align over a batch of pairs, one aligner per OpenMP
 * thread, exactly as tools/align_benchmark/align_benchmark.c:504-586 does, and
 * time only the wavefront_align calls (benchmark_gap_affine.c:128-154).
 */
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <omp.h>
#include "wavefront/wavefront_align.h"

/* Penalty vector layout: {match, mismatch, gap_open1, gap_ext1, gap_open2, gap_ext2}.
 * form (may be NULL = end-to-end): {span, pattern_begin_free, pattern_end_free,
 * text_begin_free, text_end_free, extension}.
 * heur (may be NULL = none): {strategy, min_wavefront_length,
 * max_distance_threshold, steps_between_cutoffs, min_k, max_k, xdrop, zdrop}. */
static void ref_shim_attributes(
    wavefront_aligner_attr_t* const attr,
    const int metric,
    const int* const pen,
    const int memory_mode,
    const int score_only,
    const int max_steps,
    const int* const form,
    const int* const heur) {
  *attr = wavefront_aligner_attr_default;
  attr->heuristic.strategy = wf_heuristic_none;
  if (form != NULL) {
    attr->alignment_form.span = form[0] ? alignment_endsfree : alignment_end2end;
    attr->alignment_form.pattern_begin_free = form[1];
    attr->alignment_form.pattern_end_free = form[2];
    attr->alignment_form.text_begin_free = form[3];
    attr->alignment_form.text_end_free = form[4];
    attr->alignment_form.extension = form[5] != 0;
  }
  if (heur != NULL && heur[0] != 0) {
    attr->heuristic.strategy = (wf_heuristic_strategy)heur[0];
    attr->heuristic.min_wavefront_length = heur[1];
    attr->heuristic.max_distance_threshold = heur[2];
    attr->heuristic.steps_between_cutoffs = heur[3];
    attr->heuristic.min_k = heur[4];
    attr->heuristic.max_k = heur[5];
    attr->heuristic.xdrop = heur[6];
    attr->heuristic.zdrop = heur[7];
  }
  attr->alignment_scope = score_only ? compute_score : compute_alignment;
  attr->memory_mode = (wavefront_memory_t)memory_mode;
  if (max_steps > 0) attr->system.max_alignment_steps = max_steps;
  if (metric == 0) {
    attr->distance_metric = indel;
  } else if (metric == 1) {
    attr->distance_metric = edit;
  } else if (metric == 2) {
    attr->distance_metric = gap_linear;
    attr->linear_penalties.match = pen[0];
    attr->linear_penalties.mismatch = pen[1];
    attr->linear_penalties.indel = pen[2];
  } else if (metric == 4) {
    attr->distance_metric = gap_affine_2p;
    attr->affine2p_penalties.match = pen[0];
    attr->affine2p_penalties.mismatch = pen[1];
    attr->affine2p_penalties.gap_opening1 = pen[2];
    attr->affine2p_penalties.gap_extension1 = pen[3];
    attr->affine2p_penalties.gap_opening2 = pen[4];
    attr->affine2p_penalties.gap_extension2 = pen[5];
  } else {
    attr->distance_metric = gap_affine;
    attr->affine_penalties.match = pen[0];
    attr->affine_penalties.mismatch = pen[1];
    attr->affine_penalties.gap_opening = pen[2];
    attr->affine_penalties.gap_extension = pen[3];
  }
}

/*
 * Align n pairs.  Pair i is pattern = buf[p_off[i] .. p_off[i]+plen[i]) and
 * text = buf[t_off[i] .. t_off[i]+tlen[i]).  Outputs per pair: the returned
 * status, cigar->score, the CIGAR re-scored as align_benchmark writes it to
 * .alg files (benchmark_utils.c:143-179), and the raw operations string
 * (one of M/X/I/D per column) stored at ops_out[ops_off[i]..+ops_len[i]).
 * ops_off[i] must leave room for plen[i]+tlen[i] bytes.  Returns the wall
 * time in seconds of the parallel alignment loop.
 */
/* Pairs per dynamic-schedule chunk: 64 for large batches, fewer when the
 * batch is small, so that every thread gets pairs (a batch of 32 long pairs
 * in chunks of 64 would run on one thread). */
static int shim_chunk(const int n,const int nt) {
  int c = n / (8*nt);
  if (c < 1) c = 1;
  if (c > 64) c = 64;
  return c;
}

double ref_align_batch_ex(
    const int metric,
    const int* const pen,
    const int memory_mode,
    const int score_only,
    const int max_steps,
    const int* const form,
    const int* const heur,
    const int num_threads,
    const int n,
    const char* const buf,
    const int64_t* const p_off,
    const int* const plen,
    const int64_t* const t_off,
    const int* const tlen,
    int* const status,
    int* const score,
    int* const alg_score,
    char* const ops_out,
    const int64_t* const ops_off,
    int* const ops_len) {
  wavefront_aligner_attr_t attr;
  ref_shim_attributes(&attr,metric,pen,memory_mode,score_only,max_steps,form,heur);
  const int nt = (num_threads > 0) ? num_threads : 1;
  wavefront_aligner_t** aligners = calloc(nt,sizeof(wavefront_aligner_t*));
  int t;
  for (t=0;t<nt;++t) aligners[t] = wavefront_aligner_new(&attr);
  const int chunk = shim_chunk(n,nt);
  const double t0 = omp_get_wtime();
  int i;
  #pragma omp parallel for num_threads(nt) schedule(dynamic,chunk)
  for (i=0;i<n;++i) {
    wavefront_aligner_t* const wf = aligners[omp_get_thread_num()];
    status[i] = wavefront_align(wf,buf+p_off[i],plen[i],buf+t_off[i],tlen[i]);
    cigar_t* const cigar = wf->cigar;
    score[i] = cigar->score;
    const int len = (score_only) ? 0 : cigar->end_offset - cigar->begin_offset;
    ops_len[i] = (len > 0) ? len : 0;
    if (len > 0) {
      memcpy(ops_out+ops_off[i],cigar->operations+cigar->begin_offset,len);
      alg_score[i] = (metric <= 1) ? cigar_score_edit(cigar) :
          (metric == 2) ? cigar_score_gap_linear(cigar,&attr.linear_penalties) :
          (metric == 4) ? cigar_score_gap_affine2p(cigar,&attr.affine2p_penalties) :
          cigar_score_gap_affine(cigar,&attr.affine_penalties);
    } else {
      alg_score[i] = (score_only) ? cigar->score : -1;
    }
  }
  const double t1 = omp_get_wtime();
  for (t=0;t<nt;++t) wavefront_aligner_delete(aligners[t]);
  free(aligners);
  return t1 - t0;
}

/* End-to-end, heuristic none. */
double ref_align_batch(
    const int metric,
    const int* const pen,
    const int memory_mode,
    const int score_only,
    const int max_steps,
    const int num_threads,
    const int n,
    const char* const buf,
    const int64_t* const p_off,
    const int* const plen,
    const int64_t* const t_off,
    const int* const tlen,
    int* const status,
    int* const score,
    int* const alg_score,
    char* const ops_out,
    const int64_t* const ops_off,
    int* const ops_len) {
  return ref_align_batch_ex(metric,pen,memory_mode,score_only,max_steps,NULL,NULL,num_threads,
      n,buf,p_off,plen,t_off,tlen,status,score,alg_score,ops_out,ops_off,ops_len);
}

/* Time-only variant for the CPU baseline: aligns and discards the results. */
double ref_align_time(
    const int metric,
    const int* const pen,
    const int memory_mode,
    const int num_threads,
    const int n,
    const char* const buf,
    const int64_t* const p_off,
    const int* const plen,
    const int64_t* const t_off,
    const int* const tlen,
    int64_t* const checksum) {
  wavefront_aligner_attr_t attr;
  ref_shim_attributes(&attr,metric,pen,memory_mode,0,0,NULL,NULL);
  const int nt = (num_threads > 0) ? num_threads : 1;
  wavefront_aligner_t** aligners = calloc(nt,sizeof(wavefront_aligner_t*));
  int t;
  for (t=0;t<nt;++t) aligners[t] = wavefront_aligner_new(&attr);
  int64_t sum = 0;
  const int chunk = shim_chunk(n,nt);
  const double t0 = omp_get_wtime();
  int i;
  #pragma omp parallel for num_threads(nt) schedule(dynamic,chunk) reduction(+:sum)
  for (i=0;i<n;++i) {
    wavefront_aligner_t* const wf = aligners[omp_get_thread_num()];
    wavefront_align(wf,buf+p_off[i],plen[i],buf+t_off[i],tlen[i]);
    sum += wf->cigar->score;
  }
  const double t1 = omp_get_wtime();
  for (t=0;t<nt;++t) wavefront_aligner_delete(aligners[t]);
  free(aligners);
  *checksum = sum;
  return t1 - t0;
}
