/*
 * A reference-style C caller, compiled unchanged against include/wfa_mi355x.h
 * and linked with libwfa2_amd.so.  It follows the README usage of WFA2-lib
 * (README.md:88-120: attributes from the default, gap-affine penalties,
 * wavefront_aligner_new, wavefront_align, reading wf_aligner->cigar) and then
 * aligns pairs read from stdin with wavefront_align_batch.
 *
 * stdin: one pair per line, "pattern<TAB>text".  stdout: one ".alg" line per
 * pair, "score<TAB>CIGAR" (run-length ops), the README pair first.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wfa_mi355x.h"

static void print_alg(const cigar_t* cigar, const affine_penalties_t* pen) {
  static char buf[1 << 20];
  const int score = cigar_score_gap_affine(cigar, pen);
  cigar_sprint(buf, cigar, true);
  printf("%d\t%s\n", score, buf);
}

int main(void) {
  const char* pattern = "TCTTTACTCGCGCGTTGGAGAAATACAATAGT";
  const char* text = "TCTATACTGCGCGTTTGGAGAAATAAAATAGT";
  wavefront_aligner_attr_t attributes = wavefront_aligner_attr_default;
  attributes.distance_metric = gap_affine;
  attributes.affine_penalties.match = 0;
  attributes.affine_penalties.mismatch = 4;
  attributes.affine_penalties.gap_opening = 6;
  attributes.affine_penalties.gap_extension = 2;
  attributes.heuristic.strategy = wf_heuristic_none; /* exact alignment */
  wavefront_aligner_t* const wf_aligner = wavefront_aligner_new(&attributes);
  if (wf_aligner == NULL) return 2;
  const int st = wavefront_align(wf_aligner, pattern, (int)strlen(pattern), text, (int)strlen(text));
  if (st != WF_STATUS_ALG_COMPLETED) {
    fprintf(stderr, "wavefront_align: %s\n", wavefront_align_strerror(st));
    return 3;
  }
  if (!cigar_check_alignment(stderr, pattern, (int)strlen(pattern), text, (int)strlen(text),
                             wf_aligner->cigar, true))
    return 4;
  print_alg(wf_aligner->cigar, &attributes.affine_penalties);

  /* Batch of pairs from stdin. */
  size_t cap = 1024, n = 0;
  wfa_pair_t* pairs = malloc(cap * sizeof(wfa_pair_t));
  char* line = NULL;
  size_t line_cap = 0;
  ssize_t got;
  while ((got = getline(&line, &line_cap, stdin)) > 0) {
    if (line[got - 1] == '\n') line[--got] = 0;
    char* tab = strchr(line, '\t');
    if (!tab) continue;
    *tab = 0;
    if (n == cap) pairs = realloc(pairs, (cap *= 2) * sizeof(wfa_pair_t));
    pairs[n].pattern = strdup(line);
    pairs[n].pattern_length = (int)strlen(line);
    pairs[n].text = strdup(tab + 1);
    pairs[n].text_length = (int)strlen(tab + 1);
    ++n;
  }
  free(line);
  if (n > 0) {
    wfa_result_t* results = malloc(n * sizeof(wfa_result_t));
    const size_t arena_bytes = wavefront_align_batch_arena_bytes(pairs, (int)n);
    char* arena = malloc(arena_bytes ? arena_bytes : 1);
    const int bst = wavefront_align_batch(wf_aligner, pairs, (int)n, results, arena, arena_bytes);
    if (bst != WF_STATUS_ALG_COMPLETED) {
      fprintf(stderr, "wavefront_align_batch: %s\n", wavefront_align_strerror(bst));
      return 5;
    }
    for (size_t i = 0; i < n; ++i) {
      cigar_t c;
      memset(&c, 0, sizeof(c));
      c.operations = arena + results[i].cigar_offset;
      c.max_operations = results[i].cigar_length;
      c.begin_offset = 0;
      c.end_offset = results[i].cigar_length;
      c.score = results[i].score;
      c.end_v = results[i].end_v;
      c.end_h = results[i].end_h;
      print_alg(&c, &attributes.affine_penalties);
      free((void*)pairs[i].pattern);
      free((void*)pairs[i].text);
    }
    free(results);
    free(arena);
  }
  free(pairs);
  wavefront_aligner_delete(wf_aligner);
  return 0;
}
