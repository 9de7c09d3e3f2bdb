/*
 * Host-side C parts of the drop-in API that never touch the device:
 * attribute defaults, configuration setters, status messages and the CIGAR
 * utilities callers use on results (alignment/cigar.c of the reference).
 */
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "wfa_mi355x.h"

/* Same values as wavefront_attributes.c:38-100. */
wavefront_aligner_attr_t wavefront_aligner_attr_default = {
  .distance_metric = gap_affine,
  .alignment_scope = compute_alignment,
  .alignment_form = {
    .span = alignment_end2end,
    .extension = false,
    .pattern_begin_free = 0,
    .pattern_end_free = 0,
    .text_begin_free = 0,
    .text_end_free = 0,
  },
  .linear_penalties = { .match = 0, .mismatch = 4, .indel = 2 },
  .affine_penalties = { .match = 0, .mismatch = 4, .gap_opening = 6, .gap_extension = 2 },
  .affine2p_penalties = {
    .match = 0, .mismatch = 4,
    .gap_opening1 = 6, .gap_extension1 = 2,
    .gap_opening2 = 24, .gap_extension2 = 1,
  },
  .heuristic = {
    .strategy = wf_heuristic_wfadaptive,
    .min_wavefront_length = 10,
    .max_distance_threshold = 50,
    .steps_between_cutoffs = 1,
  },
  .memory_mode = wavefront_memory_high,
  .mm_allocator = NULL,
  .plot = { .enabled = false, .resolution_points = 2000, .align_level = 0 },
  .system = {
    .max_alignment_steps = INT_MAX,
    .probe_interval_global = 3000,
    .probe_interval_compact = 6000,
    .max_memory_compact = (uint64_t)-1,
    .max_memory_resident = (uint64_t)-1,
    .max_memory_abort = UINT64_MAX,
    .verbose = 0,
    .check_alignment_correct = false,
    .max_num_threads = 1,
    .min_offsets_per_thread = 500,
  },
};

/* ----------------------------------------------------------- setters
 * wavefront_aligner.c:508-643: the form and heuristic members are plain
 * configuration; limits apply to the next alignment. */
void wavefront_aligner_set_alignment_end_to_end(wavefront_aligner_t* const wf) {
  wf->alignment_form.span = alignment_end2end;
  wf->alignment_form.extension = false;
}
void wavefront_aligner_set_alignment_free_ends(wavefront_aligner_t* const wf,
    const int pattern_begin_free, const int pattern_end_free,
    const int text_begin_free, const int text_end_free) {
  wf->alignment_form.span = alignment_endsfree;
  wf->alignment_form.extension = false;
  wf->alignment_form.pattern_begin_free = pattern_begin_free;
  wf->alignment_form.pattern_end_free = pattern_end_free;
  wf->alignment_form.text_begin_free = text_begin_free;
  wf->alignment_form.text_end_free = text_end_free;
}
void wavefront_aligner_set_alignment_extension(wavefront_aligner_t* const wf) {
  wf->alignment_form.span = alignment_endsfree;
  wf->alignment_form.extension = true;
}
/* The heuristic setters (wavefront_aligner.c:534-600 over
   wavefront_heuristic.c:40-113) add their strategy to the ones already set
   and overwrite the parameters they take (steps_between_cutoffs is shared);
   set_none clears the strategy only. */
void wavefront_aligner_set_heuristic_none(wavefront_aligner_t* const wf) {
  wf->heuristic.strategy = wf_heuristic_none;
}
void wavefront_aligner_set_heuristic_wfadaptive(wavefront_aligner_t* const wf,
    const int min_wavefront_length, const int max_distance_threshold, const int score_steps) {
  wf->heuristic.strategy |= wf_heuristic_wfadaptive;
  wf->heuristic.min_wavefront_length = min_wavefront_length;
  wf->heuristic.max_distance_threshold = max_distance_threshold;
  wf->heuristic.steps_between_cutoffs = score_steps;
  wf->heuristic.steps_wait = score_steps;
}
void wavefront_aligner_set_heuristic_wfmash(wavefront_aligner_t* const wf,
    const int min_wavefront_length, const int max_distance_threshold, const int score_steps) {
  wf->heuristic.strategy |= wf_heuristic_wfmash;
  wf->heuristic.min_wavefront_length = min_wavefront_length;
  wf->heuristic.max_distance_threshold = max_distance_threshold;
  wf->heuristic.steps_between_cutoffs = score_steps;
  wf->heuristic.steps_wait = score_steps;
}
void wavefront_aligner_set_heuristic_xdrop(wavefront_aligner_t* const wf,
    const int xdrop, const int score_steps) {
  wf->heuristic.strategy |= wf_heuristic_xdrop;
  wf->heuristic.xdrop = xdrop;
  wf->heuristic.steps_between_cutoffs = score_steps;
  wf->heuristic.steps_wait = score_steps;
}
void wavefront_aligner_set_heuristic_zdrop(wavefront_aligner_t* const wf,
    const int ydrop, const int score_steps) {
  wf->heuristic.strategy |= wf_heuristic_zdrop;
  wf->heuristic.zdrop = ydrop;
  wf->heuristic.steps_between_cutoffs = score_steps;
  wf->heuristic.steps_wait = score_steps;
}
void wavefront_aligner_set_heuristic_banded_static(wavefront_aligner_t* const wf,
    const int band_min_k, const int band_max_k) {
  wf->heuristic.strategy |= wf_heuristic_banded_static;
  wf->heuristic.min_k = band_min_k;
  wf->heuristic.max_k = band_max_k;
}
void wavefront_aligner_set_heuristic_banded_adaptive(wavefront_aligner_t* const wf,
    const int band_min_k, const int band_max_k, const int score_steps) {
  wf->heuristic.strategy |= wf_heuristic_banded_adaptive;
  wf->heuristic.min_k = band_min_k;
  wf->heuristic.max_k = band_max_k;
  wf->heuristic.steps_between_cutoffs = score_steps;
  wf->heuristic.steps_wait = score_steps;
}
void wavefront_aligner_set_max_alignment_steps(wavefront_aligner_t* const wf,
    const int max_alignment_steps) {
  wf->system.max_alignment_steps = max_alignment_steps;
}
void wavefront_aligner_set_max_memory(wavefront_aligner_t* const wf,
    const uint64_t max_memory_resident, const uint64_t max_memory_abort) {
  wf->system.max_memory_resident = max_memory_resident;
  wf->system.max_memory_abort = max_memory_abort;
}
void wavefront_aligner_set_max_num_threads(wavefront_aligner_t* const wf, const int max_num_threads) {
  wf->system.max_num_threads = max_num_threads;  /* accepted; the device path ignores it */
}
void wavefront_aligner_set_min_offsets_per_thread(wavefront_aligner_t* const wf,
    const int min_offsets_per_thread) {
  wf->system.min_offsets_per_thread = min_offsets_per_thread;
}

/* ----------------------------------------------------------- messages
 * Same texts as wavefront_aligner.c:45-98. */
char* wavefront_align_strerror(const int error_code) {
  switch (error_code) {
    case WF_STATUS_ALG_COMPLETED: return "[WFA] Alignment completed successfully";
    case WF_STATUS_ALG_PARTIAL: return "[WFA] Alignment end-to-end not found. Partial alignment found";
    case WF_STATUS_MAX_STEPS_REACHED: return "[WFA] Alignment failed. Maximum WFA-steps limit reached";
    case WF_STATUS_OOM: return "[WFA] Alignment failed. Maximum memory limit reached";
    case WF_STATUS_UNATTAINABLE: return "[WFA] Alignment failed. Unattainable under configured heuristics";
    case WF_STATUS_END_REACHED: return "[WFA] Alignment end reached";
    case WF_STATUS_END_UNREACHABLE:
      return "[WFA] Alignment end unreachable under current configuration (due to heuristics like Z-drop)";
    case WF_STATUS_UNSUPPORTED:
      return "[WFA] Configuration outside the device path (lambda sequences, or a form/heuristic/memory-mode "
             "combination the device declines)";
    default: return "[WFA] Unknown error code";
  }
}
char* wavefront_align_strerror_short(const int error_code) {
  switch (error_code) {
    case WF_STATUS_ALG_COMPLETED: return "OK.Complete";
    case WF_STATUS_ALG_PARTIAL: return "OK.Partial";
    case WF_STATUS_MAX_STEPS_REACHED: return "FAILED.MaxSteps";
    case WF_STATUS_OOM: return "FAILED.OOM";
    case WF_STATUS_UNATTAINABLE: return "FAILED.Unattainable";
    case WF_STATUS_END_REACHED: return "INTERNAL.Reached";
    case WF_STATUS_END_UNREACHABLE: return "INTERNAL.Dropped";
    case WF_STATUS_UNSUPPORTED: return "FAILED.Unsupported";
    default: return "Unknown";
  }
}

/* ----------------------------------------------------------- CIGAR
 * Scoring follows alignment/cigar.c:244-342; printing :436-470. */
bool cigar_is_null(const cigar_t* const cigar) {
  return cigar->begin_offset >= cigar->end_offset;
}

static void cigar_unknown_op(void) {
  fprintf(stderr, "[CIGAR] Computing CIGAR score: Unknown operation\n");
  exit(1);
}

/* Edit (and indel) distance: one per non-match operation. */
int cigar_score_edit(const cigar_t* const cigar) {
  int score = 0;
  for (int i = cigar->begin_offset; i < cigar->end_offset; ++i) {
    const char op = cigar->operations[i];
    if (op == 'X' || op == 'I' || op == 'D') ++score;
    else if (op != 'M') cigar_unknown_op();
  }
  return score;
}

int cigar_score_gap_linear(const cigar_t* const cigar, const linear_penalties_t* const pen) {
  int score = 0;
  for (int i = cigar->begin_offset; i < cigar->end_offset; ++i) {
    switch (cigar->operations[i]) {
      case 'M': score -= pen->match; break;
      case 'X': score -= pen->mismatch; break;
      case 'I':
      case 'D': score -= pen->indel; break;
      default: cigar_unknown_op();
    }
  }
  return score;
}

int cigar_score_gap_affine(const cigar_t* const cigar, const affine_penalties_t* const pen) {
  char last = '\0';
  int score = 0;
  for (int i = cigar->begin_offset; i < cigar->end_offset; ++i) {
    const char op = cigar->operations[i];
    switch (op) {
      case 'M': score -= pen->match; break;
      case 'X': score -= pen->mismatch; break;
      case 'D': score -= pen->gap_extension + ((last == 'D') ? 0 : pen->gap_opening); break;
      case 'I': score -= pen->gap_extension + ((last == 'I') ? 0 : pen->gap_opening); break;
      default:
        fprintf(stderr, "[CIGAR] Computing CIGAR score: Unknown operation\n");
        exit(1);
    }
    last = op;
  }
  return score;
}

static int affine2p_run_cost(char op, int len, const affine2p_penalties_t* const pen) {
  switch (op) {
    case 'M': return pen->match * len;
    case 'X': return pen->mismatch * len;
    case 'D':
    case 'I': {
      const int c1 = pen->gap_opening1 + pen->gap_extension1 * len;
      const int c2 = pen->gap_opening2 + pen->gap_extension2 * len;
      return c1 < c2 ? c1 : c2;
    }
    default:
      fprintf(stderr, "[CIGAR] Computing CIGAR score: Unknown operation\n");
      exit(1);
  }
}

int cigar_score_gap_affine2p(const cigar_t* const cigar, const affine2p_penalties_t* const pen) {
  int score = 0;
  int i = cigar->begin_offset;
  const int end = cigar->end_offset;
  if (i >= end) return -affine2p_run_cost('\0', 0, pen); /* reference: last op '\0' */
  while (i < end) {
    int j = i;
    while (j < end && cigar->operations[j] == cigar->operations[i]) ++j;
    score -= affine2p_run_cost(cigar->operations[i], j - i, pen);
    i = j;
  }
  return score;
}

bool cigar_check_alignment(FILE* const stream,
    const char* const pattern, const int pattern_length,
    const char* const text, const int text_length,
    const cigar_t* const cigar, const bool verbose) {
  int pp = 0, tp = 0;
  for (int i = cigar->begin_offset; i < cigar->end_offset; ++i) {
    switch (cigar->operations[i]) {
      case 'M':
        if (pattern[pp] != text[tp]) {
          if (verbose) fprintf(stream, "[CIGAR] Alignment not matching (pattern[%d]=%c != text[%d]=%c)\n",
                               pp, pattern[pp], tp, text[tp]);
          return false;
        }
        ++pp; ++tp; break;
      case 'X':
        if (pattern[pp] == text[tp]) {
          if (verbose) fprintf(stream, "[CIGAR] Alignment not mismatching (pattern[%d]=%c == text[%d]=%c)\n",
                               pp, pattern[pp], tp, text[tp]);
          return false;
        }
        ++pp; ++tp; break;
      case 'I': ++tp; break;
      case 'D': ++pp; break;
      default:
        fprintf(stream, "[CIGAR] Unknown edit operation '%c'\n", cigar->operations[i]);
        exit(1);
    }
  }
  if (pp != pattern_length) {
    if (verbose) fprintf(stream, "[CIGAR] Alignment incorrect length (pattern-aligned=%d,pattern-length=%d)\n",
                         pp, pattern_length);
    return false;
  }
  if (tp != text_length) {
    if (verbose) fprintf(stream, "[CIGAR] Alignment incorrect length (text-aligned=%d,text-length=%d)\n",
                         tp, text_length);
    return false;
  }
  return true;
}

int cigar_sprint(char* const buffer, const cigar_t* const cigar, const bool print_matches) {
  if (cigar_is_null(cigar)) { buffer[0] = '\0'; return 0; }
  const char* ops = cigar->operations;
  int cursor = 0;
  int i = cigar->begin_offset;
  while (i < cigar->end_offset) {
    int j = i;
    while (j < cigar->end_offset && ops[j] == ops[i]) ++j;
    if (print_matches || ops[i] != 'M') cursor += sprintf(buffer + cursor, "%d%c", j - i, ops[i]);
    i = j;
  }
  buffer[cursor] = '\0';
  return cursor;
}

/* SAM / BAM operation codes (alignment/cigar.c:38-56). */
static uint32_t sam_op_code(char op) {
  switch (op) {
    case 'M': return 0;
    case 'I': return 1;
    case 'D': return 2;
    case 'N': return 3;
    case '=': return 7;
    case 'X': return 8;
    default: return 15;
  }
}

/* Run-length SAM CIGAR, (length << 4 | code) per run (alignment/cigar.c:
   180-239): mismatches fold into 'M' unless show_mismatches, which writes
   matches as '='.  As the reference, the first operation is taken as it is
   (a leading 'X' stays 'X'), and the result is cached in the cigar until
   the next alignment or a different show_mismatches. */
void cigar_get_CIGAR(cigar_t* const cigar, const bool show_mismatches, uint32_t** const cigar_buffer,
    int* const cigar_length) {
  if (cigar->cigar_length == 0 || cigar->has_misms != show_mismatches) {
    const int n = cigar->end_offset - cigar->begin_offset;
    cigar->cigar_length = 0;
    if (n > 0) {
      uint32_t* const out = cigar->cigar_buffer;
      const char* const ops = cigar->operations + cigar->begin_offset;
      int len = 0;
      char last = ops[0];
      uint32_t run = 1;
      for (int i = 1; i <= n; ++i) {
        char op = 0;
        if (i < n) {
          op = ops[i];
          if (!show_mismatches && op == 'X') op = 'M';
          if (op == last) { ++run; continue; }
        }
        out[len++] = (run << 4) | ((show_mismatches && last == 'M') ? 7u : sam_op_code(last));
        last = op;
        run = 1;
      }
      cigar->has_misms = show_mismatches;
      cigar->cigar_length = len;
    }
  }
  *cigar_buffer = cigar->cigar_buffer;
  *cigar_length = cigar->cigar_length;
}

int cigar_sprint_SAM_CIGAR(char* const buffer, cigar_t* const cigar, const bool show_mismatches) {
  uint32_t* runs;
  int n;
  cigar_get_CIGAR(cigar, show_mismatches, &runs, &n);
  int cursor = 0;
  for (int i = 0; i < n; ++i) {
    const uint32_t code = runs[i] & 0xf;
    cursor += sprintf(buffer + cursor, "%u%c", runs[i] >> 4, code <= 8 ? "MIDN---=X"[code] : '?');
  }
  buffer[cursor] = '\0';
  return cursor;
}

void cigar_print_SAM_CIGAR(FILE* const stream, cigar_t* const cigar, const bool show_mismatches) {
  if (cigar_is_null(cigar)) return;
  char* const buffer = malloc(4 * (size_t)(cigar->end_offset - cigar->begin_offset) + 20);
  cigar_sprint_SAM_CIGAR(buffer, cigar, show_mismatches);
  fprintf(stream, "%s", buffer);
  free(buffer);
}

void cigar_print(FILE* const stream, const cigar_t* const cigar, const bool print_matches) {
  if (cigar_is_null(cigar)) return;
  char* const buffer = malloc(2 * (cigar->end_offset - cigar->begin_offset) + 20);
  cigar_sprint(buffer, cigar, print_matches);
  fprintf(stream, "%s", buffer);
  free(buffer);
}

/* Three-row view (pattern / match bars / text), as cigar_print_pretty. */
void cigar_print_pretty(FILE* const stream, const cigar_t* const cigar,
    const char* const pattern, const int pattern_length,
    const char* const text, const int text_length) {
  const int n = cigar->end_offset - cigar->begin_offset;
  if (n <= 0) return;
  char* rows = malloc(3 * (size_t)(n + 1));
  char* pr = rows;
  char* mr = rows + n + 1;
  char* tr = rows + 2 * (n + 1);
  int pp = 0, tp = 0;
  for (int i = 0; i < n; ++i) {
    const char op = cigar->operations[cigar->begin_offset + i];
    switch (op) {
      case 'M': pr[i] = pp < pattern_length ? pattern[pp++] : '?'; tr[i] = tp < text_length ? text[tp++] : '?'; mr[i] = '|'; break;
      case 'X': pr[i] = pp < pattern_length ? pattern[pp++] : '?'; tr[i] = tp < text_length ? text[tp++] : '?'; mr[i] = ' '; break;
      case 'I': pr[i] = '-'; tr[i] = tp < text_length ? text[tp++] : '?'; mr[i] = ' '; break;
      case 'D': pr[i] = pp < pattern_length ? pattern[pp++] : '?'; tr[i] = '-'; mr[i] = ' '; break;
      default: pr[i] = tr[i] = mr[i] = '?'; break;
    }
  }
  pr[n] = mr[n] = tr[n] = '\0';
  fprintf(stream, "      ALIGNMENT ");
  cigar_print(stream, cigar, true);
  fprintf(stream, "\n      PATTERN    %s\n", pr);
  fprintf(stream, "                 %s\n", mr);
  fprintf(stream, "      TEXT       %s\n", tr);
  free(rows);
}
