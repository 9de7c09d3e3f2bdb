/*
 * WFAligner.hpp -- C++ interface of the MI355X-native WFA library.
 *
 * The class interface of the reference's C++ binding
 * (bindings/cpp/WFAligner.hpp:37-262, implemented in WFAligner.cpp:40-470):
 * the same class names, constructors, enums, alignment methods, heuristic
 * setters and accessors, so that C++ callers switch by including this header
 * and linking libwfa2_amd.so.  Header-only, over the C ABI of wfa_mi355x.h.
 *
 * Added here: alignEnd2EndBatch / alignEndsFreeBatch / alignExtensionBatch,
 * which align many pairs in one device batch (wavefront_align_batch) with
 * results identical to calling the single-pair method on each pair.
 */
#ifndef WFA_MI355X_WFALIGNER_HPP_
#define WFA_MI355X_WFALIGNER_HPP_

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "wfa_mi355x.h"

namespace wfa {

class WFAligner {
 public:
  enum MemoryModel { MemoryHigh, MemoryMed, MemoryLow, MemoryUltralow };
  enum AlignmentScope { Score, Alignment };
  enum AlignmentStatus {
    StatusAlgCompleted = WF_STATUS_ALG_COMPLETED,
    StatusAlgPartial = WF_STATUS_ALG_PARTIAL,
    StatusMaxStepsReached = WF_STATUS_MAX_STEPS_REACHED,
    StatusOOM = WF_STATUS_OOM,
  };
  // One pair of a batch: status, score (as getAlignmentScore) and the
  // alignment operations (as getAlignment).
  struct BatchAlignment {
    AlignmentStatus status;
    int score;
    std::string alignment;
  };

  // Align end-to-end
  AlignmentStatus alignEnd2End(const char* const pattern, const int patternLength, const char* const text,
                               const int textLength) {
    wavefront_aligner_set_alignment_end_to_end(wfAligner);
    return (AlignmentStatus)wavefront_align(wfAligner, pattern, patternLength, text, textLength);
  }
  AlignmentStatus alignEnd2End(const std::string& pattern, const std::string& text) {
    return alignEnd2End(pattern.c_str(), (int)pattern.length(), text.c_str(), (int)text.length());
  }
  AlignmentStatus alignEnd2End(const uint8_t* const pattern, const int patternLength, const uint8_t* const text,
                               const int textLength) {
    wavefront_aligner_set_alignment_end_to_end(wfAligner);
    return (AlignmentStatus)wavefront_align_packed2bits(wfAligner, pattern, patternLength, text, textLength);
  }
  AlignmentStatus alignEnd2End(int (*matchFunct)(int, int, void*), void* matchFunctArguments,
                               const int patternLength, const int textLength) {
    wavefront_aligner_set_alignment_end_to_end(wfAligner);
    return (AlignmentStatus)wavefront_align_lambda(wfAligner, matchFunct, matchFunctArguments, patternLength,
                                                   textLength);
  }
  // Align ends-free
  AlignmentStatus alignEndsFree(const char* const pattern, const int patternLength, const int patternBeginFree,
                                const int patternEndFree, const char* const text, const int textLength,
                                const int textBeginFree, const int textEndFree) {
    wavefront_aligner_set_alignment_free_ends(wfAligner, patternBeginFree, patternEndFree, textBeginFree,
                                              textEndFree);
    return (AlignmentStatus)wavefront_align(wfAligner, pattern, patternLength, text, textLength);
  }
  AlignmentStatus alignEndsFree(const std::string& pattern, const int patternBeginFree, const int patternEndFree,
                                const std::string& text, const int textBeginFree, const int textEndFree) {
    return alignEndsFree(pattern.c_str(), (int)pattern.length(), patternBeginFree, patternEndFree, text.c_str(),
                         (int)text.length(), textBeginFree, textEndFree);
  }
  AlignmentStatus alignEndsFree(const uint8_t* const pattern, const int patternLength, const int patternBeginFree,
                                const int patternEndFree, const uint8_t* const text, const int textLength,
                                const int textBeginFree, const int textEndFree) {
    wavefront_aligner_set_alignment_free_ends(wfAligner, patternBeginFree, patternEndFree, textBeginFree,
                                              textEndFree);
    return (AlignmentStatus)wavefront_align_packed2bits(wfAligner, pattern, patternLength, text, textLength);
  }
  AlignmentStatus alignEndsFree(int (*matchFunct)(int, int, void*), void* matchFunctArguments,
                                const int patternLength, const int patternBeginFree, const int patternEndFree,
                                const int textLength, const int textBeginFree, const int textEndFree) {
    wavefront_aligner_set_alignment_free_ends(wfAligner, patternBeginFree, patternEndFree, textBeginFree,
                                              textEndFree);
    return (AlignmentStatus)wavefront_align_lambda(wfAligner, matchFunct, matchFunctArguments, patternLength,
                                                   textLength);
  }
  // Alignment extension
  AlignmentStatus alignExtension(const char* const pattern, const int patternLength, const char* const text,
                                 const int textLength) {
    wavefront_aligner_set_alignment_extension(wfAligner);
    return (AlignmentStatus)wavefront_align(wfAligner, pattern, patternLength, text, textLength);
  }
  AlignmentStatus alignExtension(std::string& pattern, std::string& text) {
    return alignExtension(pattern.c_str(), (int)pattern.length(), text.c_str(), (int)text.length());
  }
  AlignmentStatus alignExtension(const uint8_t* const pattern, const int patternLength, const uint8_t* const text,
                                 const int textLength) {
    wavefront_aligner_set_alignment_extension(wfAligner);
    return (AlignmentStatus)wavefront_align_packed2bits(wfAligner, pattern, patternLength, text, textLength);
  }
  AlignmentStatus alignExtension(int (*matchFunct)(int, int, void*), void* matchFunctArguments,
                                 const int patternLength, const int textLength) {
    wavefront_aligner_set_alignment_extension(wfAligner);
    return (AlignmentStatus)wavefront_align_lambda(wfAligner, matchFunct, matchFunctArguments, patternLength,
                                                   textLength);
  }
  // Batches (new): one device batch per call; returns the first failing
  // status (StatusAlgCompleted when every pair completed or was partial).
  AlignmentStatus alignEnd2EndBatch(const std::vector<std::string>& patterns, const std::vector<std::string>& texts,
                                    std::vector<BatchAlignment>& out) {
    wavefront_aligner_set_alignment_end_to_end(wfAligner);
    return alignBatch(patterns, texts, out);
  }
  AlignmentStatus alignEndsFreeBatch(const std::vector<std::string>& patterns, const int patternBeginFree,
                                     const int patternEndFree, const std::vector<std::string>& texts,
                                     const int textBeginFree, const int textEndFree,
                                     std::vector<BatchAlignment>& out) {
    wavefront_aligner_set_alignment_free_ends(wfAligner, patternBeginFree, patternEndFree, textBeginFree,
                                              textEndFree);
    return alignBatch(patterns, texts, out);
  }
  AlignmentStatus alignExtensionBatch(const std::vector<std::string>& patterns,
                                      const std::vector<std::string>& texts, std::vector<BatchAlignment>& out) {
    wavefront_aligner_set_alignment_extension(wfAligner);
    return alignBatch(patterns, texts, out);
  }
  // Heuristics
  void setHeuristicNone() { wavefront_aligner_set_heuristic_none(wfAligner); }
  void setHeuristicBandedStatic(const int band_min_k, const int band_max_k) {
    wavefront_aligner_set_heuristic_banded_static(wfAligner, band_min_k, band_max_k);
  }
  void setHeuristicBandedAdaptive(const int band_min_k, const int band_max_k, const int steps_between_cutoffs = 1) {
    wavefront_aligner_set_heuristic_banded_adaptive(wfAligner, band_min_k, band_max_k, steps_between_cutoffs);
  }
  void setHeuristicWFadaptive(const int min_wavefront_length, const int max_distance_threshold,
                              const int steps_between_cutoffs = 1) {
    wavefront_aligner_set_heuristic_wfadaptive(wfAligner, min_wavefront_length, max_distance_threshold,
                                               steps_between_cutoffs);
  }
  void setHeuristicWFmash(const int min_wavefront_length, const int max_distance_threshold,
                          const int steps_between_cutoffs = 1) {
    wavefront_aligner_set_heuristic_wfmash(wfAligner, min_wavefront_length, max_distance_threshold,
                                           steps_between_cutoffs);
  }
  void setHeuristicXDrop(const int xdrop, const int steps_between_cutoffs = 1) {
    wavefront_aligner_set_heuristic_xdrop(wfAligner, xdrop, steps_between_cutoffs);
  }
  void setHeuristicZDrop(const int zdrop, const int steps_between_cutoffs = 1) {
    wavefront_aligner_set_heuristic_zdrop(wfAligner, zdrop, steps_between_cutoffs);
  }
  // Limits
  void setMaxAlignmentSteps(const int maxAlignmentSteps) {
    wavefront_aligner_set_max_alignment_steps(wfAligner, maxAlignmentSteps);
  }
  void setMaxMemory(const uint64_t maxMemoryResident, const uint64_t maxMemoryAbort) {
    wavefront_aligner_set_max_memory(wfAligner, maxMemoryResident, maxMemoryAbort);
  }
  void setMaxNumThreads(const int maxNumThreads) { wavefront_aligner_set_max_num_threads(wfAligner, maxNumThreads); }
  // Accessors
  int getAlignmentStatus() { return wfAligner->align_status.status; }
  int getAlignmentScore() { return wfAligner->cigar->score; }
  void getAlignment(char** const cigarOperations, int* cigarLength) {
    *cigarOperations = wfAligner->cigar->operations + wfAligner->cigar->begin_offset;
    *cigarLength = wfAligner->cigar->end_offset - wfAligner->cigar->begin_offset;
  }
  std::string getAlignment() {
    const cigar_t* c = wfAligner->cigar;
    return c->end_offset > c->begin_offset ? std::string(c->operations + c->begin_offset, c->end_offset - c->begin_offset)
                                           : std::string();
  }
  void getCIGAR(const bool showMismatches, uint32_t** const cigarOperations, int* const numCigarOperations) {
    cigar_get_CIGAR(wfAligner->cigar, showMismatches, cigarOperations, numCigarOperations);
  }
  std::string getCIGAR(const bool showMismatches) {
    const int n = wfAligner->cigar->end_offset - wfAligner->cigar->begin_offset;
    if (n <= 0) return std::string();
    std::vector<char> buffer(4 * (size_t)n + 16);
    const int len = cigar_sprint_SAM_CIGAR(buffer.data(), wfAligner->cigar, showMismatches);
    return std::string(buffer.data(), len);
  }
  // Display
  void printPretty(FILE* const stream, const char* const pattern, const int patternLength, const char* const text,
                   const int textLength) {
    cigar_print_pretty(stream, wfAligner->cigar, pattern, patternLength, text, textLength);
  }
  // Misc
  char* strStatus(const AlignmentStatus status) { return wavefront_align_strerror((int)status); }
  void debugTag(char* const debugTag) { wfAligner->align_mode_tag = debugTag; }

  virtual ~WFAligner() { wavefront_aligner_delete(wfAligner); }

 protected:
  wavefront_aligner_attr_t attributes;
  wavefront_aligner_t* wfAligner;

  WFAligner(const AlignmentScope alignmentScope, const MemoryModel memoryModel = MemoryHigh) {
    attributes = wavefront_aligner_attr_default;
    switch (memoryModel) {
      case MemoryMed: attributes.memory_mode = wavefront_memory_med; break;
      case MemoryLow: attributes.memory_mode = wavefront_memory_low; break;
      case MemoryUltralow: attributes.memory_mode = wavefront_memory_ultralow; break;
      default: attributes.memory_mode = wavefront_memory_high; break;
    }
    attributes.alignment_scope = (alignmentScope == Score) ? compute_score : compute_alignment;
    wfAligner = nullptr;
  }

 private:
  WFAligner(const WFAligner&);

  AlignmentStatus alignBatch(const std::vector<std::string>& patterns, const std::vector<std::string>& texts,
                             std::vector<BatchAlignment>& out) {
    const int n = (int)patterns.size();
    std::vector<wfa_pair_t> pairs(n);
    for (int i = 0; i < n; ++i)
      pairs[i] = wfa_pair_t{patterns[i].data(), (int)patterns[i].size(), texts[i].data(), (int)texts[i].size()};
    std::vector<wfa_result_t> res(n);
    const size_t bytes = wavefront_align_batch_arena_bytes(pairs.data(), n);
    std::vector<char> arena(bytes > 0 ? bytes : 1);
    const int rc = wavefront_align_batch(wfAligner, pairs.data(), n, res.data(), arena.data(), arena.size());
    out.resize(n);
    for (int i = 0; i < n; ++i) {
      out[i].status = (AlignmentStatus)res[i].status;
      out[i].score = res[i].score;
      out[i].alignment.assign(arena.data() + res[i].cigar_offset, (size_t)res[i].cigar_length);
    }
    return (AlignmentStatus)rc;
  }
};

// Indel aligner (longest common subsequence)
class WFAlignerIndel : public WFAligner {
 public:
  WFAlignerIndel(const AlignmentScope alignmentScope, const MemoryModel memoryModel = MemoryHigh)
      : WFAligner(alignmentScope, memoryModel) {
    attributes.distance_metric = indel;
    wfAligner = wavefront_aligner_new(&attributes);
  }
};

// Edit aligner (Levenshtein)
class WFAlignerEdit : public WFAligner {
 public:
  WFAlignerEdit(const AlignmentScope alignmentScope, const MemoryModel memoryModel = MemoryHigh)
      : WFAligner(alignmentScope, memoryModel) {
    attributes.distance_metric = edit;
    wfAligner = wavefront_aligner_new(&attributes);
  }
};

// Gap-linear aligner (Needleman-Wunsch)
class WFAlignerGapLinear : public WFAligner {
 public:
  WFAlignerGapLinear(const int mismatch, const int indel, const AlignmentScope alignmentScope,
                     const MemoryModel memoryModel = MemoryHigh)
      : WFAlignerGapLinear(0, mismatch, indel, alignmentScope, memoryModel) {}
  WFAlignerGapLinear(const int match, const int mismatch, const int indel, const AlignmentScope alignmentScope,
                     const MemoryModel memoryModel = MemoryHigh)
      : WFAligner(alignmentScope, memoryModel) {
    attributes.distance_metric = gap_linear;
    attributes.linear_penalties.match = match;
    attributes.linear_penalties.mismatch = mismatch;
    attributes.linear_penalties.indel = indel;
    wfAligner = wavefront_aligner_new(&attributes);
  }
};

// Gap-affine aligner (Smith-Waterman-Gotoh)
class WFAlignerGapAffine : public WFAligner {
 public:
  WFAlignerGapAffine(const int mismatch, const int gapOpening, const int gapExtension,
                     const AlignmentScope alignmentScope, const MemoryModel memoryModel = MemoryHigh)
      : WFAlignerGapAffine(0, mismatch, gapOpening, gapExtension, alignmentScope, memoryModel) {}
  WFAlignerGapAffine(const int match, const int mismatch, const int gapOpening, const int gapExtension,
                     const AlignmentScope alignmentScope, const MemoryModel memoryModel = MemoryHigh)
      : WFAligner(alignmentScope, memoryModel) {
    attributes.distance_metric = gap_affine;
    attributes.affine_penalties.match = match;
    attributes.affine_penalties.mismatch = mismatch;
    attributes.affine_penalties.gap_opening = gapOpening;
    attributes.affine_penalties.gap_extension = gapExtension;
    wfAligner = wavefront_aligner_new(&attributes);
  }
};

// Dual-cost gap-affine aligner (concave, two pieces)
class WFAlignerGapAffine2Pieces : public WFAligner {
 public:
  WFAlignerGapAffine2Pieces(const int mismatch, const int gapOpening1, const int gapExtension1,
                            const int gapOpening2, const int gapExtension2, const AlignmentScope alignmentScope,
                            const MemoryModel memoryModel = MemoryHigh)
      : WFAlignerGapAffine2Pieces(0, mismatch, gapOpening1, gapExtension1, gapOpening2, gapExtension2,
                                  alignmentScope, memoryModel) {}
  WFAlignerGapAffine2Pieces(const int match, const int mismatch, const int gapOpening1, const int gapExtension1,
                            const int gapOpening2, const int gapExtension2, const AlignmentScope alignmentScope,
                            const MemoryModel memoryModel = MemoryHigh)
      : WFAligner(alignmentScope, memoryModel) {
    attributes.distance_metric = gap_affine_2p;
    attributes.affine2p_penalties.match = match;
    attributes.affine2p_penalties.mismatch = mismatch;
    attributes.affine2p_penalties.gap_opening1 = gapOpening1;
    attributes.affine2p_penalties.gap_extension1 = gapExtension1;
    attributes.affine2p_penalties.gap_opening2 = gapOpening2;
    attributes.affine2p_penalties.gap_extension2 = gapExtension2;
    wfAligner = wavefront_aligner_new(&attributes);
  }
};

}  // namespace wfa

#endif  // WFA_MI355X_WFALIGNER_HPP_
