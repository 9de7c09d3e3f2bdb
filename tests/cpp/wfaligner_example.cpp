// A C++ caller in the style of the reference's C++ binding examples
// (bindings/cpp/WFAligner.hpp), compiled against include/WFAligner.hpp and
// linked with libwfa2_amd.so.
//
// argv[1] selects the aligner and form: "affine" (WFAlignerGapAffine 4,6,2
// with the default wf-adaptive heuristic, end-to-end), "affine-exact" (the
// same after setHeuristicNone), "endsfree" (affine, exact, text ends free up
// to 10), "extension" (affine, exact), "edit" (WFAlignerEdit, exact).
// argv[2] == "batch" aligns all pairs with one batch call instead of one call
// per pair.  stdin: "pattern<TAB>text" lines.  stdout, per pair:
// "status<TAB>score<TAB>SAM CIGAR<TAB>SAM CIGAR with =/X<TAB>operations".
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "WFAligner.hpp"

using wfa::WFAligner;

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "affine";
  const bool batch = argc > 2 && std::string(argv[2]) == "batch";
  std::unique_ptr<WFAligner> al;
  if (mode == "edit") {
    al.reset(new wfa::WFAlignerEdit(WFAligner::Alignment, WFAligner::MemoryHigh));
  } else {
    al.reset(new wfa::WFAlignerGapAffine(4, 6, 2, WFAligner::Alignment, WFAligner::MemoryHigh));
  }
  if (mode != "affine") al->setHeuristicNone();
  std::vector<std::string> pats, txts;
  std::string line;
  while (std::getline(std::cin, line)) {
    const size_t tab = line.find('\t');
    if (tab == std::string::npos) continue;
    pats.push_back(line.substr(0, tab));
    txts.push_back(line.substr(tab + 1));
  }
  if (batch) {
    std::vector<WFAligner::BatchAlignment> out;
    if (mode == "endsfree") al->alignEndsFreeBatch(pats, 0, 0, txts, 10, 10, out);
    else if (mode == "extension") al->alignExtensionBatch(pats, txts, out);
    else al->alignEnd2EndBatch(pats, txts, out);
    for (const auto& r : out) std::cout << r.status << '\t' << r.score << "\t\t\t" << r.alignment << '\n';
    return 0;
  }
  for (size_t i = 0; i < pats.size(); ++i) {
    WFAligner::AlignmentStatus st;
    if (mode == "endsfree") st = al->alignEndsFree(pats[i], 0, 0, txts[i], 10, 10);
    else if (mode == "extension") st = al->alignExtension(pats[i], txts[i]);
    else st = al->alignEnd2End(pats[i], txts[i]);
    std::cout << st << '\t' << al->getAlignmentScore() << '\t' << al->getCIGAR(false) << '\t'
              << al->getCIGAR(true) << '\t' << al->getAlignment() << '\n';
  }
  return 0;
}
