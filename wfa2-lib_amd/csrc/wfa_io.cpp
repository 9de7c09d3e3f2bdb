// Streaming .seq -> .alg alignment (the input and output formats of the
// reference's align_benchmark tool): a reader thread parses line pairs into
// batches, the calling thread aligns each batch on the device with
// wavefront_align_batch, and a writer thread formats the result lines, so
// parsing, device work and formatting of consecutive batches overlap.
//
// Input (tools/align_benchmark/align_benchmark.c:313-339): pairs of lines,
// the first character of each line dropped ('>' pattern, '<' text) and one
// trailing '\n' removed.  Output (benchmark_utils.c:103-179): per pair
// "score\tCIGAR", or "plen\ttlen\tscore\tpattern\ttext\tCIGAR" in the full
// form; the score is the CIGAR re-scored under the metric (the aligner's
// cigar->score when it computes scores only), -1 without a CIGAR; the CIGAR
// is run-length with matches, "-" when empty.
#include <algorithm>
#include <chrono>
#include <functional>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "wfa_internal.h"

namespace {

double now_s() {
  using clock = std::chrono::steady_clock;
  return std::chrono::duration<double>(clock::now().time_since_epoch()).count();
}

struct Batch {
  std::string seqs;                 // pattern and text bytes of every pair
  std::vector<int64_t> poff, toff;
  std::vector<int> plen, tlen;
  std::vector<wfa_result_t> res;
  std::vector<char> arena;          // CIGAR operations (byte form)
  std::vector<uint32_t> runs;       // or run-length CIGARs made on the device
  bool by_runs = false;             // res[i].cigar_offset / cigar_length index runs
  int n() const { return (int)plen.size(); }
};

// Bounded queue between two pipeline stages (nullptr ends the stream).
class Channel {
 public:
  explicit Channel(size_t depth) : depth_(depth) {}
  void put(std::unique_ptr<Batch> b) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return q_.size() < depth_; });
    q_.push_back(std::move(b));
    cv_.notify_all();
  }
  std::unique_ptr<Batch> get() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return !q_.empty(); });
    std::unique_ptr<Batch> b = std::move(q_.front());
    q_.pop_front();
    cv_.notify_all();
    return b;
  }

 private:
  size_t depth_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<std::unique_ptr<Batch>> q_;
};

// Lines of a FILE read through a large buffer.  A line is returned as a view
// into the buffer (with its '\n'), valid until the next call; one that
// straddles two reads is assembled in a side string.
class LineReader {
 public:
  explicit LineReader(FILE* f) : f_(f), buf_(size_t(1) << 22) {}
  bool next(const char** p, size_t* len) {
    if (carry_done_) {
      carry_.clear();
      carry_done_ = false;
    }
    for (;;) {
      if (pos_ < end_) {
        const char* b = buf_.data() + pos_;
        const void* nl = memchr(b, '\n', end_ - pos_);
        if (nl) {
          const size_t take = (size_t)((const char*)nl - b) + 1;
          pos_ += take;
          if (carry_.empty()) {
            *p = b;
            *len = take;
            return true;
          }
          carry_.append(b, take);
          return give_carry(p, len);
        }
        carry_.append(b, end_ - pos_);
        pos_ = end_;
      }
      end_ = fread(buf_.data(), 1, buf_.size(), f_);
      pos_ = 0;
      bytes_ += end_;
      if (end_ == 0) return carry_.empty() ? false : give_carry(p, len);
    }
  }
  int64_t bytes() const { return bytes_; }

 private:
  bool give_carry(const char** p, size_t* len) {
    *p = carry_.data();
    *len = carry_.size();
    carry_done_ = true;
    return true;
  }
  FILE* f_;
  std::vector<char> buf_;
  size_t pos_ = 0, end_ = 0;
  int64_t bytes_ = 0;
  std::string carry_;
  bool carry_done_ = false;
};

// A line of the .seq format: drop the tag character and one trailing '\n'.
void strip(const char* line, size_t size, const char** p, int* len) {
  int n = (int)size - 1;
  if (size > 0 && line[size - 1] == '\n') --n;
  *p = line + (size == 0 ? 0 : 1);
  *len = n > 0 ? n : 0;
}

// cigar_score_* of the metric (alignment/cigar.c:244-342) with the user penalties.
int cigar_score_of(const wavefront_penalties_t& p, const char* ops, int n) {
  cigar_t c{};
  c.operations = const_cast<char*>(ops);
  c.begin_offset = 0;
  c.end_offset = n;
  switch (p.distance_metric) {
    case indel:
    case edit: return cigar_score_edit(&c);
    case gap_linear: {
      linear_penalties_t lp = p.linear_penalties;
      return cigar_score_gap_linear(&c, &lp);
    }
    case gap_affine: {
      affine_penalties_t ap = p.affine_penalties;
      return cigar_score_gap_affine(&c, &ap);
    }
    default: {
      affine2p_penalties_t ap = p.affine2p_penalties;
      return cigar_score_gap_affine2p(&c, &ap);
    }
  }
}

// The same scores over run-length CIGARs ((length << 4) | SAM code: 7 '=',
// 8 X, 1 I, 2 D): a run of one gap type opens once, as the byte loops of
// cigar_score_gap_affine see it (no two runs of one type are adjacent).
int runs_score_of(const wavefront_penalties_t& p, const uint32_t* runs, int n) {
  long long score = 0;
  for (int i = 0; i < n; ++i) {
    const int len = (int)(runs[i] >> 4);
    const uint32_t c = runs[i] & 15u;
    const bool match = c == 7u, mism = c == 8u, gap = c == 1u || c == 2u;
    switch (p.distance_metric) {
      case indel:
      case edit:
        if (!match) score += len;
        break;
      case gap_linear:
        score -= (long long)len * (match ? p.linear_penalties.match
                                         : mism ? p.linear_penalties.mismatch : p.linear_penalties.indel);
        break;
      case gap_affine: {
        const affine_penalties_t& a = p.affine_penalties;
        if (match) score -= (long long)a.match * len;
        else if (mism) score -= (long long)a.mismatch * len;
        else if (gap) score -= a.gap_opening + (long long)a.gap_extension * len;
        break;
      }
      default: {
        const affine2p_penalties_t& a = p.affine2p_penalties;
        if (match) score -= (long long)a.match * len;
        else if (mism) score -= (long long)a.mismatch * len;
        else if (gap) score -= std::min((long long)a.gap_opening1 + (long long)a.gap_extension1 * len,
                                        (long long)a.gap_opening2 + (long long)a.gap_extension2 * len);
      }
    }
  }
  return (int)score;
}

void append_int(std::string& out, long long v) {
  char tmp[24];
  int n = 0;
  const bool neg = v < 0;
  unsigned long long u = neg ? 0ull - (unsigned long long)v : (unsigned long long)v;
  do {
    tmp[23 - n++] = (char)('0' + u % 10);
    u /= 10;
  } while (u);
  if (neg) tmp[23 - n++] = '-';
  out.append(tmp + 24 - n, n);
}

// The .alg lines of pairs [lo, hi) of a batch.
void format_lines(const Batch& b, int lo, int hi, bool score_only, bool full_output, const wavefront_penalties_t& pen,
                  std::string& text) {
  text.clear();
  static const char kRunOp[16] = {'M', 'I', 'D', 'N', '?', '?', '?', 'M', 'X', '?', '?', '?', '?', '?', '?', '?'};
  for (int i = lo; i < hi; ++i) {
    const wfa_result_t& r = b.res[i];
    const bool has_cigar = !score_only && r.cigar_length > 0;
    const char* ops = b.by_runs ? nullptr : b.arena.data() + (has_cigar ? r.cigar_offset : 0);
    const uint32_t* runs = b.by_runs ? b.runs.data() + (has_cigar ? r.cigar_offset : 0) : nullptr;
    const int score = score_only ? r.score
                                 : (!has_cigar ? -1
                                               : (b.by_runs ? runs_score_of(pen, runs, r.cigar_length)
                                                            : cigar_score_of(pen, ops, r.cigar_length)));
    if (full_output) {
      append_int(text, b.plen[i]);
      text += '\t';
      append_int(text, b.tlen[i]);
      text += '\t';
    }
    append_int(text, score);
    text += '\t';
    if (full_output) {
      text.append(b.seqs.data() + b.poff[i], b.plen[i]);
      text += '\t';
      text.append(b.seqs.data() + b.toff[i], b.tlen[i]);
      text += '\t';
    }
    if (!has_cigar) {
      text += '-';
    } else if (b.by_runs) {
      for (int j = 0; j < r.cigar_length; ++j) {
        append_int(text, (long long)(runs[j] >> 4));
        text += kRunOp[runs[j] & 15u];
      }
    } else {
      int j = 0;
      while (j < r.cigar_length) {
        int k = j;
        while (k < r.cigar_length && ops[k] == ops[j]) ++k;
        append_int(text, k - j);
        text += ops[j];
        j = k;
      }
    }
    text += '\n';
  }
}

}  // namespace

extern "C" int wfa_align_seq_file(wavefront_aligner_t* const wf, const char* in_path, const char* out_path,
                                  int batch_pairs, int full_output, wfa_io_stats_t* stats) {
  if (!wf || !in_path || !out_path) return WF_STATUS_UNATTAINABLE;
  FILE* in = fopen(in_path, "rb");
  if (!in) {
    fprintf(stderr, "[WFA] cannot open %s\n", in_path);
    return WF_STATUS_UNATTAINABLE;
  }
  FILE* out = fopen(out_path, "wb");
  if (!out) {
    fprintf(stderr, "[WFA] cannot create %s\n", out_path);
    fclose(in);
    return WF_STATUS_UNATTAINABLE;
  }
  if (batch_pairs <= 0) batch_pairs = 1 << 16;
  const bool score_only = wf->alignment_scope == compute_score;
  const wavefront_penalties_t pen = wf->penalties;
  const double t0 = now_s();
  Channel to_align(2), to_write(2);
  int64_t pairs = 0, bytes_out = 0, bytes_in = 0;
  double read_s = 0, align_s = 0, write_s = 0;

  std::thread reader([&] {
    LineReader lr(in);
    bool eof = false;
    for (;;) {
      const double t = now_s();
      auto b = std::make_unique<Batch>();
      b->seqs.reserve(size_t(8) << 20);
      b->poff.reserve(batch_pairs);
      b->toff.reserve(batch_pairs);
      b->plen.reserve(batch_pairs);
      b->tlen.reserve(batch_pairs);
      while (!eof && b->n() < batch_pairs && b->seqs.size() < (size_t(64) << 20)) {
        // each line is copied out before the next is read (views die then)
        const char* l;
        size_t ll;
        const char* p;
        int pl;
        if (!lr.next(&l, &ll)) {
          eof = true;
          break;
        }
        strip(l, ll, &p, &pl);
        const size_t at = b->seqs.size();
        b->seqs.append(p, pl);
        if (!lr.next(&l, &ll)) {  // a pattern without its text is dropped
          b->seqs.resize(at);
          eof = true;
          break;
        }
        const char* q;
        int ql;
        strip(l, ll, &q, &ql);
        b->poff.push_back((int64_t)at);
        b->plen.push_back(pl);
        b->toff.push_back((int64_t)b->seqs.size());
        b->tlen.push_back(ql);
        b->seqs.append(q, ql);
      }
      read_s += now_s() - t;
      const bool last = b->n() == 0;
      to_align.put(last ? nullptr : std::move(b));
      if (last) break;
    }
    bytes_in = lr.bytes();
  });

  // The writer formats each batch in parallel slices (formatting is the
  // costly part of the output) and writes the slices in order.
  const int nfmt = std::max(1, std::min(8, (int)std::thread::hardware_concurrency() / 2));
  std::thread writer([&] {
    std::vector<std::string> parts(nfmt);
    for (;;) {
      std::unique_ptr<Batch> b = to_write.get();
      if (!b) break;
      const double t = now_s();
      const int n = b->n();
      const int step = (n + nfmt - 1) / nfmt;
      std::vector<std::thread> fmt;
      for (int w = 1; w < nfmt && w * step < n; ++w)
        fmt.emplace_back(format_lines, std::cref(*b), w * step, std::min(n, (w + 1) * step), score_only,
                         full_output != 0, std::cref(pen), std::ref(parts[w]));
      format_lines(*b, 0, std::min(n, step), score_only, full_output != 0, pen, parts[0]);
      for (auto& th : fmt) th.join();
      for (int w = 0; w < nfmt && w * step < n; ++w) {
        fwrite(parts[w].data(), 1, parts[w].size(), out);
        bytes_out += (int64_t)parts[w].size();
      }
      write_s += now_s() - t;
    }
  });

  // CIGARs come back run-length (made on the device) except where the host
  // trims them (extension form, z-drop).
  const bool by_runs = !score_only && !wf->alignment_form.extension &&
                       !((wf->heuristic.strategy & wf_heuristic_zdrop) && wf->memory_mode != wavefront_memory_ultralow);
  wfa_host::Engine* e = wfa_host::engine_of(wf);
  std::unique_ptr<wfa_device_batch_t> db;
  int first_fail = WF_STATUS_ALG_COMPLETED;
  for (;;) {
    std::unique_ptr<Batch> b = to_align.get();
    if (!b) break;
    const double t = now_s();
    const int n = b->n();
    std::vector<wfa_pair_t> pr(n);
    size_t ops = 0;
    for (int i = 0; i < n; ++i) {
      pr[i] = wfa_pair_t{b->seqs.data() + b->poff[i], b->plen[i], b->seqs.data() + b->toff[i], b->tlen[i]};
      ops += (size_t)b->plen[i] + b->tlen[i];
    }
    b->res.resize(n);
    int rc;
    if (by_runs && e && wfa_host::engine_init(e)) {
      if (!db) db.reset(new (std::nothrow) wfa_device_batch_t());
      rc = WF_STATUS_OOM;
      if (db && db->load(e, pr.data(), n, true)) {
        rc = wfa_device_batch_run(wf, db.get());
        if (db->have_results) {
          const int64_t total = wfa_device_batch_cigar_runs(wf, db.get());
          if (total >= 0) {
            b->runs.resize((size_t)std::max<int64_t>(total, 1));
            rc = wfa_device_batch_download_runs(wf, db.get(), b->res.data(), b->runs.data(), b->runs.size());
            b->by_runs = true;
          } else {
            rc = (int)total;
          }
        }
      }
      if (!b->by_runs)
        for (auto& r : b->res) r = wfa_result_t{rc, INT32_MIN, -1, -1, 0, 0};
    } else {
      b->arena.resize(std::max<size_t>(ops, 1));
      rc = wavefront_align_batch(wf, pr.data(), n, b->res.data(), score_only ? nullptr : b->arena.data(),
                                 score_only ? 0 : b->arena.size());
    }
    if (rc != WF_STATUS_ALG_COMPLETED && first_fail == WF_STATUS_ALG_COMPLETED) first_fail = rc;
    pairs += n;
    align_s += now_s() - t;
    to_write.put(std::move(b));
  }
  to_write.put(nullptr);
  reader.join();
  writer.join();
  fclose(in);
  const bool write_ok = fclose(out) == 0;
  if (stats) {
    stats->pairs = pairs;
    stats->bytes_in = bytes_in;
    stats->bytes_out = bytes_out;
    stats->seconds = now_s() - t0;
    stats->read_seconds = read_s;
    stats->align_seconds = align_s;
    stats->write_seconds = write_s;
  }
  if (!write_ok) {
    fprintf(stderr, "[WFA] cannot write %s\n", out_path);
    return WF_STATUS_UNATTAINABLE;
  }
  return first_fail;
}
