"""Host-side checks of two pieces whose edge cases the GPU tests reach only
by chance, compiled with g++ from the text of their sources:

* the CIGAR chunk writers of the backtrace kernel (csrc/wfa_kernels.hip,
  OpChunk / OpWriter / OpWriterBack): random sequences of single operations
  and runs, forward and backward, at every alignment of the slot, must write
  exactly the bytes of a plain byte-by-byte writer and nothing outside the
  slot;
* the line reader of the .seq streamer (csrc/wfa_io.cpp, LineReader and
  strip): lines longer than its 4 MB buffer, lines that straddle two reads
  and a last line without '\\n' must come back whole."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "wfa2-lib_amd", "csrc")
GXX = shutil.which("g++")


def between(text, start, end_marker):
    a = text.index(start)
    b = text.index(end_marker, a)
    return text[a:b]


def build_and_run(tmp_path, name, source):
    src = tmp_path / (name + ".cpp")
    exe = tmp_path / name
    src.write_text(source)
    subprocess.check_call([GXX, "-O2", "-std=c++17", str(src), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok"), out.stdout


WRITERS_MAIN = r"""
int main() {
  std::mt19937 rng(11);
  alignas(16) static unsigned char buf[4096], ref[4096];
  for (int t = 0; t < 100000; ++t) {
    // forward: a slot from w0 at any alignment
    memset(buf, 0xEE, sizeof buf);
    memset(ref, 0xEE, sizeof ref);
    unsigned char* o = buf + t % 16;
    unsigned char* r = ref + t % 16;
    const int w0 = 64 + (int)(rng() % 64);
    OpWriter ow{o, w0, w0, OpChunk{}};
    int w = w0;
    const int nops = (int)(rng() % 40);
    for (int i = 0; i < nops; ++i) {
      const unsigned c = "MXID"[rng() % 4];
      if (rng() % 2) {
        ow.put(c);
        r[w++] = (unsigned char)c;
      } else {
        const int m = (int)(rng() % 50);
        ow.run(c, m);
        for (int j = 0; j < m; ++j) r[w++] = (unsigned char)c;
      }
      if (ow.w != w) { printf("forward position\n"); return 1; }
    }
    ow.finish();
    if (memcmp(buf, ref, sizeof buf)) { printf("forward bytes, case %d\n", t); return 1; }
    // backward: below top at any alignment (bytes from top up untouched)
    memset(buf, 0xEE, sizeof buf);
    memset(ref, 0xEE, sizeof ref);
    const int top = 2900 + (int)(rng() % 64);
    OpWriterBack bw{o, top, top, OpChunk{}};
    int p = top;
    const int n = (int)(rng() % 200);
    for (int i = 0; i < n; ++i) {
      const unsigned c = (unsigned)("XID"[rng() % 3] | (rng() % 2 ? 0x80 : 0));
      bw.put(c);
      r[--p] = (unsigned char)c;
      if (bw.p != p) { printf("backward position\n"); return 1; }
    }
    bw.finish();
    if (memcmp(buf, ref, sizeof buf)) { printf("backward bytes, case %d\n", t); return 1; }
  }
  printf("ok\n");
  return 0;
}
"""


@pytest.mark.skipif(GXX is None, reason="g++ not available")
def test_cigar_chunk_writers_write_exactly_their_bytes(tmp_path):
    text = open(os.path.join(CSRC, "wfa_kernels.hip")).read()
    writers = between(text, "struct OpChunk {", "// One lane per pair, in two passes")
    prelude = "\n".join([
        "#include <algorithm>", "#include <cstdint>", "#include <cstdio>", "#include <cstring>",
        "#include <random>",
        "#define __device__", "#define __forceinline__ inline",
        "using std::max; using std::min;",
        "struct uint4 { uint32_t x, y, z, w; };",
        "static uint4 make_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return {a, b, c, d}; }",
    ])
    build_and_run(tmp_path, "writers", prelude + "\n" + writers + "\n" + WRITERS_MAIN)


READER_MAIN = r"""
int main() {
  std::mt19937 rng(5);
  for (int trial = 0; trial < 6; ++trial) {
    std::string content;
    std::vector<std::string> want;
    const size_t target = trial < 3 ? (size_t)9 << 20 : 1000;
    while (content.size() < target) {
      const int len = rng() % 7 == 0 ? (int)(rng() % 3000000) : (int)(rng() % 400);
      want.push_back(std::string(len, (char)('A' + rng() % 4)));
      content += want.back();
      content += '\n';
    }
    const bool no_final_newline = trial % 2 == 1;
    if (no_final_newline) content.pop_back();
    FILE* f = tmpfile();
    fwrite(content.data(), 1, content.size(), f);
    rewind(f);
    LineReader lr(f);
    const char* l;
    size_t ll;
    size_t i = 0;
    while (lr.next(&l, &ll)) {
      const bool last = i + 1 == want.size();
      const std::string expect = want[i] + (last && no_final_newline ? "" : "\n");
      if (std::string(l, ll) != expect) { printf("trial %d line %zu\n", trial, i); return 1; }
      const char* p;
      int pl;
      strip(l, ll, &p, &pl);
      const int keep = want[i].empty() ? 0 : (int)want[i].size() - 1;
      if (pl != keep || (pl > 0 && memcmp(p, want[i].data() + 1, pl))) {
        printf("trial %d strip %zu\n", trial, i);
        return 1;
      }
      ++i;
    }
    if (i != want.size()) { printf("trial %d: %zu of %zu lines\n", trial, i, want.size()); return 1; }
    fclose(f);
  }
  printf("ok\n");
  return 0;
}
"""


@pytest.mark.skipif(GXX is None, reason="g++ not available")
def test_seq_line_reader_across_buffer_refills(tmp_path):
    text = open(os.path.join(CSRC, "wfa_io.cpp")).read()
    reader = between(text, "class LineReader {", "// A line of the .seq format")
    strip = between(text, "void strip(const char* line", "\n}\n") + "\n}\n"
    prelude = "\n".join(["#include <cstdint>", "#include <cstdio>", "#include <cstring>", "#include <random>",
                         "#include <string>", "#include <vector>"])
    build_and_run(tmp_path, "reader", prelude + "\n" + reader + "\n" + strip + "\n" + READER_MAIN)
