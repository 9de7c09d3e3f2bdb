/*
 * Seeded synthetic pair generator for the benchmark workloads.
 *
 * Every pair is generated from its own SplitMix64 stream seeded by
 * (seed, pair index), so any pair regenerates independently of the others.
 * The text is uniform ACGT; the pattern is a copy of the text with error
 * events applied one after another at uniform positions of the current
 * pattern: a substitution always changes the base, an insertion inserts
 * uniform bases, a deletion removes bases (the error model of
 * tools/generate_dataset/generate_dataset.c:144-217, seeded, and with
 * geometric indel lengths where a workload asks for them).
 *
 * Workloads (shape ids):
 *   0  1,000 bp, 5% errors, 1-base indels, ceil(e*L) events
 *   1  L ~ U{100..300}, e ~ U[0,0.10], 1-base indels, ceil(e*L) events
 *   2  L log-uniform [1,000, 20,000], e ~ U[0.01,0.15], indel length
 *      Geometric(mean 3), round(e*L) events
 *   3  100,000 bp, 15% errors, indel length Geometric(mean 2), round(e*L) events
 *   4  serving mixture: 60% shape 0, 35% shape 2, 5% shape 3
 * Event types are substitution / insertion / deletion with probability 1/3 each.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  char* buf;          /* patterns and texts, back to back */
  int64_t bytes;
  int n;
  int* plen;
  int* tlen;
  int64_t* p_off;
  int64_t* t_off;
  float* err;         /* error rate drawn for the pair */
  int* shape;         /* shape id actually used (differs from the request for 4) */
} wfa_gen_t;

static inline uint64_t sm64(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline double unif(uint64_t* s) { return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0); }
static inline int64_t below(uint64_t* s, int64_t n) { return (int64_t)(unif(s) * (double)n); }
static inline int geometric(uint64_t* s, double mean) {
  if (mean <= 1.0) return 1;
  const double p = 1.0 / mean;
  double u = unif(s);
  if (u <= 0.0) u = 1e-300;
  return 1 + (int)floor(log(u) / log(1.0 - p));
}

static const char kBases[4] = {'A', 'C', 'G', 'T'};

typedef struct { char* s; int64_t len, cap; } strbuf_t;
static void sb_reserve(strbuf_t* b, int64_t need) {
  if (need <= b->cap) return;
  int64_t c = b->cap ? b->cap : 1024;
  while (c < need) c *= 2;
  b->s = realloc(b->s, c);
  b->cap = c;
}

static void make_pair(int shape, uint64_t* rng, strbuf_t* text, strbuf_t* pat, float* err) {
  int64_t L;
  double e, mean_indel;
  int use_round;
  switch (shape) {
    case 0: L = 1000; e = 0.05; mean_indel = 1.0; use_round = 0; break;
    case 1: L = 100 + below(rng, 201); e = 0.10 * unif(rng); mean_indel = 1.0; use_round = 0; break;
    case 2: L = (int64_t)llround(exp(log(1000.0) + unif(rng) * (log(20000.0) - log(1000.0))));
            e = 0.01 + 0.14 * unif(rng); mean_indel = 3.0; use_round = 1; break;
    default: L = 100000; e = 0.15; mean_indel = 2.0; use_round = 1; break;
  }
  *err = (float)e;
  sb_reserve(text, L);
  for (int64_t i = 0; i < L; ++i) text->s[i] = kBases[sm64(rng) & 3];
  text->len = L;
  sb_reserve(pat, L);
  memcpy(pat->s, text->s, L);
  pat->len = L;
  const int64_t events = use_round ? (int64_t)llround(e * (double)L) : (int64_t)ceil(e * (double)L - 1e-9);
  for (int64_t ev = 0; ev < events; ++ev) {
    const int64_t kind = below(rng, 3);
    if (kind == 0) {                       /* substitution */
      if (pat->len == 0) continue;
      const int64_t pos = below(rng, pat->len);
      int c;
      do { c = (int)(sm64(rng) & 3); } while (kBases[c] == pat->s[pos]);
      pat->s[pos] = kBases[c];
    } else if (kind == 1) {                /* insertion */
      const int ln = geometric(rng, mean_indel);
      const int64_t pos = below(rng, pat->len + 1);
      sb_reserve(pat, pat->len + ln);
      memmove(pat->s + pos + ln, pat->s + pos, pat->len - pos);
      for (int i = 0; i < ln; ++i) pat->s[pos + i] = kBases[sm64(rng) & 3];
      pat->len += ln;
    } else {                               /* deletion */
      if (pat->len == 0) continue;
      int64_t ln = geometric(rng, mean_indel);
      const int64_t pos = below(rng, pat->len);
      if (ln > pat->len - pos) ln = pat->len - pos;
      memmove(pat->s + pos, pat->s + pos + ln, pat->len - pos - ln);
      pat->len -= ln;
    }
  }
}

wfa_gen_t* wfa_gen(int shape, uint64_t seed, int64_t first, int n) {
  wfa_gen_t* g = calloc(1, sizeof(wfa_gen_t));
  g->n = n;
  g->plen = malloc(sizeof(int) * (n > 0 ? n : 1));
  g->tlen = malloc(sizeof(int) * (n > 0 ? n : 1));
  g->p_off = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  g->t_off = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  g->err = malloc(sizeof(float) * (n > 0 ? n : 1));
  g->shape = malloc(sizeof(int) * (n > 0 ? n : 1));
  strbuf_t out = {0}, text = {0}, pat = {0};
  for (int i = 0; i < n; ++i) {
    uint64_t rng = seed * 0xD1B54A32D192ED03ull + (uint64_t)(first + i) * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    (void)sm64(&rng);
    int sh = shape;
    if (shape == 4) {
      const double u = unif(&rng);
      sh = (u < 0.60) ? 0 : (u < 0.95 ? 2 : 3);
    }
    make_pair(sh, &rng, &text, &pat, &g->err[i]);
    g->shape[i] = sh;
    sb_reserve(&out, out.len + pat.len + text.len);
    g->p_off[i] = out.len;
    g->plen[i] = (int)pat.len;
    memcpy(out.s + out.len, pat.s, pat.len);
    out.len += pat.len;
    g->t_off[i] = out.len;
    g->tlen[i] = (int)text.len;
    memcpy(out.s + out.len, text.s, text.len);
    out.len += text.len;
  }
  free(text.s);
  free(pat.s);
  g->buf = out.s;
  g->bytes = out.len;
  return g;
}

void wfa_gen_free(wfa_gen_t* g) {
  if (!g) return;
  free(g->buf);
  free(g->plen);
  free(g->tlen);
  free(g->p_off);
  free(g->t_off);
  free(g->err);
  free(g->shape);
  free(g);
}
