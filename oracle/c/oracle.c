/* oracle.c - C/OpenMP restatement of the alpacc runtime (test infrastructure and CPU baseline).
 *
 * Same data-parallel algorithm as futhark/lexer.fut and futhark/parser.fut, over the generated
 * tables in the reference layouts:
 *   lexer  : map bytes to encoded endomorphisms, inclusive scan with `compose`
 *            (compositions[idx(b)*e + idx(a)], lexer.fut:67-81), token-end flags, (+) scan,
 *            take-right scan of token starts, scatter (lexer.fut:91-142); the chunk loop of
 *            lex_int_flag (lexer.fut:144-172) including its re-shift of earlier spans and the
 *            scatter of non-token elements to index prev_size - 1 (highest index wins);
 *   parser : keys over [start] ++ tokens ++ [end], open-addressing lookup with the FNV-style hash
 *            (parser.fut:110-151), segmented copies (:216-233), depth scan (:153-166), the implicit
 *            min-tree and find_previous (:72-108), bracket matching (:174-184), parents
 *            (:286-296), terminal nodes (:305-332).
 * Parallel scans split the input into one contiguous chunk per thread (reduce, scan of the chunk
 * totals, re-scan), as a multicore scan does.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  /* lexer */
  int64_t e;
  const uint32_t *comp;  /* [e*e], entry [b*e + a] = encoded (a then b) */
  const uint32_t *trans; /* [256] encoded */
  const uint8_t *accept; /* [e] */
  uint32_t identity;
  uint32_t index_mask, t_mask, t_off, p_off;
  int32_t ignore;
  /* parser */
  int64_t q, k, start_t, end_t, empty_t;
  int64_t hsize, max_iters;
  const int64_t *hkeys; /* [hsize*(q+k)] */
  const uint8_t *hused; /* [hsize] */
  const int64_t *hvals; /* [hsize*4] */
  const int64_t *stacks;
  int64_t top_bit;
  const int64_t *prods;
  const int32_t *arity;
  const int32_t *p2t;
} otables;

static inline uint32_t to_index(const otables *t, uint32_t a) { return a & t->index_mask; }
static inline uint32_t to_term(const otables *t, uint32_t a) { return (a & t->t_mask) >> t->t_off; }
static inline int is_produce(const otables *t, uint32_t a) { return (a >> t->p_off) & 1u; }
static inline int is_ignore(const otables *t, uint32_t term) {
  return t->ignore >= 0 && (int32_t)term == t->ignore;
}
static inline uint32_t compose(const otables *t, uint32_t a, uint32_t b) {
  return t->comp[(int64_t)to_index(t, b) * t->e + to_index(t, a)];
}

int oracle_threads(void) { return omp_get_max_threads(); }

/* threads of the following calls (the 1-thread CPU baseline); returns the new count */
int oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

/* chunk bounds for thread p of P over [0, n) */
static inline void chunk_of(int64_t n, int p, int P, int64_t *lo, int64_t *hi) {
  *lo = n * p / P;
  *hi = n * (p + 1) / P;
}

/* traverse (lexer.fut:73-81): states[i] = prev o x0 o ... o xi */
static void traverse(const otables *t, uint32_t prev, const uint8_t *s, int64_t n, uint32_t *states) {
  int P = omp_get_max_threads();
  if (n < 4096) P = 1;
  uint32_t *agg = (uint32_t *)malloc(sizeof(uint32_t) * (P + 1));
#pragma omp parallel num_threads(P)
  {
    int p = omp_get_thread_num();
    int64_t lo, hi;
    chunk_of(n, p, P, &lo, &hi);
    uint32_t a = t->identity;
    for (int64_t i = lo; i < hi; ++i) {
      uint32_t x = t->trans[s[i]];
      if (i == 0) x = compose(t, prev, x);
      a = compose(t, a, x);
    }
    agg[p] = a;
#pragma omp barrier
#pragma omp single
    {
      uint32_t run = t->identity;
      for (int j = 0; j < P; ++j) {
        uint32_t v = agg[j];
        agg[j] = run;
        run = compose(t, run, v);
      }
    }
    uint32_t run = agg[p];
    for (int64_t i = lo; i < hi; ++i) {
      uint32_t x = t->trans[s[i]];
      if (i == 0) x = compose(t, prev, x);
      run = compose(t, run, x);
      states[i] = run;
    }
  }
  free(agg);
}

typedef struct {
  uint32_t *kind;
  int64_t *start, *end;
  int64_t cap;
} dest_t;

static void dest_grow(dest_t *d, int64_t need) {
  if (d->cap >= need) return;
  int64_t cap = 2 * need;
  d->kind = (uint32_t *)realloc(d->kind, sizeof(uint32_t) * cap);
  d->start = (int64_t *)realloc(d->start, sizeof(int64_t) * cap);
  d->end = (int64_t *)realloc(d->end, sizeof(int64_t) * cap);
  for (int64_t i = d->cap; i < cap; ++i) {
    d->kind[i] = 0;
    d->start[i] = 0;
    d->end[i] = 0;
  }
  d->cap = cap;
}

/* lex_step (lexer.fut:91-142) */
static void lex_step(const otables *t, int64_t offset, uint32_t *state, int64_t *start, int64_t *size,
                     dest_t *dest, const uint8_t *s, int64_t n, uint32_t *states, int64_t remap_len) {
  traverse(t, *state, s, n, states);
  int P = omp_get_max_threads();
  if (n < 4096) P = 1;
  int64_t *cnt = (int64_t *)calloc(P + 1, sizeof(int64_t));
  int64_t *lst = (int64_t *)malloc(sizeof(int64_t) * (P + 1));
  const int64_t NONE = INT64_MAX;
  const int64_t prev_size = *size;
  const int64_t prev_start = *start;
#pragma omp parallel num_threads(P)
  {
    int p = omp_get_thread_num();
    int64_t lo, hi;
    chunk_of(n, p, P, &lo, &hi);
    int64_t c = 0, last = NONE;
    for (int64_t i = lo; i < hi; ++i) {
      uint32_t st = states[i];
      if (i != n - 1 && is_produce(t, states[i + 1]) && !is_ignore(t, to_term(t, st))) ++c;
      if (is_produce(t, st) && !is_ignore(t, to_term(t, st)))
        last = i;
      else if (i == 0)
        last = prev_start - offset;
    }
    cnt[p] = c;
    lst[p] = last;
#pragma omp barrier
#pragma omp single
    {
      int64_t run = 0, rl = NONE;
      for (int j = 0; j < P; ++j) {
        int64_t v = cnt[j], l = lst[j];
        cnt[j] = run;
        run += v;
        int64_t tmp = lst[j];
        lst[j] = rl;
        if (tmp != NONE) rl = tmp;
        (void)l;
      }
      cnt[P] = run;
      lst[P] = rl;
      dest_grow(dest, prev_size + run + 1);
    }
    int64_t o = cnt[p];
    int64_t cur = lst[p];
    for (int64_t i = lo; i < hi; ++i) {
      uint32_t st = states[i];
      if (is_produce(t, st) && !is_ignore(t, to_term(t, st)))
        cur = i;
      else if (i == 0)
        cur = prev_start - offset;
      if (i != n - 1 && is_produce(t, states[i + 1]) && !is_ignore(t, to_term(t, st))) {
        int64_t j = prev_size + o;
        dest->kind[j] = to_term(t, st);
        dest->start[j] = cur;
        dest->end[j] = i;
        ++o;
      }
    }
  }
  /* scatter of the non-token elements to prev_size - 1: the highest such element wins (it is
   * always the last element n - 1, which is never a token end) */
  if (prev_size - 1 >= 0) {
    int64_t i = n - 1;
    /* its take-right start */
    int64_t cur = lst[P];
    {
      uint32_t st = states[i];
      if (is_produce(t, st) && !is_ignore(t, to_term(t, st))) cur = i;
      else if (i == 0) cur = prev_start - offset;
    }
    dest->kind[prev_size - 1] = to_term(t, states[i]);
    dest->start[prev_size - 1] = cur;
    dest->end[prev_size - 1] = i;
  }
  /* re-map of the destination (lexer.fut:131-133): every entry, not only the new ones */
  int64_t m = dest->cap < remap_len ? dest->cap : remap_len;
  if (remap_len < 0) m = dest->cap;
#pragma omp parallel for
  for (int64_t j = 0; j < m; ++j) {
    dest->start[j] += offset;
    dest->end[j] += 1 + offset;
  }
  int64_t last_start = lst[P];
  {
    uint32_t st = states[n - 1];
    if (is_produce(t, st) && !is_ignore(t, to_term(t, st))) last_start = n - 1;
    else if (n - 1 == 0) last_start = prev_start - offset;
  }
  *state = states[n - 1];
  *start = last_start != NONE ? offset + last_start : prev_start;
  *size = prev_size + cnt[P];
  free(cnt);
  free(lst);
}

/* lex_int_flag (lexer.fut:144-172); chunk 0 = one chunk over the whole input.
 * Returns 1 accepted / 0 rejected; outputs are malloc'ed (caller frees with oracle_free). */
int oracle_lex(const otables *t, const uint8_t *s, int64_t n, int64_t chunk, uint32_t **kinds,
               int64_t **starts, int64_t **ends, int64_t *ntok) {
  int64_t C = chunk == 0 ? n : chunk;
  dest_t dest = {0, 0, 0, 0};
  dest_grow(&dest, 1);
  uint32_t state = t->identity;
  int64_t start = 0, size = 0;
  uint32_t *states = (uint32_t *)malloc(sizeof(uint32_t) * ((C + 1) < n ? (C + 1) : (n > 0 ? n : 1)));
  for (int64_t off = 0; n > 0 && off < n; off += (C > 0 ? C : 1)) {
    int64_t m = off + C + 1 < n ? off + C + 1 : n;
    lex_step(t, off, &state, &start, &size, &dest, s + off, m - off, states, chunk == 0 ? size + n + 1 : -1);
  }
  free(states);
  uint32_t last_t = to_term(t, state);
  dest_grow(&dest, size + 1);
  if (!is_ignore(t, last_t)) {
    dest.kind[size] = last_t;
    dest.start[size] = start;
    dest.end[size] = n;
    size++;
  }
  *kinds = dest.kind;
  *starts = dest.start;
  *ends = dest.end;
  *ntok = size;
  return t->accept[to_index(t, state)] ? 1 : 0;
}

void oracle_free(void *p) { free(p); }

/* ------------------------------------------------------------------------------------------ */
/* parser */

static inline uint64_t fnv(const int64_t *key, int64_t m) {
  uint64_t h = 14695981039346656037ull;
  for (int64_t j = 0; j < m; ++j) {
    h ^= (uint64_t)key[j];
    h *= 1099511628211ull;
  }
  return h;
}

/* lookup (parser.fut:130-142); returns 0 when missing */
static int lookup(const otables *t, const int64_t *key, int64_t out[4]) {
  const int64_t m = t->q + t->k;
  if (t->hsize == 0) return 0;
  uint64_t h = fnv(key, m) % (uint64_t)t->hsize;
  int found = 0, is_found = 0;
  int64_t i = 0;
  out[0] = out[1] = out[2] = out[3] = -1;
  while (is_found || i < t->max_iters) {
    const int64_t *k2 = t->hkeys + (int64_t)h * m;
    int valid = t->hused[h];
    for (int64_t j = 0; valid && j < m; ++j) valid = k2[j] == key[j];
    if (valid) {
      memcpy(out, t->hvals + (int64_t)h * 4, 4 * sizeof(int64_t));
      found = 1;
    }
    is_found = valid;
    ++i;
    h = (h + 1) % (uint64_t)t->hsize;
  }
  return found;
}

/* exclusive (+) scan in place, returns the total */
static int64_t exscan_i64(int64_t *a, int64_t n) {
  int P = omp_get_max_threads();
  if (n < 65536) P = 1;
  int64_t *part = (int64_t *)calloc(P + 1, sizeof(int64_t));
#pragma omp parallel num_threads(P)
  {
    int p = omp_get_thread_num();
    int64_t lo, hi;
    chunk_of(n, p, P, &lo, &hi);
    int64_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += a[i];
    part[p] = s;
#pragma omp barrier
#pragma omp single
    {
      int64_t run = 0;
      for (int j = 0; j < P; ++j) {
        int64_t v = part[j];
        part[j] = run;
        run += v;
      }
      part[P] = run;
    }
    int64_t run = part[p];
    for (int64_t i = lo; i < hi; ++i) {
      int64_t v = a[i];
      a[i] = run;
      run += v;
    }
  }
  int64_t total = part[P];
  free(part);
  return total;
}

/* mk_tree (parser.fut:72-90) with op = min, ne = i64.highest; returns tree, sets *tsize */
static int64_t *mk_tree(const int64_t *arr, int64_t n, int64_t *tsize) {
  int temp = n == 0 ? 0 : 64 - __builtin_clzll((uint64_t)n);
  int h = __builtin_popcountll((uint64_t)n) == 1 ? temp : temp + 1;
  int64_t size = ((int64_t)1 << h) - 1;
  int64_t offset = ((int64_t)1 << (h - 1)) - 1;
  int64_t *tree = (int64_t *)malloc(sizeof(int64_t) * (size > 0 ? size : 1));
#pragma omp parallel for
  for (int64_t i = 0; i < size; ++i) tree[i] = INT64_MAX;
#pragma omp parallel for
  for (int64_t i = 0; i < n; ++i) tree[offset + i] = arr[i];
  for (int level = h - 2; level >= 0; --level) {
    int64_t off = ((int64_t)1 << level) - 1;
    int64_t cnt = (int64_t)1 << level;
    int64_t child = ((int64_t)1 << (level + 1)) - 1;
#pragma omp parallel for
    for (int64_t i = 0; i < cnt; ++i) {
      int64_t a = tree[child + 2 * i], b = tree[child + 2 * i + 1];
      tree[off + i] = a < b ? a : b;
    }
  }
  *tsize = size;
  return tree;
}

/* find_previous (<=) (parser.fut:92-108) */
static int64_t find_previous_le(const int64_t *tree, int64_t n, int64_t idx) {
  int h = 64 - __builtin_clzll((uint64_t)n);
  int64_t offset = ((int64_t)1 << (h - 1)) - 1;
  int64_t i = offset + idx;
  int64_t v = tree[i];
  while (i != 0) {
    int64_t sib = (i % 2 == 0) ? i - 1 : i + 1;
    int is_left = i % 2 == 1;
    if (!(is_left || !(tree[sib] <= v))) break;
    i = (i - 1) / 2;
  }
  if (i == 0) return -1;
  int64_t j = (i % 2 == 0) ? i - 1 : i + 1;
  while (j < offset) j = 2 * j + 1 + (tree[2 * j + 2] <= v ? 1 : 0);
  return j - offset;
}

/* parse_int (parser.fut:316-336) over T tokens. Outputs (parent, node, span) like the C ABI:
 * node bit 31 = terminal. Returns 1 accepted, 0 rejected. Outputs malloc'ed. */
int oracle_parse(const otables *t, const uint32_t *kinds, const int64_t *starts, const int64_t *ends,
                 int64_t T, uint32_t **parent_out, uint32_t **node_out, uint32_t **span_out,
                 int64_t *nnodes) {
  const int64_t P = T + 2, m = t->q + t->k;
  int64_t *spans = (int64_t *)malloc(sizeof(int64_t) * 4 * P);
  int64_t *slen = (int64_t *)malloc(sizeof(int64_t) * (P + 1));
  int64_t *plen = (int64_t *)malloc(sizeof(int64_t) * (P + 1));
  int ok = 1;
#pragma omp parallel for reduction(&& : ok)
  for (int64_t p = 0; p < P; ++p) {
    int64_t key[64];
    for (int64_t j = 0; j < m; ++j) {
      int64_t pos = p + j - t->q;
      int64_t v;
      if (pos < 0 || pos >= P) v = t->empty_t;
      else if (pos == 0) v = t->start_t;
      else if (pos == P - 1) v = t->end_t;
      else v = kinds[pos - 1];
      key[j] = v;
    }
    int f = lookup(t, key, spans + 4 * p);
    ok = ok && f;
    slen[p] = f ? spans[4 * p + 1] - spans[4 * p] : 0;
    plen[p] = f ? spans[4 * p + 3] - spans[4 * p + 2] : 0;
  }
  if (!ok) {
    free(spans);
    free(slen);
    free(plen);
    return 0;
  }
  int64_t B = exscan_i64(slen, P);
  int64_t N = exscan_i64(plen, P);
  int64_t *br = (int64_t *)malloc(sizeof(int64_t) * (B > 0 ? B : 1));
  int64_t *pr = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
#pragma omp parallel for
  for (int64_t p = 0; p < P; ++p) {
    for (int64_t s = spans[4 * p]; s < spans[4 * p + 1]; ++s) br[slen[p] + s - spans[4 * p]] = t->stacks[s];
    for (int64_t s = spans[4 * p + 2]; s < spans[4 * p + 3]; ++s) pr[plen[p] + s - spans[4 * p + 2]] = t->prods[s];
  }
  free(spans);
  free(slen);
  free(plen);
  /* depths (parser.fut:153-166) */
  int64_t *sc = (int64_t *)malloc(sizeof(int64_t) * (B > 0 ? B : 1));
#pragma omp parallel for
  for (int64_t i = 0; i < B; ++i) sc[i] = (br[i] & t->top_bit) ? 1 : -1;
  int64_t total = exscan_i64(sc, B);
  int bad = 0;
#pragma omp parallel for reduction(|| : bad)
  for (int64_t i = 0; i < B; ++i) {
    int left = (br[i] & t->top_bit) != 0;
    int64_t incl = sc[i] + (left ? 1 : -1);
    bad = bad || incl < 0;
    sc[i] = incl - (left ? 1 : 0);
  }
  if (bad || (B != 0 && total != 0)) {
    free(br);
    free(pr);
    free(sc);
    return 0;
  }
  if (B > 0) {
    int64_t ts;
    int64_t *tree = mk_tree(sc, B, &ts);
    int mism = 0;
#pragma omp parallel for reduction(|| : mism)
    for (int64_t i = 0; i < B; ++i) {
      if (br[i] & t->top_bit) continue;
      int64_t j = find_previous_le(tree, ts, i);
      mism = mism || j < 0 || ((br[i] & ~t->top_bit) != (br[j] & ~t->top_bit));
    }
    free(tree);
    if (mism) {
      free(br);
      free(pr);
      free(sc);
      return 0;
    }
  }
  free(br);
  free(sc);
  /* parents (parser.fut:286-296) */
  int64_t *s = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
  int64_t *rank = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
#pragma omp parallel for
  for (int64_t i = 0; i < N; ++i) {
    s[i] = (int64_t)t->arity[pr[i]] - 1;
    rank[i] = t->p2t[pr[i]] >= 0 ? 1 : 0;
  }
  exscan_i64(s, N);
  int64_t nterm = exscan_i64(rank, N);
  uint32_t *parent = (uint32_t *)malloc(sizeof(uint32_t) * (N > 0 ? N : 1));
  uint32_t *node = (uint32_t *)malloc(sizeof(uint32_t) * (N > 0 ? N : 1));
  uint32_t *span = (uint32_t *)malloc(sizeof(uint32_t) * 2 * (N > 0 ? N : 1));
  if (N > 0) {
    int64_t ts;
    int64_t *tree = mk_tree(s, N, &ts);
    const int subst = nterm == T; /* safe_zip (parser.fut:300-303) */
#pragma omp parallel for
    for (int64_t i = 0; i < N; ++i) {
      int64_t par = i == 0 ? 0 : find_previous_le(tree, ts, i);
      parent[i] = (uint32_t)par;
      int32_t tt = t->p2t[pr[i]];
      if (subst && tt >= 0) {
        node[i] = 0x80000000u | (uint32_t)tt;
        span[2 * i] = (uint32_t)starts[rank[i]];
        span[2 * i + 1] = (uint32_t)ends[rank[i]];
      } else {
        node[i] = (uint32_t)pr[i];
        span[2 * i] = 0;
        span[2 * i + 1] = 0;
      }
    }
    free(tree);
  }
  free(s);
  free(rank);
  free(pr);
  *parent_out = parent;
  *node_out = node;
  *span_out = span;
  *nnodes = N;
  return 1;
}
