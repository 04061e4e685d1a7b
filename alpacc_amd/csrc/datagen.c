/* datagen.c - seeded synthetic documents for the benchmark and test workloads.
 *
 * Every generator fills exactly `n` bytes: it emits whole values while they fit, closes the
 * document, and pads the tail with spaces (an ignore token in every target grammar).
 *   gen_json  : a top-level array of JSON values; nesting depth <= 16, container fan-out 1-8
 *               (geometric), printable-ASCII strings of length 0-32 with ~5% two-byte escapes
 *               from {\" \/ \n \t \r \b \f} (never a backslash before the closing quote), integers
 *               and decimals without exponent, true/false/null, whitespace runs of 0-4 bytes
 *               from {space, \n, \t, \r} making up ~10% of the bytes.
 *   gen_arith : a chain t1 op t2 op ... of random expression trees of depth <= 12; leaves are
 *               -?[0-9]{1,6} or [a-z]{1,8} (50/50), operators + - * / uniform, parentheses with
 *               p = 0.2, one space between tokens with p = 0.3, and a forced space after a binary
 *               '-' that precedes a digit (otherwise the lexer reads a negative number).
 *   gen_sexp  : one top-level list of s-expressions; atoms [a-z0-9]{1,12}, list depth <= 24,
 *               fan-out 0-10, atoms always separated by whitespace.
 * The PRNG is xoshiro256** seeded through splitmix64.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed) {
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&seed);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t *r) {
  uint64_t *s = r->s;
  const uint64_t result = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

/* uniform in [0, m) */
static inline uint32_t rng_below(rng_t *r, uint32_t m) {
  return (uint32_t)(((rng_next(r) >> 32) * (uint64_t)m) >> 32);
}

/* uniform in [0, 1) with 24 bits */
static inline double rng_unit(rng_t *r) { return (double)(rng_next(r) >> 40) / 16777216.0; }

typedef struct {
  uint8_t *buf;
  size_t len, cap;
  int overflow;
} out_t;

static inline void put(out_t *o, uint8_t c) {
  if (o->len < o->cap)
    o->buf[o->len++] = c;
  else
    o->overflow = 1;
}

static inline void puts_(out_t *o, const char *s) {
  while (*s) put(o, (uint8_t)*s++);
}

/* ------------------------------------------------------------------------------------------ */
/* JSON */

static const char WS[4] = {' ', '\n', '\t', '\r'};

static void json_ws(out_t *o, rng_t *r) {
  /* whitespace run of 0-4 bytes; about 10% of the document */
  if (rng_unit(r) < 0.25) {
    uint32_t k = 1 + rng_below(r, 4);
    for (uint32_t i = 0; i < k; ++i) put(o, (uint8_t)WS[rng_below(r, 4)]);
  }
}

static const char *ESC[7] = {"\\\"", "\\/", "\\n", "\\t", "\\r", "\\b", "\\f"};

static void json_string(out_t *o, rng_t *r) {
  put(o, '"');
  uint32_t len = rng_below(r, 33);
  for (uint32_t i = 0; i < len; ++i) {
    if (rng_unit(r) < 0.05) {
      puts_(o, ESC[rng_below(r, 7)]);
    } else {
      uint8_t c;
      do {
        c = (uint8_t)(0x20 + rng_below(r, 95));
      } while (c == '"' || c == '\\');
      put(o, c);
    }
  }
  put(o, '"');
}

static void json_number(out_t *o, rng_t *r) {
  if (rng_unit(r) < 0.3) put(o, '-');
  uint32_t d = 1 + rng_below(r, 8);
  put(o, (uint8_t)('1' + rng_below(r, 9)));
  for (uint32_t i = 1; i < d; ++i) put(o, (uint8_t)('0' + rng_below(r, 10)));
  if (rng_unit(r) < 0.5) {
    put(o, '.');
    uint32_t f = 1 + rng_below(r, 6);
    for (uint32_t i = 0; i < f; ++i) put(o, (uint8_t)('0' + rng_below(r, 10)));
  }
}

static uint32_t geometric_fanout(rng_t *r) {
  uint32_t k = 1;
  while (k < 8 && rng_unit(r) < 0.6) ++k;
  return k;
}

static void json_value(out_t *o, rng_t *r, int depth);

static void json_container(out_t *o, rng_t *r, int depth, int is_obj) {
  put(o, is_obj ? '{' : '[');
  json_ws(o, r);
  uint32_t k = geometric_fanout(r);
  for (uint32_t i = 0; i < k && !o->overflow; ++i) {
    if (i) {
      put(o, ',');
      json_ws(o, r);
    }
    if (is_obj) {
      json_string(o, r);
      json_ws(o, r);
      put(o, ':');
      json_ws(o, r);
    }
    json_value(o, r, depth + 1);
    json_ws(o, r);
  }
  put(o, is_obj ? '}' : ']');
}

static void json_value(out_t *o, rng_t *r, int depth) {
  double u = rng_unit(r);
  double p_container = depth < 16 ? 0.28 : 0.0;
  if (u < p_container) {
    json_container(o, r, depth, rng_below(r, 2));
    return;
  }
  u = rng_unit(r);
  if (u < 0.4)
    json_string(o, r);
  else if (u < 0.8)
    json_number(o, r);
  else {
    static const char *KW[3] = {"true", "false", "null"};
    puts_(o, KW[rng_below(r, 3)]);
  }
}

size_t gen_json(uint8_t *buf, size_t n, uint64_t seed) {
  rng_t r;
  rng_seed(&r, seed);
  if (n < 2) return 0;
  size_t limit = n - 1; /* room for the closing bracket */
  out_t o = {buf, 0, n, 0};
  put(&o, '[');
  uint8_t *tmp = (uint8_t *)malloc(1 << 20);
  if (!tmp) return 0;
  int first = 1;
  while (1) {
    out_t t = {tmp, 0, 1 << 20, 0};
    if (!first) {
      put(&t, ',');
      json_ws(&t, &r);
    }
    json_value(&t, &r, 1);
    json_ws(&t, &r);
    if (t.overflow) continue;
    if (o.len + t.len > limit) break;
    memcpy(o.buf + o.len, tmp, t.len);
    o.len += t.len;
    first = 0;
  }
  free(tmp);
  put(&o, ']');
  while (o.len < n) o.buf[o.len++] = ' ';
  return o.len;
}

/* ------------------------------------------------------------------------------------------ */
/* arithmetic */

static int arith_space(out_t *o, rng_t *r) {
  if (rng_unit(r) < 0.3) {
    put(o, ' ');
    return 1;
  }
  return 0;
}

/* emits a leaf; returns 1 if it starts with a digit */
static void arith_leaf(out_t *o, rng_t *r) {
  if (rng_below(r, 2)) {
    if (rng_below(r, 2)) put(o, '-');
    uint32_t d = 1 + rng_below(r, 6);
    for (uint32_t i = 0; i < d; ++i) put(o, (uint8_t)('0' + rng_below(r, 10)));
  } else {
    uint32_t d = 1 + rng_below(r, 8);
    for (uint32_t i = 0; i < d; ++i) put(o, (uint8_t)('a' + rng_below(r, 26)));
  }
}

/* last_minus: the previous token was a binary '-' without a space after it */
static void arith_tree(out_t *o, rng_t *r, int depth, int *last_minus);

static void arith_operand(out_t *o, rng_t *r, int depth, int *last_minus) {
  if (rng_unit(r) < 0.2) {
    put(o, '(');
    *last_minus = 0;
    arith_space(o, r);
    arith_tree(o, r, depth + 1, last_minus);
    arith_space(o, r);
    put(o, ')');
    *last_minus = 0;
    return;
  }
  arith_tree(o, r, depth, last_minus);
}

static void arith_tree(out_t *o, rng_t *r, int depth, int *last_minus) {
  double p_leaf = depth >= 12 ? 1.0 : 0.55 + 0.03 * depth;
  if (rng_unit(r) < p_leaf) {
    /* a leaf right after a binary '-' must not start with a digit or '-' */
    size_t at = o->len;
    arith_leaf(o, r);
    if (*last_minus && at < o->len && at < o->cap &&
        ((o->buf[at] >= '0' && o->buf[at] <= '9') || o->buf[at] == '-')) {
      /* insert the forced space */
      if (o->len < o->cap) {
        memmove(o->buf + at + 1, o->buf + at, o->len - at);
        o->buf[at] = ' ';
        o->len++;
      } else {
        o->overflow = 1;
      }
    }
    *last_minus = 0;
    return;
  }
  arith_operand(o, r, depth + 1, last_minus);
  int sp = arith_space(o, r);
  (void)sp;
  static const char OPS[4] = {'+', '-', '*', '/'};
  char op = OPS[rng_below(r, 4)];
  put(o, (uint8_t)op);
  *last_minus = (op == '-') && !arith_space(o, r);
  if (op != '-') arith_space(o, r);
  arith_operand(o, r, depth + 1, last_minus);
}

size_t gen_arith(uint8_t *buf, size_t n, uint64_t seed) {
  rng_t r;
  rng_seed(&r, seed);
  if (n < 1) return 0;
  out_t o = {buf, 0, n, 0};
  uint8_t *tmp = (uint8_t *)malloc(1 << 20);
  if (!tmp) return 0;
  int first = 1;
  int last_minus = 0;
  while (1) {
    out_t t = {tmp, 0, 1 << 20, 0};
    int lm = 0;
    if (!first) {
      arith_space(&t, &r);
      static const char OPS[4] = {'+', '-', '*', '/'};
      char op = OPS[rng_below(&r, 4)];
      put(&t, (uint8_t)op);
      lm = (op == '-') && !arith_space(&t, &r);
      if (op != '-') arith_space(&t, &r);
    }
    arith_tree(&t, &r, 0, &lm);
    if (t.overflow) continue;
    if (o.len + t.len > n) break;
    memcpy(o.buf + o.len, tmp, t.len);
    o.len += t.len;
    first = 0;
    last_minus = lm;
  }
  (void)last_minus;
  free(tmp);
  if (first) { /* nothing fitted: a single digit */
    o.buf[0] = '0';
    o.len = 1;
  }
  while (o.len < n) o.buf[o.len++] = ' ';
  return o.len;
}

/* ------------------------------------------------------------------------------------------ */
/* s-expressions */

static void sexp_ws(out_t *o, rng_t *r) {
  uint32_t k = 1 + (rng_unit(r) < 0.2 ? rng_below(r, 3) : 0);
  for (uint32_t i = 0; i < k; ++i) put(o, (uint8_t)WS[rng_below(r, 4)]);
}

static void sexp_atom(out_t *o, rng_t *r) {
  static const char AL[] = "abcdefghijklmnopqrstuvwxyz0123456789";
  uint32_t d = 1 + rng_below(r, 12);
  for (uint32_t i = 0; i < d; ++i) put(o, (uint8_t)AL[rng_below(r, 36)]);
}

static void sexp_expr(out_t *o, rng_t *r, int depth) {
  if (depth >= 24 || rng_unit(r) < 0.85) {
    sexp_atom(o, r);
    return;
  }
  put(o, '(');
  uint32_t k = rng_below(r, 11);
  for (uint32_t i = 0; i < k && !o->overflow; ++i) {
    if (i) sexp_ws(o, r);
    sexp_expr(o, r, depth + 1);
  }
  put(o, ')');
}

size_t gen_sexp(uint8_t *buf, size_t n, uint64_t seed) {
  rng_t r;
  rng_seed(&r, seed);
  if (n < 2) return 0;
  size_t limit = n - 1;
  out_t o = {buf, 0, n, 0};
  put(&o, '(');
  uint8_t *tmp = (uint8_t *)malloc(1 << 20);
  if (!tmp) return 0;
  int first = 1;
  while (1) {
    out_t t = {tmp, 0, 1 << 20, 0};
    if (!first) sexp_ws(&t, &r);
    sexp_expr(&t, &r, 1);
    if (t.overflow) continue;
    if (o.len + t.len > limit) break;
    memcpy(o.buf + o.len, tmp, t.len);
    o.len += t.len;
    first = 0;
  }
  free(tmp);
  put(&o, ')');
  while (o.len < n) o.buf[o.len++] = ' ';
  return o.len;
}
