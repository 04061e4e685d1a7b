/*
 * Stand-alone lexer/parser executable for one grammar, built on the C ABI of libalpacc_hip.
 *
 * It takes the place of the executable the reference's CUDA backend generates (cuda/lexer.cu:
 * lexer_stream reads stdin and prints one "start end terminal" line per token, :421-425, then exits
 * with 0 when the input is accepted and -1 otherwise, :532). The grammar's tables come from a
 * header written by alpacc_amd.gen.emit_c; compile with
 *   -DALPACC_TABLES_HEADER='"json_tables.h"' -DALPACC_PREFIX=json
 *
 * usage: alpacc_<grammar> [lex|parse|test] [--names|--readme] < input
 *   lex   : "start end terminal" per token (the reference's WriteAscii format)
 *   parse : "parent T terminal start end" or "parent P production" per node, in preorder
 *   test  : the artifact's `test` entry (futhark/test.fut): wire-format inputs on stdin,
 *           wire-format outputs on stdout
 * --names prints terminal/production names instead of ids.
 * --readme prints the result as README.md:84-106 shows a parse: "#some [(0, #production (#E0_0)),
 *   (2, #terminal (#num) (0, 2)), ...]" one item a line (for lex "(#num, (0, 2))" items), or
 *   "#none" for a rejected input.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include ALPACC_TABLES_HEADER

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define TABLES CAT(ALPACC_PREFIX, _tables)
#define TERMINAL_NAMES CAT(ALPACC_PREFIX, _terminal_names)
#define PRODUCTION_NAMES CAT(ALPACC_PREFIX, _production_names)

static uint8_t *read_all(FILE *f, uint64_t *n) {
  size_t cap = 1 << 20, len = 0;
  uint8_t *buf = (uint8_t *)malloc(cap);
  if (!buf) return NULL;
  for (;;) {
    if (len == cap) {
      cap *= 2;
      uint8_t *nb = (uint8_t *)realloc(buf, cap);
      if (!nb) {
        free(buf);
        return NULL;
      }
      buf = nb;
    }
    size_t r = fread(buf + len, 1, cap - len, f);
    len += r;
    if (r == 0) break;
  }
  *n = len;
  return buf;
}

static void *upload(const void *h, uint64_t bytes) {
  void *d = alpacc_device_alloc(bytes ? bytes : 1);
  if (d && bytes && alpacc_copy_to_device(d, h, bytes, NULL) != 0) {
    alpacc_device_free(d);
    return NULL;
  }
  return d;
}

static int fail(const char *what, alpacc_ctx *ctx) {
  fprintf(stderr, "alpacc: %s%s%s\n", what, ctx ? ": " : "", ctx ? alpacc_last_error(ctx) : "");
  return 2;
}

int main(int argc, char **argv) {
  int parse = 0, names = 0, test = 0, readme = 0;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "parse")) parse = 1;
    else if (!strcmp(argv[i], "lex")) parse = 0;
    else if (!strcmp(argv[i], "test")) test = 1;
    else if (!strcmp(argv[i], "--names")) names = 1;
    else if (!strcmp(argv[i], "--readme")) readme = 1;
    else {
      fprintf(stderr, "usage: %s [lex|parse|test] [--names|--readme] < input\n", argv[0]);
      return 2;
    }
  }
  uint64_t n = 0;
  uint8_t *in = read_all(stdin, &n);
  if (!in) return fail("out of host memory", NULL);
  if (alpacc_device_count() < 1 || alpacc_set_device(0) != 0) return fail("no ROCm device", NULL);
  alpacc_ctx *ctx = alpacc_create(&TABLES, 0);
  if (!ctx) {
    fprintf(stderr, "alpacc: %s\n", alpacc_create_error());
    return 2;
  }
  if (test) {
    uint8_t *out = NULL;
    uint64_t out_n = 0;
    if (alpacc_test(ctx, in, n, &out, &out_n, NULL) < 0) return fail("test entry failed", ctx);
    fwrite(out, 1, out_n, stdout);
    fflush(stdout);
    alpacc_free(out);
    alpacc_destroy(ctx);
    free(in);
    return 0;
  }
  uint8_t *d_in = (uint8_t *)upload(in, n);
  if (!d_in) return fail("device upload failed", ctx);
  int rc;
  if (!parse) {
    const uint64_t cap = n ? n : 1;
    uint32_t *d = (uint32_t *)alpacc_device_alloc(3 * cap * sizeof(uint32_t));
    if (!d) return fail("device allocation failed", ctx);
    uint64_t T = 0;
    rc = alpacc_lex(ctx, d_in, n, d, d + cap, d + 2 * cap, cap, &T, NULL);
    if (rc < 0) return fail("lex failed", ctx);
    if (rc == ALPACC_ACCEPT) {
      uint32_t *h = (uint32_t *)malloc(3 * cap * sizeof(uint32_t));
      if (!h || alpacc_copy_to_host(h, d, 3 * cap * sizeof(uint32_t), NULL) != 0)
        return fail("download failed", ctx);
      for (uint64_t i = 0; i < T; ++i) {
        if (readme)
          printf("%s(#%s, (%u, %u))", i ? ",\n       " : "#some [", TERMINAL_NAMES[h[i]], h[cap + i],
                 h[2 * cap + i]);
        else if (names)
          printf("%u %u %s\n", h[cap + i], h[2 * cap + i], TERMINAL_NAMES[h[i]]);
        else
          printf("%u %u %u\n", h[cap + i], h[2 * cap + i], h[i]);
      }
      if (readme) printf(T ? "]\n" : "#some []\n");
      free(h);
    }
    alpacc_device_free(d);
  } else {
    uint64_t cap = n + 16, N = 0;
    for (;;) {
      uint32_t *d = (uint32_t *)alpacc_device_alloc(4 * cap * sizeof(uint32_t));
      if (!d) return fail("device allocation failed", ctx);
      rc = alpacc_parse(ctx, d_in, n, d, d + cap, d + 2 * cap, cap, &N, NULL);
      if (rc == ALPACC_ERR_CAPACITY) {
        alpacc_device_free(d);
        cap = N;
        continue;
      }
      if (rc < 0) return fail("parse failed", ctx);
      if (rc == ALPACC_ACCEPT) {
        uint32_t *h = (uint32_t *)malloc(4 * cap * sizeof(uint32_t));
        if (!h || alpacc_copy_to_host(h, d, 4 * cap * sizeof(uint32_t), NULL) != 0)
          return fail("download failed", ctx);
        const uint32_t *par = h, *node = h + cap, *span = h + 2 * cap;
        for (uint64_t i = 0; i < N; ++i) {
          const uint32_t id = node[i] & ~ALPACC_NODE_TERMINAL;
          if (readme) {
            printf("%s(%u, ", i ? ",\n       " : "#some [", par[i]);
            if (node[i] & ALPACC_NODE_TERMINAL)
              printf("#terminal (#%s) (%u, %u))", TERMINAL_NAMES[id], span[2 * i], span[2 * i + 1]);
            else
              printf("#production (#%s))", PRODUCTION_NAMES[id]);
          } else if (node[i] & ALPACC_NODE_TERMINAL) {
            if (names)
              printf("%u T %s %u %u\n", par[i], TERMINAL_NAMES[id], span[2 * i], span[2 * i + 1]);
            else
              printf("%u T %u %u %u\n", par[i], id, span[2 * i], span[2 * i + 1]);
          } else if (names) {
            printf("%u P %s\n", par[i], PRODUCTION_NAMES[id]);
          } else {
            printf("%u P %u\n", par[i], id);
          }
        }
        if (readme) printf(N ? "]\n" : "#some []\n");
        free(h);
      }
      alpacc_device_free(d);
      break;
    }
  }
  if (readme && rc != ALPACC_ACCEPT) printf("#none\n");
  fflush(stdout);
  alpacc_device_free(d_in);
  alpacc_destroy(ctx);
  free(in);
  return rc == ALPACC_ACCEPT ? 0 : -1;
}
