/*
 * alpacc_hip.h - C ABI of the MI355X (gfx950) lexer/parser backend for alpacc grammars.
 *
 * The reference exposes its runtime as the entries of a generated Futhark program
 * (src/Alpacc/Generator/Futhark/Generator.hs:19-92):
 *   lex_int   : []u8 -> ([]terminal_int, []i64, []i64), empty arrays on rejection
 *   parse_int : []u8 -> opt [](i64, node terminal_int production_int)
 * built from the table bundle of src/Alpacc/Generator/Analyzer.hs:33-82 and the generic runtime
 * futhark/lexer.fut + futhark/parser.fut. This header replaces those entries with plain C calls:
 * tables go in once (alpacc_create), documents are device buffers, outputs are caller-allocated
 * device buffers, and rejection is an explicit return value instead of an empty/none result.
 *
 * Return codes: 1 = accepted, 0 = rejected (the reference's #none), < 0 = error.
 * Device pointers are HIP device (or host-coherent) addresses; `stream` is a hipStream_t (or NULL
 * for the default stream). Calls block until their results are known. One context may be used by
 * one host thread at a time.
 */
#ifndef ALPACC_HIP_H
#define ALPACC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ALPACC_ACCEPT 1
#define ALPACC_REJECT 0
#define ALPACC_ERR_ARG (-1)
#define ALPACC_ERR_HIP (-2)
#define ALPACC_ERR_CAPACITY (-3)
#define ALPACC_ERR_TIMEOUT (-4)
#define ALPACC_ERR_NO_PARSER (-5)

/* Node word of a parse tree: bit 31 set = terminal node (low bits = terminal id), else a
 * production node (low bits = production id). */
#define ALPACC_NODE_TERMINAL 0x80000000u

/* The generated tables of one grammar (the reference's Analyzer record, Analyzer.hs:33-82),
 * in the device layouts described in alpacc_amd/gen/tables.py. Host pointers; alpacc_create
 * copies them. Parser fields may be zero/NULL for a lexer-only bundle. A caller compiled against
 * an older copy of this header passes a shorter struct: rebuild callers with the header. */
typedef struct alpacc_tables {
  /* lexer (Analyzer.hs:33-41, futhark/lexer.fut:8-27) */
  uint32_t num_endos;        /* e: endomorphisms (identity = 0, dead = 1) */
  uint32_t num_classes;      /* C: byte classes */
  uint32_t num_states;       /* m + 1: lexer DFA states plus the dead state 0 */
  uint32_t init_state;       /* DFA initial state (in 1..m) */
  int32_t ignore_terminal;   /* -1 when the grammar has no ignore token */
  uint32_t dead_terminal;
  const uint8_t *byte_class;  /* [256] */
  const uint16_t *endo_step;  /* [e*C]  endo a, then one byte of class c */
  const uint16_t *endo_comp;  /* [e*e]  endo a then endo b, row a */
  const uint8_t *endo_apply;  /* [e*(m+1)] endo a applied to state s */
  const uint16_t *dfa_step;   /* [(m+1)*C] next | produce<<8 | label(next)<<9 */
  const uint8_t *state_accept;/* [m+1] */
  /* parser (Analyzer.hs:52-66, futhark/parser.fut:9-33) */
  uint32_t q, k;
  uint32_t num_terminals;    /* terminal symbol ids are [0, num_terminals) */
  uint32_t start_terminal, end_terminal, empty_terminal;
  uint32_t num_productions;
  const int32_t *arities;          /* [P] */
  const int32_t *prod_to_terminal; /* [P], -1 = not a terminal production */
  uint64_t pair_index_len;         /* num_terminals^(q+k), or 0 when the table is hashed */
  const uint32_t *pair_index;      /* dense pair table: entry id or 0xFFFFFFFF */
  uint32_t num_pair_entries;
  const uint32_t *pair_entry;      /* [E*4] stack start/end, production start/end */
  uint32_t stacks_len;
  const uint16_t *stacks;          /* brackets: symbol | 0x8000 for left */
  uint32_t productions_len;
  const uint16_t *productions;
  /* the pair table hashed instead of dense (pair_index_len 0, pair_index NULL), for grammars with
   * more than 2^26 possible key windows: pair_hash_len slots (a power of two), each the dense
   * index of a window (sum of terminal * num_terminals^position, first terminal most significant)
   * or 0xFFFFFFFFFFFFFFFF when free, and its entry id; a window's home slot is
   * (index * 0x9E3779B97F4A7C15 mod 2^64) >> (64 - log2(pair_hash_len)), linear probing. */
  uint64_t pair_hash_len;
  const uint64_t *pair_hash_keys;
  const uint32_t *pair_hash_vals;
} alpacc_tables;

typedef struct alpacc_ctx alpacc_ctx;

/* Upload a table bundle to `device`; NULL on failure (see alpacc_create_error). */
alpacc_ctx *alpacc_create(const alpacc_tables *t, int device);
const char *alpacc_create_error(void);
void alpacc_destroy(alpacc_ctx *ctx);
const char *alpacc_last_error(alpacc_ctx *ctx);

/* lex_int (futhark/lexer.fut:174-178; entry Generator.hs:70-75), unchunked semantics.
 * in: device bytes [n]; outputs: device u32 [cap] each (cap >= n suffices). On return
 * *num_tokens holds the token count (also when rejected, in which case outputs are undefined). */
int alpacc_lex(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint32_t *d_kinds,
               uint32_t *d_starts, uint32_t *d_ends, uint64_t cap, uint64_t *num_tokens,
               void *stream);

/* Allocation-free form of alpacc_lex, for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It has no counterpart
 * in the reference. Same arguments, semantics and results as alpacc_lex, for documents of at most
 * ALPACC_NOALLOC_MAX_BYTES; larger ones return ALPACC_ERR_ARG before any GPU work. Its scratch
 * (about 4.3 MiB) is one static device array per GPU, held until the process ends, so the call
 * allocates nothing (no hipMalloc, hipFree or allocating library routine). Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other threads, other contexts)
 * race and must be serialised by the caller. */
#define ALPACC_NOALLOC_MAX_BYTES (64ull << 20)
int alpacc_lex_noalloc(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint32_t *d_kinds,
                       uint32_t *d_starts, uint32_t *d_ends, uint64_t cap, uint64_t *num_tokens,
                       void *stream);

/* parse_int of a lexer+parser artifact (Generator.hs:35-44: lex_int then parser.parse_int,
 * futhark/parser.fut:316-336). Outputs device u32 arrays: parent[cap], node[cap], span[2*cap]
 * (start, end; zeros for production nodes). If cap is too small the call returns
 * ALPACC_ERR_CAPACITY with *num_nodes set to the size needed. */
int alpacc_parse(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint32_t *d_parent,
                 uint32_t *d_node, uint32_t *d_span, uint64_t cap, uint64_t *num_nodes,
                 void *stream);

/* Parser stage alone on a device token stream (parse_int of futhark/parser.fut:334-336):
 * kinds/starts/ends are the lexer outputs of T tokens. */
int alpacc_parse_tokens(alpacc_ctx *ctx, const uint32_t *d_kinds, const uint32_t *d_starts,
                        const uint32_t *d_ends, uint64_t num_tokens, uint32_t *d_parent,
                        uint32_t *d_node, uint32_t *d_span, uint64_t cap, uint64_t *num_nodes,
                        void *stream);

/* The reference's chunk size: its generated entries call lexer.lex_int 16777216
 * (src/Alpacc/Generator/Futhark/Generator.hs:24-76). */
#define ALPACC_REFERENCE_CHUNK 16777216ull

/* lex_int with the reference's chunk loop (lex_int_flag, futhark/lexer.fut:144-172) over chunks of
 * chunk_bytes: the byte shared by neighbouring chunks is run through the DFA twice, the element that
 * ends each chunk clobbers the last token of the chunks before it, and spans carry the shifts the
 * loop applies; starts/ends are therefore 64-bit (i64 in the reference). Equals alpacc_lex (as
 * 64-bit spans) whenever n <= chunk_bytes. Outputs: device kinds u32 [cap], starts/ends u64 [cap],
 * cap >= n + ceil(n / chunk_bytes). */
int alpacc_lex_chunked(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint64_t chunk_bytes,
                       uint32_t *d_kinds, uint64_t *d_starts, uint64_t *d_ends, uint64_t cap,
                       uint64_t *num_tokens, void *stream);

/* parse_int of the reference's lexer+parser artifact exactly as generated (Generator.hs:24-44:
 * lexer.lex_int 16777216 then parser.parse_int) for chunk_bytes = ALPACC_REFERENCE_CHUNK: the
 * chunked lexer followed by the parser. span is u64 [2*cap]. */
int alpacc_parse_chunked(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint64_t chunk_bytes,
                         uint32_t *d_parent, uint32_t *d_node, uint64_t *d_span, uint64_t cap,
                         uint64_t *num_nodes, void *stream);

/* Per-document result of a batched call: status ALPACC_ACCEPT / ALPACC_REJECT, the document's
 * first output slot (token slot or tree node) and its count (0 when rejected). */
typedef struct alpacc_doc_result {
  int32_t status;
  uint32_t reserved;
  uint64_t offset;
  uint64_t count;
} alpacc_doc_result;

/* Batched lex_int over num_docs documents of one device buffer (document d = bytes
 * [offsets[d], offsets[d] + lengths[d]) of d_in; host arrays), each lexed on its own with the
 * unchunked semantics. Document d's tokens are written from slot sum_{d' < d}(lengths[d'] + 1) of
 * d_kinds/d_starts/d_ends (u32, positions relative to the document); cap >= sum(lengths + 1).
 * results[d] (host) receives status, first slot and token count. Returns ALPACC_ACCEPT when the
 * batch ran (per-document verdicts are in results), < 0 on error. There is no counterpart entry in
 * the reference: it replaces one lex_int call per document. */
int alpacc_lex_batch(alpacc_ctx *ctx, const uint8_t *d_in, const uint64_t *offsets,
                     const uint64_t *lengths, uint32_t num_docs, uint32_t *d_kinds,
                     uint32_t *d_starts, uint32_t *d_ends, uint64_t cap,
                     alpacc_doc_result *results, void *stream);

/* Batched parse_int (lexer + parser) over num_docs documents, as alpacc_lex_batch. The trees are
 * written one after the other to d_parent/d_node/d_span (u32, span u32x2); results[d].offset is
 * the document's first node, parents are relative to it (each document's first node is its own
 * parent) and spans relative to the document. *total_nodes receives the nodes written (also the
 * size needed when the call returns ALPACC_ERR_CAPACITY). Replaces one parse_int call per
 * document of the reference's generated entry (Generator.hs:35-44). */
int alpacc_parse_batch(alpacc_ctx *ctx, const uint8_t *d_in, const uint64_t *offsets,
                       const uint64_t *lengths, uint32_t num_docs, uint32_t *d_parent,
                       uint32_t *d_node, uint32_t *d_span, uint64_t cap, uint64_t *total_nodes,
                       alpacc_doc_result *results, void *stream);

/* The `test` entry of the generated artifact (futhark/test.fut:31-180, entries Generator.hs:19-92):
 * blob = u64 count, then per input u64 length + bytes (or, for a parser-only bundle, u64 count of
 * terminals + that many u64 terminals); all integers big-endian. The output, allocated here (free
 * with alpacc_free), is u64 count, then per input one byte accepted and, when accepted:
 *   lexer bundle     : u64 T, T x (u64 kind, start, end)               (lexer_test)
 *   lexer + parser   : u64 N, N x (u8 tag, u64 parent, id, start, end)  (lexer_parser_test)
 *   parser only      : u64 N, N x u64 production                        (parser_test)
 * Lexing uses the reference entries' 16 MiB chunk. */
int alpacc_test(alpacc_ctx *ctx, const uint8_t *blob, uint64_t size, uint8_t **out,
                uint64_t *out_size, void *stream);
void alpacc_free(void *ptr);

/* Number of kernels the last call on `ctx` launched. */
int alpacc_last_launch_count(alpacc_ctx *ctx);

/* Testing hook (no counterpart in the reference): with on != 0, every lexer pass of the following
 * calls on `ctx` raises its look-back time-out flag on the device right after clearing it, as a
 * bounded look-back wait that gave up would. Every kernel still runs to its normal end; the call
 * must then return ALPACC_ERR_TIMEOUT instead of results built on a wrong prefix. */
int alpacc_test_force_timeout(alpacc_ctx *ctx, int on);

/* Launch choices of a context (no counterpart in the reference, whose generated code fixes them
 * at generation time: cuda/lexer.cu:1-48 derives its items per thread from the shared memory per
 * block). alpacc_amd/autotune.py times candidates per grammar and stores the best.
 *   graph_mode        : batched lex + parse (alpacc_parse_batch). 0 = queue the kernels directly;
 *                       1 (default) = capture the call's queue into a HIP graph when its plan
 *                       (sizes, buffers, flags) repeats the previous call's, and replay that graph
 *                       while it keeps repeating; 2 = capture every new plan.
 *   lex_blocks_per_cu : persistent lex_reduce workgroups per CU (1..8, default 2)
 *   keys_blocks_per_cu: persistent parse_keys2 workgroups per CU (1..8, default 2)
 *   search_max_blocks : largest grid of bracket_match / parse_parents, in workgroups of 256
 *                       threads (64..2^20, default 16384)
 *   search_block_threads: workgroup size of bracket_match / parse_parents (256, 512 or 1024;
 *                       default 256)
 *   brackets_per_kib  : brackets per KiB of input the batched bracket scratch is first sized for
 *                       (1..2^20, default 256); a batch with more runs again sized exactly, and
 *                       the estimate grows to what calls needed
 *   lexer_tables_in_lds: 1 (default) = the lexer's tables in LDS where they fit (the
 *                       (endo x class) table up to 160 KiB, the DFA table up to 4096 entries),
 *                       read from L2 otherwise; 0 = every lexer table read from L2
 * graph_launches / graph_captures count the context's batched calls that ran as a graph and the
 * graphs captured (read by alpacc_get_tuning, ignored by alpacc_set_tuning). */
typedef struct alpacc_tuning {
  int32_t graph_mode;
  uint32_t lex_blocks_per_cu;
  uint32_t keys_blocks_per_cu;
  uint32_t search_max_blocks;
  uint32_t brackets_per_kib;
  uint32_t search_block_threads;
  int32_t lexer_tables_in_lds;
  uint32_t reserved;
  uint64_t graph_launches;
  uint64_t graph_captures;
} alpacc_tuning;
int alpacc_set_tuning(alpacc_ctx *ctx, const alpacc_tuning *t);
int alpacc_get_tuning(alpacc_ctx *ctx, alpacc_tuning *t);

/* Device memory a context holds: its tables and its grow-only scratch arenas (the serving loop
 * admits requests by it). */
uint64_t alpacc_device_bytes(alpacc_ctx *ctx);

/* Element counts of the last parser call on `ctx`: counts[0] tokens, [1] key positions (T + 2;
 * tile-padded in batched calls), [2] brackets written for the bracket search (those left after
 * neighbouring positions' matching brackets cancel), [3] productions (tree nodes). */
int alpacc_last_counts(alpacc_ctx *ctx, uint64_t *counts);

/* Device-time profiling: when enabled, every call records HIP events on its stream around each
 * stage (lex_reduce, lex_scan, lex_emit, parse, and inside parse: parse_keys, parse_expand,
 * bracket_match, parse_parents). alpacc_profile_read fills up to max_stages
 * names, accumulated milliseconds and call counts since the last read, resets them, and returns
 * the number of stages filled. */
int alpacc_profile_enable(alpacc_ctx *ctx, int on);
int alpacc_profile_read(alpacc_ctx *ctx, const char **names, double *ms, uint64_t *calls,
                        int max_stages);

/* Device plumbing for hosts that have no HIP binding of their own (the Python host, ctypes
 * callers, FFI callers). They share this library's HIP runtime instance, which a caller's
 * separately loaded HIP runtime would not. All return 0 / non-NULL on success. */
int alpacc_device_count(void);
int alpacc_set_device(int device);
void *alpacc_device_alloc(uint64_t bytes);
void alpacc_device_free(void *ptr);
void *alpacc_host_alloc(uint64_t bytes); /* pinned host memory */
void alpacc_host_free(void *ptr);
int alpacc_copy_to_device(void *dst, const void *src, uint64_t bytes, void *stream);
int alpacc_copy_to_host(void *dst, const void *src, uint64_t bytes, void *stream);
int alpacc_copy_device(void *dst, const void *src, uint64_t bytes, void *stream); /* async */
int alpacc_device_memset(void *dst, int value, uint64_t bytes, void *stream);
void *alpacc_stream_create(void);
void alpacc_stream_destroy(void *stream);
int alpacc_stream_sync(void *stream);
void *alpacc_event_create(void);
void alpacc_event_destroy(void *event);
int alpacc_event_record(void *event, void *stream);
int alpacc_event_sync(void *event); /* waits until the recorded work is done */
/* milliseconds between two recorded events (waits for `end`); negative on error */
float alpacc_event_elapsed_ms(void *start, void *end);

#ifdef __cplusplus
}
#endif

#endif /* ALPACC_HIP_H */
