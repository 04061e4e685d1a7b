// alpacc_hip.hip - MI355X (gfx950, CDNA4, wave64) lexer and LLP parser for alpacc grammars.
//
// Device path for the runtime of the reference (futhark/lexer.fut, futhark/parser.fut):
//
//   lexer : three phases, the input read twice.
//           lex_reduce - every wavefront owns a 2 KiB tile (64 lanes x 32 bytes); each lane walks
//             its bytes through the (endomorphism x byte class) table in LDS, and a wave64
//             Blelloch scan composes the 64 lane endomorphisms with the (endo x endo) table.
//           lex_scan_* / lex_tile_states - exclusive composition of the tile endomorphisms and
//             the DFA state at every tile start (n / 2048 elements).
//           lex_emit - lanes walk the lexer DFA from their known start states once (one walk
//             word per (state, class): next row, kept / boundary / accept bits, label),
//             scan (token count, last token start) across the block and, by a decoupled
//             look-back, across blocks, and write (kind, start, end) of every non-ignored token
//             through LDS staging, coalesced.
//   parser: one scan at the level of key positions. parse_keys2 maps every window of q + k token
//           kinds to its pair entry; pscan_* scan the entries' totals (brackets, productions,
//           depth, lowest depth, sum of arity - 1); parse_expand2 writes every position's
//           brackets and productions (load-balanced, with two levels of adjacent bracket
//           cancellation) and the tree's node and span words; bracket_match and parse_parents
//           run a wave-cooperative "previous smaller-or-equal" search over 64-element chunks and
//           a 64-ary min hierarchy.
//
// All arithmetic is integer. Tables are produced by alpacc_amd/gen (the restated generator).

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/alpacc_hip.h"

namespace {

constexpr int WAVE = 64;
constexpr int LEX_IPT = 32;                      // bytes per lane
constexpr int LEX_TILE = WAVE * LEX_IPT;         // bytes per wavefront tile
constexpr int LEX_BLOCK = 1024;                  // 16 wavefronts share one LDS table copy
constexpr uint32_t ST_AGG = 1u << 30;
constexpr uint32_t ST_INC = 2u << 30;
constexpr uint64_t CT_AGG = 1ull << 62;
constexpr uint64_t CT_INC = 2ull << 62;
constexpr uint32_t SPIN_LIMIT = 1u << 22;        // bounded waits: a stuck look-back ends the kernel
constexpr int LEX_CT_WAVES = LEX_BLOCK / WAVE;   // wave tiles per count tile (one block)
constexpr int LEX_SCAN_BLOCK = 1024;             // wave tiles per K2a scan block
constexpr int LEX_MAX_DFA = 4096;                // max (m+1)*C entries of the DFA table in LDS
constexpr int LEX_STAGE = 384;                   // tokens staged per wavefront and round

// ----------------------------------------------------------------------------------------------
// small device helpers

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_relaxed64(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed64(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Cross-lane steps through DPP row operations (register-to-register) instead of ds_bpermute.
// update_dpp(old, src, ctrl, row_mask, bank_mask, bound_ctrl = false): a lane whose source lies
// outside its 16-lane row, or whose row the mask disables, keeps `old` (the scan identity).
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114,
              DPP_ROW_SHR8 = 0x118, DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143,
              DPP_WAVE_SHR1 = 0x138, DPP_WAVE_SHL1 = 0x130;

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int32_t dpp_i32(int32_t old, int32_t v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xF, false);
}

// Inclusive scan in lane order: Hillis-Steele inside the rows (shr 1, 2, 4, 8), then row 0's
// total into row 1 and row 2's into row 3 (bcast15), then rows 0-1 into rows 2-3 (bcast31).
// `op(earlier, later)` may be non-commutative; `id` is its identity.
template <class Op>
__device__ __forceinline__ int32_t wave_incl_i32(int32_t v, int32_t id, Op op) {
  v = op(dpp_i32<DPP_ROW_SHR1>(id, v), v);
  v = op(dpp_i32<DPP_ROW_SHR2>(id, v), v);
  v = op(dpp_i32<DPP_ROW_SHR4>(id, v), v);
  v = op(dpp_i32<DPP_ROW_SHR8>(id, v), v);
  v = op(dpp_i32<DPP_ROW_BCAST15, 0xA>(id, v), v);
  v = op(dpp_i32<DPP_ROW_BCAST31, 0xC>(id, v), v);
  return v;
}

// the value of the lane before (lane 0 gets id)
__device__ __forceinline__ int32_t wave_prev_i32(int32_t v, int32_t id) {
  return dpp_i32<DPP_WAVE_SHR1>(id, v);
}

struct OpAdd {
  __device__ int32_t operator()(int32_t a, int32_t b) const { return a + b; }
};
struct OpMax {
  __device__ int32_t operator()(int32_t a, int32_t b) const { return max(a, b); }
};
struct OpMin {
  __device__ int32_t operator()(int32_t a, int32_t b) const { return min(a, b); }
};

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readlane(wave_incl_i32((int32_t)v, 0, OpAdd()), WAVE - 1);
}
__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
  return __builtin_amdgcn_readlane(wave_incl_i32(v, INT32_MIN, OpMax()), WAVE - 1);
}
__device__ __forceinline__ int32_t wave_min_i32(int32_t v) {
  return __builtin_amdgcn_readlane(wave_incl_i32(v, INT32_MAX, OpMin()), WAVE - 1);
}

// orders LDS writes of some lanes before reads by other lanes of the same wavefront
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct LexArgs {
  const uint8_t *in;
  uint64_t n;
  uint32_t *kinds;
  uint8_t *kinds8;            // when set (unchunked): token kinds as bytes here instead of kinds
  uint32_t *starts;
  uint32_t *ends;
  const uint8_t *byte_class;
  const uint16_t *endo_step;
  const uint16_t *endo_comp;
  const uint8_t *endo_apply;
  const uint32_t *dfa_walk;   // [(m+1)*C] lex_emit's walk words (laid out where alpacc_create
                              // builds them from dfa_step)
  uint32_t e, C, m1, init;
  int32_t ignore;
  uint32_t num_tiles;         // wave tiles of LEX_TILE bytes
  uint32_t num_ctiles;        // count tiles of LEX_CT_WAVES wave tiles
  uint32_t aligned;           // input pointer is 16-byte aligned
  uint16_t *lane_excl;        // [num_tiles*64] exclusive endomorphism of each lane in its wave tile
  uint16_t *tile_agg;         // [num_tiles] endomorphism of each wave tile
  uint16_t *tile_excl;        // [num_tiles] exclusive endomorphism within its scan block
  uint16_t *block_agg;        // [num scan blocks]
  uint8_t *tile_state;        // [num_tiles] lexer DFA state at the start of each wave tile
  uint32_t *tile_counter;     // [1] dynamic count-tile index
  uint64_t *st_count;         // [num_ctiles] look-back words of the count scan
  uint32_t *result;           // [0] accept, [1] tokens, [2] timeout flag
  // chunked mode (the reference's chunk loop, lexer.fut:144-172): `in` is the virtual string in
  // which every chunk [j*chunk, min(j*chunk + chunk + 1, n_orig)) is laid end to end, so the byte
  // shared by two neighbouring chunks appears twice.
  uint32_t chunk;             // chunk size in bytes of the original string
  uint32_t nchunks;           // K = ceil(n_orig / chunk)
  uint64_t n_orig;
  uint64_t *starts64;         // [cap] token starts as the reference computes them (i64 there)
  uint64_t *ends64;
  uint64_t *chunk_info;       // [K*4] per chunk: tokens through its end, kind, start, end of its
                              // last element (the element that the next chunk's scatter clobbers)
  // batched mode: many documents in one pass. Every document starts a new wave tile (tiles are
  // numbered through the documents in order), the DFA starts afresh and token positions restart
  // at 0 in every document, and document d's tokens go to slots [tok_base, tok_base + len + 1).
  const struct LexDoc *docs;  // [ndocs]
  const uint32_t *tile_doc;   // [num_tiles] document of every wave tile
  const uint4 *tile_rec;      // [num_tiles] the document's byte offset (x, y), length, tile in it
  const uint2 *chunks;        // [nchunks] first wave tile, tile count of every scan chunk
  uint32_t ndocs;
  uint32_t *doc_res;          // [2*ndocs] accept flag, token count
};

struct LexDoc {
  uint64_t offset;    // first byte in `in`
  uint64_t len;       // bytes
  uint64_t tok_base;  // first token slot
  uint32_t tile_base; // first wave tile
  uint32_t ntiles;
  uint32_t chunk_base;  // first scan chunk (LEX_SCAN_BLOCK wave tiles of this document)
  uint32_t nchunks;
};

// The bytes a wave tile covers: document base pointer and length, the tile's index inside its
// document, whether it is the document's first tile.
struct TileView {
  const uint8_t *ptr;
  uint64_t n;
  uint64_t tok_base;
  uint32_t doc;
  uint32_t ltile;
  bool aligned;
};

template <bool BATCH>
__device__ __forceinline__ TileView tile_view(const LexArgs &p, uint32_t tile) {
  TileView v;
  if (!BATCH) {
    v.ptr = p.in;
    v.n = p.n;
    v.tok_base = 0;
    v.doc = 0;
    v.ltile = tile;
    v.aligned = p.aligned != 0;
  } else {
    // the input loads wait on one record (offset, length, tile in the document); the document's
    // index and token base are only needed later, when the tokens are written
    const uint4 r = p.tile_rec[tile];
    const uint32_t d = p.tile_doc[tile];
    v.ptr = p.in + (((uint64_t)r.y << 32) | r.x);
    v.n = r.z;
    v.tok_base = p.docs[d].tok_base;
    v.doc = d;
    v.ltile = r.w;
    v.aligned = ((uintptr_t)v.ptr & 15) == 0;
  }
  return v;
}

// Span of a token of the chunked mode: the token ends at virtual position v, its start is the
// virtual position w. lex_step (lexer.fut:91-142) of chunk j writes (start - j*chunk, i) with i the
// chunk-local position, and the map at :131-133 of every later chunk j' >= j adds j'*chunk to the
// start and 1 + j'*chunk to the end.
__device__ __forceinline__ void chunk_span(const LexArgs &p, uint32_t w, uint32_t v, uint64_t &s,
                                           uint64_t &e) {
  const uint32_t cp1 = p.chunk + 1;
  const uint32_t j = v / cp1;
  const uint32_t i = v - j * cp1;
  const uint64_t orig_w = (uint64_t)(w - w / cp1);
  const uint64_t K = p.nchunks;
  const uint64_t shift = (uint64_t)p.chunk * (((K - 1 + j) * (K - j)) / 2);
  s = orig_w + shift - (uint64_t)j * p.chunk;
  e = (uint64_t)i + (K - j) + shift;
}

__device__ __forceinline__ uint32_t comp_lookup(const uint16_t *tab, uint32_t e, uint32_t a,
                                                uint32_t b) {
  return tab[a * e + b];
}

#define LEX_CLS(j) ((cls[(j) >> 2] >> (8 * ((j)&3))) & 0xFFu)

// bytes of one lane -> byte classes, four per register
__device__ __forceinline__ void classes_of_words(const uint32_t (&wds)[8], const uint8_t *s_cls,
                                                 uint32_t (&cls)[LEX_IPT / 4]) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const uint32_t x = wds[q];
    cls[q] = (uint32_t)s_cls[x & 0xFF] | ((uint32_t)s_cls[(x >> 8) & 0xFF] << 8) |
             ((uint32_t)s_cls[(x >> 16) & 0xFF] << 16) | ((uint32_t)s_cls[x >> 24] << 24);
  }
}

// The byte classes of in[base .. base + valid) (n: the document length), class 0 past valid. A
// document that starts off a 16-byte boundary is read through the aligned 16-byte words around
// each lane's 32 bytes (the offset is the same for all lanes of a document). A document's first
// and last lanes read the same words, the last one clamped to the word holding the document's
// last byte: an aligned 16-byte word that holds a byte of the document lies inside one mapped
// page, so these loads cannot fault, and the bytes they read outside the document are masked.
__device__ __forceinline__ void load_classes(const uint8_t *in, bool aligned, const uint8_t *s_cls,
                                             uint64_t base, int valid, uint32_t (&cls)[LEX_IPT / 4],
                                             uint64_t n) {
  if (valid == LEX_IPT && aligned) {
    const uint4 a = *reinterpret_cast<const uint4 *>(in + base);
    const uint4 b = *reinterpret_cast<const uint4 *>(in + base + 16);
    const uint32_t wds[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    classes_of_words(wds, s_cls, cls);
  } else if (valid == LEX_IPT && base >= 16 && base + LEX_IPT + 16 <= n) {
    const uint8_t *q = in + base;
    const uint32_t off = (uint32_t)((uintptr_t)q & 15);
    const uint4 *a4 = reinterpret_cast<const uint4 *>(q - off);  // inside [in, in + n)
    const uint4 x0 = a4[0], x1 = a4[1], x2 = a4[2];
    const uint32_t w[12] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w,
                            x2.x, x2.y, x2.z, x2.w};
    const uint32_t sh = off & 3;
    uint32_t wds[8];
    // word r of the window is bytes 4r .. 4r+3 from q: words off/4 + r and + r + 1, shifted
#define LEX_ALIGN_CASE(W0)                                                              \
  _Pragma("unroll") for (int r = 0; r < 8; ++r) wds[r] =                                \
      __builtin_amdgcn_alignbyte(w[(W0) + r + 1], w[(W0) + r], sh);
    switch (off >> 2) {
      case 0: LEX_ALIGN_CASE(0) break;
      case 1: LEX_ALIGN_CASE(1) break;
      case 2: LEX_ALIGN_CASE(2) break;
      default: LEX_ALIGN_CASE(3) break;
    }
#undef LEX_ALIGN_CASE
    classes_of_words(wds, s_cls, cls);
  } else if (valid > 0) {
    // a document's first or last lanes: three word loads in flight together, as in the interior
    // (byte loads, or a guarded load per byte, hold up the wave tile's count tile, and with it
    // every count tile that looks back to it)
    const uint8_t *q = in + base;
    const uint32_t off = (uint32_t)((uintptr_t)q & 15);
    const uint4 *a4 = reinterpret_cast<const uint4 *>(q - off);
    // index (from a4) of the word holding the document's last byte; q is inside the document
    const uint32_t lw = (uint32_t)(((uintptr_t)(in + n - 1) >> 4) - ((uintptr_t)q >> 4));
    const uint4 x0 = a4[0], x1 = a4[min(1u, lw)], x2 = a4[min(2u, lw)];
    const uint32_t w[12] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w,
                            x2.x, x2.y, x2.z, x2.w};
    const uint32_t sh = off & 3;
    uint32_t wds[8];
#define LEX_ALIGN_CASE(W0)                                                              \
  _Pragma("unroll") for (int r = 0; r < 8; ++r) wds[r] =                                \
      __builtin_amdgcn_alignbyte(w[(W0) + r + 1], w[(W0) + r], sh);
    switch (off >> 2) {
      case 0: LEX_ALIGN_CASE(0) break;
      case 1: LEX_ALIGN_CASE(1) break;
      case 2: LEX_ALIGN_CASE(2) break;
      default: LEX_ALIGN_CASE(3) break;
    }
#undef LEX_ALIGN_CASE
    classes_of_words(wds, s_cls, cls);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int rem = valid - 4 * k;  // bytes of word k inside the document
      cls[k] &= rem >= 4 ? 0xFFFFFFFFu : (rem <= 0 ? 0u : (1u << (8 * rem)) - 1u);
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) cls[q] = 0u;
  }
}

// K1: per wave tile, the endomorphism of every lane's bytes (serial walk through the
// (endo x class) table) and their wave-exclusive composition (Blelloch up/down sweep with the
// (endo x endo) table, 126 lookups per wave). Persistent, grid-stride over wave tiles.
// TAB: where the tables sit: 2 = both in LDS, 1 = the (endo x class) table in LDS and the
// composition table in L2, 0 = both in L2 (an (endo x class) table larger than LDS).
template <int TAB, bool BATCH>
__global__ void __launch_bounds__(LEX_BLOCK) __attribute__((amdgpu_waves_per_eu(8)))
lex_reduce(LexArgs p) {
  extern __shared__ __align__(16) uint8_t smem[];
  const uint32_t e = p.e, C = p.C;
  uint16_t *s_step = reinterpret_cast<uint16_t *>(smem);
  uint8_t *s_cls = TAB >= 1 ? reinterpret_cast<uint8_t *>(s_step + e * C) : smem;
  uint16_t *s_comp = reinterpret_cast<uint16_t *>(s_cls + 256);
  if (TAB >= 1)
    for (uint32_t i = threadIdx.x; i < e * C; i += blockDim.x) s_step[i] = p.endo_step[i];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_cls[i] = p.byte_class[i];
  if (TAB == 2)
    for (uint32_t i = threadIdx.x; i < e * e; i += blockDim.x) s_comp[i] = p.endo_comp[i];
  __syncthreads();
  const uint16_t *comp_tab = TAB == 2 ? s_comp : p.endo_comp;
  const uint16_t *step_tab = TAB >= 1 ? s_step : p.endo_step;
  const int lane = threadIdx.x & (WAVE - 1);
  const uint32_t wpb = blockDim.x / WAVE;
  const uint32_t stride = gridDim.x * wpb;
  // two tiles per iteration: their table-lookup chains are independent and overlap
  for (uint32_t t0 = blockIdx.x * wpb + threadIdx.x / WAVE; t0 < p.num_tiles; t0 += 2 * stride) {
    uint32_t tile[2] = {t0, t0 + stride};
    uint32_t v[2], total[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      uint32_t cls[LEX_IPT / 4];
      int valid = 0;
      if (tile[h] < p.num_tiles) {
        const TileView tv = tile_view<BATCH>(p, tile[h]);
        const uint64_t base = (uint64_t)tv.ltile * LEX_TILE + (uint64_t)lane * LEX_IPT;
        valid = base >= tv.n ? 0 : (int)min((uint64_t)LEX_IPT, tv.n - base);
        load_classes(tv.ptr, tv.aligned, s_cls, base, valid, cls, tv.n);
      }
      uint32_t x = 0;
      if (valid == LEX_IPT) {
#pragma unroll
        for (int j = 0; j < LEX_IPT; ++j) x = step_tab[__umul24(x, C) + LEX_CLS(j)];
      } else {
#pragma unroll
        for (int j = 0; j < LEX_IPT; ++j)
          if (j < valid) x = step_tab[x * C + LEX_CLS(j)];
      }
      v[h] = x;
    }
    // up-sweep: lane l with (l+1) % 2d == 0 holds the composition of lanes l-2d+1 .. l
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint32_t left = __shfl(v[h], (lane - d) & (WAVE - 1), WAVE);
        if (((lane + 1) & (2 * d - 1)) == 0) v[h] = comp_lookup(comp_tab, e, left, v[h]);
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      total[h] = __shfl(v[h], WAVE - 1, WAVE);
      if (lane == WAVE - 1) v[h] = 0;
    }
    // down-sweep: right node r gets prefix(r) then sum(left subtree); left node gets prefix(r)
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const uint32_t left = __shfl(v[h], (lane - d) & (WAVE - 1), WAVE);
        const uint32_t right = __shfl(v[h], (lane + d) & (WAVE - 1), WAVE);
        if (((lane + 1) & (2 * d - 1)) == 0)
          v[h] = comp_lookup(comp_tab, e, v[h], left);
        else if (((lane + 1 + d) & (2 * d - 1)) == 0)
          v[h] = right;
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (tile[h] < p.num_tiles) {
        p.lane_excl[(uint64_t)tile[h] * WAVE + lane] = (uint16_t)v[h];
        if (lane == 0) p.tile_agg[tile[h]] = (uint16_t)total[h];
      }
    }
  }
}

// K2a: exclusive composition of the wave-tile endomorphisms inside blocks of LEX_SCAN_BLOCK tiles
// (Blelloch in LDS), the block totals to block_agg.
__global__ void __launch_bounds__(LEX_SCAN_BLOCK) lex_scan_tiles(LexArgs p) {
  __shared__ uint16_t s[LEX_SCAN_BLOCK];
  const uint32_t t = blockIdx.x * LEX_SCAN_BLOCK + threadIdx.x;
  s[threadIdx.x] = t < p.num_tiles ? p.tile_agg[t] : 0;
  __syncthreads();
  const uint32_t i = threadIdx.x;
  for (uint32_t d = 1; d < LEX_SCAN_BLOCK; d <<= 1) {
    uint32_t nv = 0;
    const bool act = ((i + 1) & (2 * d - 1)) == 0;
    if (act) nv = comp_lookup(p.endo_comp, p.e, s[i - d], s[i]);
    __syncthreads();
    if (act) s[i] = (uint16_t)nv;
    __syncthreads();
  }
  if (i == LEX_SCAN_BLOCK - 1) {
    p.block_agg[blockIdx.x] = s[i];
    s[i] = 0;
  }
  __syncthreads();
  for (uint32_t d = LEX_SCAN_BLOCK / 2; d >= 1; d >>= 1) {
    const bool right = ((i + 1) & (2 * d - 1)) == 0;
    uint32_t l = 0, r = 0;
    if (right) {
      l = s[i - d];
      r = s[i];
    }
    __syncthreads();
    if (right) {
      s[i - d] = (uint16_t)r;
      s[i] = (uint16_t)comp_lookup(p.endo_comp, p.e, r, l);
    }
    __syncthreads();
  }
  if (t < p.num_tiles) p.tile_excl[t] = s[i];
}

// Exclusive composition of v[0 .. n) in place by one workgroup of LEX_SCAN_BLOCK threads
// (Blelloch in LDS, chunks with a carry).
__device__ void scan_excl_inplace(const LexArgs &p, uint16_t *v, uint32_t n) {
  __shared__ uint16_t s[LEX_SCAN_BLOCK];
  __shared__ uint16_t carry_s;
  const uint32_t i = threadIdx.x;
  if (i == 0) carry_s = 0;
  for (uint32_t off = 0; off < n; off += LEX_SCAN_BLOCK) {
    s[i] = off + i < n ? v[off + i] : 0;
    __syncthreads();
    for (uint32_t d = 1; d < LEX_SCAN_BLOCK; d <<= 1) {
      uint32_t nv = 0;
      const bool act = ((i + 1) & (2 * d - 1)) == 0;
      if (act) nv = comp_lookup(p.endo_comp, p.e, s[i - d], s[i]);
      __syncthreads();
      if (act) s[i] = (uint16_t)nv;
      __syncthreads();
    }
    uint32_t chunk_total = 0;
    if (i == LEX_SCAN_BLOCK - 1) {
      chunk_total = s[i];
      s[i] = 0;
    }
    __syncthreads();
    for (uint32_t d = LEX_SCAN_BLOCK / 2; d >= 1; d >>= 1) {
      const bool right = ((i + 1) & (2 * d - 1)) == 0;
      uint32_t l = 0, r = 0;
      if (right) {
        l = s[i - d];
        r = s[i];
      }
      __syncthreads();
      if (right) {
        s[i - d] = (uint16_t)r;
        s[i] = (uint16_t)comp_lookup(p.endo_comp, p.e, r, l);
      }
      __syncthreads();
    }
    const uint32_t carry = carry_s;
    if (off + i < n) v[off + i] = (uint16_t)comp_lookup(p.endo_comp, p.e, carry, s[i]);
    __syncthreads();
    if (i == LEX_SCAN_BLOCK - 1) carry_s = (uint16_t)comp_lookup(p.endo_comp, p.e, carry, chunk_total);
    __syncthreads();
  }
}

// K2b: one block, exclusive composition of the block totals in place.
__global__ void __launch_bounds__(LEX_SCAN_BLOCK) lex_scan_blocks(LexArgs p, uint32_t nblocks) {
  scan_excl_inplace(p, p.block_agg, nblocks);
}

// K2c: DFA state at the start of every wave tile.
__global__ void lex_tile_states(LexArgs p) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.num_tiles) return;
  const uint32_t pre = comp_lookup(p.endo_comp, p.e, p.block_agg[t / LEX_SCAN_BLOCK], p.tile_excl[t]);
  p.tile_state[t] = p.endo_apply[pre * p.m1 + p.init];
}

// K2 of the batched mode, spread over many workgroups however long a document is (one workgroup per
// document took 0.6 ms for a 256 MiB one): (a) every chunk of up to LEX_SCAN_BLOCK wave tiles of
// one document, exclusive composition inside the chunk and the chunk's total; (b) per document,
// exclusive composition of its chunk totals in place; (c) the DFA state at the start of every tile.
__global__ void __launch_bounds__(LEX_SCAN_BLOCK) lex_scan_chunks(LexArgs p) {
  __shared__ uint16_t s[LEX_SCAN_BLOCK];
  const uint2 ch = p.chunks[blockIdx.x];
  const uint32_t i = threadIdx.x;
  const bool in = i < ch.y;
  s[i] = in ? p.tile_agg[ch.x + i] : 0;  // 0: the identity
  __syncthreads();
  for (uint32_t d = 1; d < LEX_SCAN_BLOCK; d <<= 1) {
    uint32_t nv = 0;
    const bool act = ((i + 1) & (2 * d - 1)) == 0;
    if (act) nv = comp_lookup(p.endo_comp, p.e, s[i - d], s[i]);
    __syncthreads();
    if (act) s[i] = (uint16_t)nv;
    __syncthreads();
  }
  if (i == LEX_SCAN_BLOCK - 1) {
    p.block_agg[blockIdx.x] = s[i];
    s[i] = 0;
  }
  __syncthreads();
  for (uint32_t d = LEX_SCAN_BLOCK / 2; d >= 1; d >>= 1) {
    const bool right = ((i + 1) & (2 * d - 1)) == 0;
    uint32_t l = 0, r = 0;
    if (right) {
      l = s[i - d];
      r = s[i];
    }
    __syncthreads();
    if (right) {
      s[i - d] = (uint16_t)r;
      s[i] = (uint16_t)comp_lookup(p.endo_comp, p.e, r, l);
    }
    __syncthreads();
  }
  if (in) p.tile_excl[ch.x + i] = s[i];
}

__global__ void __launch_bounds__(LEX_SCAN_BLOCK) lex_scan_doc_chunks(LexArgs p) {
  const LexDoc doc = p.docs[blockIdx.x];
  scan_excl_inplace(p, p.block_agg + doc.chunk_base, doc.nchunks);
}

__global__ void __launch_bounds__(LEX_SCAN_BLOCK) lex_chunk_states(LexArgs p) {
  const uint2 ch = p.chunks[blockIdx.x];
  const uint32_t i = threadIdx.x;
  if (i < ch.y) {
    const uint32_t pre =
        comp_lookup(p.endo_comp, p.e, p.block_agg[blockIdx.x], p.tile_excl[ch.x + i]);
    p.tile_state[ch.x + i] = p.endo_apply[pre * p.m1 + p.init];
  }
}

// Batched mode: the document of every wave tile, and the record its input loads start from.
__global__ void lex_tile_docs(LexArgs p) {
  const LexDoc doc = p.docs[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < doc.ntiles; i += blockDim.x) {
    const_cast<uint32_t *>(p.tile_doc)[doc.tile_base + i] = blockIdx.x;
    const_cast<uint4 *>(p.tile_rec)[doc.tile_base + i] =
        make_uint4((uint32_t)doc.offset, (uint32_t)(doc.offset >> 32), (uint32_t)doc.len, i);
  }
  // and the document's scan chunks (first wave tile, tile count)
  for (uint32_t c = threadIdx.x; c < doc.nchunks; c += blockDim.x)
    const_cast<uint2 *>(p.chunks)[doc.chunk_base + c] =
        make_uint2(doc.tile_base + c * LEX_SCAN_BLOCK, min((uint32_t)LEX_SCAN_BLOCK, doc.ntiles - c * LEX_SCAN_BLOCK));
}

// Exclusive (token count, last token start) of a count tile: decoupled look-back; the two
// components combine by + and max, so the window reduces with plain wave reductions.
__device__ void lookback_count(const LexArgs &p, uint32_t tile, int lane, uint32_t &cnt,
                               int32_t &last) {
  uint32_t csum = 0;
  int32_t lmax = -1;
  int64_t j = (int64_t)tile - 1;
  uint32_t spins = 0;
  while (true) {
    const int64_t idx = j - lane;
    const uint64_t w = idx >= 0 ? ld_relaxed64(p.st_count + idx) : CT_INC;
    const uint32_t flag = (uint32_t)(w >> 62);
    const uint64_t incm = __ballot(flag == 2);
    const uint64_t zerom = __ballot(flag == 0);
    const uint64_t upto = incm ? (((incm & (~incm + 1)) << 1) - 1) : ~0ull;
    if (zerom & upto) {
      // wait on the nearest count tile not yet published, one word per poll, then read the
      // window again: every waiting count tile reading its whole window (and the shared timeout
      // word) at each poll piles requests onto the few lines that hold the recent words, and the
      // loads of the tiles still walking slow down behind them
      const int zl = __ffsll((unsigned long long)(zerom & upto)) - 1;
      const uint64_t *zw = p.st_count + (j - zl);
      bool given_up = false;
      while ((ld_relaxed64(zw) >> 62) == 0) {
        ++spins;
        if (spins > SPIN_LIMIT || ((spins & 63) == 1 && ld_relaxed(p.result + 2) != 0)) {
          given_up = true;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      if (given_up) {
        if (lane == 0) st_relaxed(p.result + 2, 1u);
        cnt = 0;
        last = -1;
        return;
      }
      continue;
    }
    const bool in = (upto >> lane) & 1ull;
    uint32_t c = in ? (uint32_t)(w & 0x7FFFFFFFull) : 0u;
    int32_t l = in ? (int32_t)((w >> 31) & 0x7FFFFFFFull) - 1 : -1;
    csum += wave_sum_u32(c);
    lmax = max(lmax, wave_max_i32(l));
    if (incm) break;
    j -= WAVE;
  }
  cnt = csum;
  last = lmax;
}

__device__ __forceinline__ uint64_t pack_count(uint32_t cnt, int32_t last, uint64_t flag) {
  return flag | ((uint64_t)(uint32_t)(last + 1) << 31) | (uint64_t)cnt;
}

// K3: walk the lexer DFA from the known lane start states, count token ends / find token starts,
// scan them across the count tile (16 wave tiles) and across count tiles (decoupled look-back),
// then write (kind, start, end) of every token. CHUNKED: the input is the virtual string of the
// chunked mode; the last element of every chunk but the last ends no token (i != n - 1 in
// lex_step, lexer.fut:104-107), spans are shifted as lex_step does, and every chunk's last
// element is recorded in chunk_info for lex_chunk_fixup.
// DFA_LDS: the walk words (dfa_walk, (m + 1) x C) are copied to LDS; else they are read from L2.
template <bool CHUNKED, bool BATCH, bool DFA_LDS>
__global__ void __launch_bounds__(LEX_BLOCK) lex_emit(LexArgs p) {
  __shared__ uint32_t s_dfa[LEX_MAX_DFA];
  __shared__ uint8_t s_cls[256];
  __shared__ uint32_t s_wcnt[LEX_CT_WAVES];
  __shared__ int32_t s_wlast[LEX_CT_WAVES];
  __shared__ uint32_t s_wreset[LEX_CT_WAVES];
  __shared__ uint32_t s_ctile;
  __shared__ uint32_t s_cpre;  // the count tile's look-back prefix (count, last start)
  __shared__ int32_t s_lpre;
  __shared__ uint32_t s_stage[LEX_CT_WAVES][2][LEX_STAGE];  // starts, ends
  __shared__ uint8_t s_kind[LEX_CT_WAVES][LEX_STAGE];
  const uint32_t C = p.C;
  if (DFA_LDS)
    for (uint32_t i = threadIdx.x; i < p.m1 * C; i += blockDim.x) s_dfa[i] = p.dfa_walk[i];
  // walk words (the next state's row as a byte offset, kept / boundary / accept bits, label)
  const uint8_t *dfa = reinterpret_cast<const uint8_t *>(DFA_LDS ? s_dfa : p.dfa_walk);
#define LEX_WALK(ROW4, CL) (*reinterpret_cast<const uint32_t *>(dfa + (ROW4) + ((CL) << 2)))
  const uint32_t C4 = C << 2;
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_cls[i] = p.byte_class[i];
  if (threadIdx.x == 0) s_ctile = atomicAdd(p.tile_counter, 1u);
  __syncthreads();
  const uint32_t ct = s_ctile;
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  const uint32_t tile = ct * LEX_CT_WAVES + w;
  const uint32_t ign = p.ignore < 0 ? 0xFFFFFFFFu : (uint32_t)p.ignore;
  // positions are relative to the tile's document
  TileView tv;
  if (tile < p.num_tiles) {
    tv = tile_view<BATCH>(p, tile);
  } else {
    tv.ptr = p.in;
    tv.n = 0;
    tv.tok_base = 0;
    tv.doc = 0;
    tv.ltile = 0;
    tv.aligned = false;
  }
  const uint64_t n = tv.n;
  const uint64_t tbase = (uint64_t)tv.ltile * LEX_TILE;
  const uint64_t base = tbase + (uint64_t)lane * LEX_IPT;
  const int valid = (tile >= p.num_tiles || base >= n) ? 0 : (int)min((uint64_t)LEX_IPT, n - base);
  uint32_t cls[LEX_IPT / 4];
  load_classes(tv.ptr, tv.aligned, s_cls, base, valid, cls, tv.n);
  uint32_t s_lane = 0;
  if (tile < p.num_tiles) {
    const uint32_t s_tile = p.tile_state[tile];
    const uint32_t ex = p.lane_excl[(uint64_t)tile * WAVE + lane];
    s_lane = p.endo_apply[ex * p.m1 + s_tile];
  }
  uint32_t c_next = __shfl_down(LEX_CLS(0), 1, WAVE);
  if (lane == WAVE - 1) {
    const uint64_t nb = tbase + LEX_TILE;
    c_next = nb < n ? s_cls[tv.ptr[nb]] : 0u;
  }

  // ---- one walk: token-end mask, token-start mask and the label after every byte
  uint32_t endm = 0, startm = 0;
  uint32_t labs[LEX_IPT / 4] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t acc_end = 0;  // whether the state after the document's last byte accepts
  uint32_t lab_end = 0;
  // the lane holding its document's last byte (no byte follows it)
  const bool doc_last = valid > 0 && base + (uint64_t)valid == n;
  if (valid > 0) {
    // every lane walks all its 32 classes, a document's last lane too (class 0 past the end, and
    // c_next is 0 there), so that no wave runs a second, bounds-checked walk for one lane; the
    // marks past the document's end are cleared below. x_j, the word that took byte j, gives the
    // label of byte j (bits 24-30) and whether it is kept (bit 18); x_{j+1} whether a token
    // boundary follows byte j (bit 19). A token ends at j where a boundary follows a kept byte and
    // starts at j + 1 where a boundary precedes a kept byte: two bit masks, ANDed after the walk.
    uint32_t x = LEX_WALK(__umul24(s_lane, C4), LEX_CLS(0));
    const uint32_t start0 = (x >> 19) & (x >> 18) & 1u;
    uint32_t pm = 0, km = 0;
    uint32_t xs[4];
#pragma unroll
    for (int j = 0; j < LEX_IPT; ++j) {
      km |= ((x >> 18) & 1u) << j;
      xs[j & 3] = x;
      if ((j & 3) == 3) {
        // the labels of bytes j-3 .. j: byte 3 of their four words
        const uint32_t t01 = __builtin_amdgcn_perm(xs[1], xs[0], 0x0c0c0703u);
        const uint32_t t23 = __builtin_amdgcn_perm(xs[3], xs[2], 0x0c0c0703u);
        labs[j >> 2] = __builtin_amdgcn_perm(t23, t01, 0x05040100u);
      }
      const uint32_t cn = j + 1 < LEX_IPT ? LEX_CLS((j + 1) & (LEX_IPT - 1)) : c_next;
      const uint32_t x2 = LEX_WALK(x & 0x3FFFFu, cn);
      pm |= ((x2 >> 19) & 1u) << j;
      if (j + 1 < LEX_IPT) x = x2;
    }
    endm = pm & km;
    startm = start0 | ((pm << 1) & km);
    acc_end = (x >> 20) & 1u;
    lab_end = x >> 24;
    if (doc_last) {
      // tokens start only inside the document, and its last byte ends no token here (the final
      // token is appended below); its label is the one after byte valid - 1. Whether the state
      // after that byte accepts is walked again once the count tile's total is out (below).
      const uint32_t keep = valid == LEX_IPT ? 0xFFFFFFFFu : (1u << valid) - 1u;
      startm &= keep;
      endm &= keep >> 1;
      uint32_t lw;
      switch ((valid - 1) >> 2) {  // a register array takes constant indexes only
        case 0: lw = labs[0]; break;
        case 1: lw = labs[1]; break;
        case 2: lw = labs[2]; break;
        case 3: lw = labs[3]; break;
        case 4: lw = labs[4]; break;
        case 5: lw = labs[5]; break;
        case 6: lw = labs[6]; break;
        default: lw = labs[7]; break;
      }
      lab_end = (lw >> (8 * ((valid - 1) & 3))) & 0xFFu;
    }
  }
  uint32_t lastm = 0;  // CHUNKED: chunk-last positions of this lane
  if (CHUNKED && valid > 0) {
    const uint32_t cp1 = p.chunk + 1;
    const uint32_t r0 = (uint32_t)(base % cp1);
    for (uint32_t j = p.chunk - r0; j < (uint32_t)valid; j += cp1) lastm |= 1u << j;
    endm &= ~lastm;
  }
  const uint32_t cnt = __popc(endm);
  const int32_t last = startm ? (int32_t)(base + 31 - __clz(startm)) : -1;

  // ---- scans: wave, count tile, across count tiles
  const uint32_t cincl = (uint32_t)wave_incl_i32((int32_t)cnt, 0, OpAdd());
  const int32_t lincl = wave_incl_i32(last, -1, OpMax());
  if (lane == WAVE - 1) {
    s_wcnt[w] = cincl;
    s_wlast[w] = lincl;
    // a wave tile that starts a document restarts the count and the last start
    s_wreset[w] = (tile < p.num_tiles && tv.ltile == 0) ? 1u : 0u;
  }
  __syncthreads();
  uint32_t ct_tot = 0;  // wave 0: the count tile's own total, last start and reset flag
  int32_t ct_last = -1;
  bool ct_reset = false, need_lb = false;
  if (w == 0) {
    // segmented scan of the 16 wave totals (a reset keeps only what follows it)
    const uint32_t wc = lane < LEX_CT_WAVES ? s_wcnt[lane] : 0u;
    const int32_t wl = lane < LEX_CT_WAVES ? s_wlast[lane] : -1;
    const bool wr = lane < LEX_CT_WAVES ? s_wreset[lane] != 0 : false;
    // 16 lanes: one DPP row (shr 1, 2, 4, 8), segmented by the reset flags
    int32_t wci = (int32_t)wc, wli = wl, wfi = wr ? 1 : 0;
#define LEX_SEG_STEP(CTRL)                                  \
    {                                                       \
      const int32_t oc = dpp_i32<CTRL>(0, wci);             \
      const int32_t ol = dpp_i32<CTRL>(-1, wli);            \
      const int32_t of = dpp_i32<CTRL>(0, wfi);             \
      if (!wfi) {                                           \
        wci += oc;                                          \
        wli = max(wli, ol);                                 \
        wfi = of;                                           \
      }                                                     \
    }
    LEX_SEG_STEP(DPP_ROW_SHR1)
    LEX_SEG_STEP(DPP_ROW_SHR2)
    LEX_SEG_STEP(DPP_ROW_SHR4)
    LEX_SEG_STEP(DPP_ROW_SHR8)
#undef LEX_SEG_STEP
    const uint32_t ctot = (uint32_t)__builtin_amdgcn_readlane(wci, LEX_CT_WAVES - 1);
    const int32_t ltot = __builtin_amdgcn_readlane(wli, LEX_CT_WAVES - 1);
    const bool ftot = __builtin_amdgcn_readlane(wfi, LEX_CT_WAVES - 1) != 0;
    // a count tile holding a document start already has its (post-reset) inclusive value
    if (lane == 0) st_relaxed64(p.st_count + ct, pack_count(ctot, ltot, ftot ? CT_INC : CT_AGG));
    ct_tot = ctot;
    ct_last = ltot;
    ct_reset = ftot;
    need_lb = ct > 0 && __builtin_amdgcn_readlane((int)wr, 0) == 0;
    // exclusive value of every wave inside the count tile; a wave with no reset before it (its
    // own included) adds the count tile's look-back prefix, which arrives later
    uint32_t xc = (uint32_t)wave_prev_i32(wci, 0);
    int32_t xl = wave_prev_i32(wli, -1);
    const bool xf = wave_prev_i32(wfi, 0) != 0;
    if (wr) {
      xc = 0;
      xl = -1;
    }
    if (lane < LEX_CT_WAVES) {
      s_wcnt[lane] = xc;
      s_wlast[lane] = xl;
      s_wreset[lane] = (wr || xf) ? 0u : 1u;
    }
  }
  // a document's last lane ending inside its 32 bytes: whether the state after its last byte
  // accepts (the walk above went on over class 0), walked now that the count tile's total is out
  if (doc_last && valid < LEX_IPT) {
    uint32_t x = LEX_WALK(__umul24(s_lane, C4), LEX_CLS(0));
#pragma unroll
    for (int j = 1; j < LEX_IPT - 1; ++j)
      if (j < valid) x = LEX_WALK(x & 0x3FFFFu, LEX_CLS(j));
    acc_end = (x >> 20) & 1u;
  }
  __syncthreads();

  // ---- write tokens: stage in LDS by wave-local index, copy out coalesced, in rounds. The first
  // round is staged before the look-back, so that its wait overlaps the staging; a lane whose
  // first token began before the count tile stages START_PENDING and the copy takes the
  // look-back's last start instead.
  constexpr uint32_t START_PENDING = 0xFFFFFFFFu;  // positions stay below 2^31
  const bool takes_prefix = s_wreset[w] != 0;
  const uint32_t wtotal = (uint32_t)__builtin_amdgcn_readlane((int32_t)cincl, WAVE - 1);
  const uint32_t lofs = cincl - cnt;                // wave-local index of this lane's first token
  const int32_t lexcl = wave_prev_i32(lincl, -1);
  const int32_t last_in_tile = max(s_wlast[w], lexcl);
  uint32_t start_in = last_in_tile >= 0 ? (uint32_t)last_in_tile
                                        : (takes_prefix ? START_PENDING : 0u);
  uint32_t pending_start = 0;
  uint32_t wout = 0;  // slot of the wave's first token
  uint8_t *stk = s_kind[w];
  uint32_t *sts = s_stage[w][0];
  uint32_t *ste = s_stage[w][1];
  for (uint32_t r0 = 0;; r0 += LEX_STAGE) {
    const bool any = r0 < wtotal;  // uniform in the wave
    if (any) {
      // this lane's tokens with wave-local index in [r0, r0 + LEX_STAGE)
      uint32_t em = endm;
      uint32_t k = lofs;
      if (k < r0) {  // skip tokens staged in earlier rounds
        const uint32_t skip = min(r0 - k, (uint32_t)__popc(em));
        for (uint32_t q = 0; q < skip; ++q) em &= em - 1;
        k += skip;
      }
      while (em && k < r0 + LEX_STAGE) {
        const int j = __ffs(em) - 1;
        em &= em - 1;
        const uint32_t sm = startm & (j == 31 ? 0xFFFFFFFFu : ((2u << j) - 1u));
        const uint32_t cur = sm ? (uint32_t)(base + 31 - __clz(sm)) : start_in;
        uint32_t lw;
        switch (j >> 2) {
          case 0: lw = labs[0]; break;
          case 1: lw = labs[1]; break;
          case 2: lw = labs[2]; break;
          case 3: lw = labs[3]; break;
          case 4: lw = labs[4]; break;
          case 5: lw = labs[5]; break;
          case 6: lw = labs[6]; break;
          default: lw = labs[7]; break;
        }
        const uint32_t o = k - r0;
        stk[o] = (uint8_t)(lw >> (8 * (j & 3)));
        sts[o] = cur;
        ste[o] = (uint32_t)(base + j + (CHUNKED ? 0 : 1));
        ++k;
      }
    }
    if (r0 == 0) {
      // the count tile's prefix: decoupled look-back by wave 0 while the staging above runs
      if (w == 0) {
        uint32_t cpre = 0;
        int32_t lpre = -1;
        if (need_lb) {
          lookback_count(p, ct, lane, cpre, lpre);
          if (lane == 0 && !ct_reset)
            st_relaxed64(p.st_count + ct, pack_count(cpre + ct_tot, max(lpre, ct_last), CT_INC));
        }
        if (lane == 0) {
          s_cpre = cpre;
          s_lpre = lpre;
        }
      }
      __syncthreads();
      const uint32_t cpre = takes_prefix ? s_cpre : 0u;
      const int32_t lpre = takes_prefix ? s_lpre : -1;
      wout = (uint32_t)tv.tok_base + s_wcnt[w] + cpre;
      pending_start = (uint32_t)max(lpre, 0);
      if (start_in == START_PENDING) start_in = pending_start;
    }
    if (!any) break;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t m = min((uint32_t)LEX_STAGE, wtotal - r0);
    for (uint32_t i = lane; i < m; i += WAVE) {
      const uint64_t g = (uint64_t)wout + r0 + i;
      uint32_t sv = sts[i];
      if (sv == START_PENDING) sv = pending_start;
      if (!CHUNKED && p.kinds8) p.kinds8[g] = stk[i];
      else p.kinds[g] = stk[i];
      if (CHUNKED) {
        uint64_t cs, ce;
        chunk_span(p, sv, ste[i], cs, ce);
        p.starts64[g] = cs;
        p.ends64[g] = ce;
      } else {
        p.starts[g] = sv;
        p.ends[g] = ste[i];
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // ---- CHUNKED: the last element of every chunk that ends in this lane
  if (CHUNKED) {
    uint32_t lm = lastm;
    while (lm) {
      const int j = __ffs(lm) - 1;
      lm &= lm - 1;
      const uint32_t le = j == 31 ? 0xFFFFFFFFu : ((2u << j) - 1u);
      const uint32_t sm = startm & le;
      const uint32_t cur = sm ? (uint32_t)(base + 31 - __clz(sm)) : (uint32_t)start_in;
      uint32_t lw;
      switch (j >> 2) {
        case 0: lw = labs[0]; break;
        case 1: lw = labs[1]; break;
        case 2: lw = labs[2]; break;
        case 3: lw = labs[3]; break;
        case 4: lw = labs[4]; break;
        case 5: lw = labs[5]; break;
        case 6: lw = labs[6]; break;
        default: lw = labs[7]; break;
      }
      const uint32_t v = (uint32_t)(base + j);
      uint64_t cs, ce;
      chunk_span(p, cur, v, cs, ce);
      uint64_t *ci = p.chunk_info + 4ull * (v / (p.chunk + 1));
      ci[0] = (uint64_t)wout + lofs + __popc(endm & le);
      ci[1] = (lw >> (8 * (j & 3))) & 0xFFu;
      ci[2] = cs;
      ci[3] = ce;
    }
  }
  // ---- the lane holding the last byte appends the final token and decides acceptance
  if (valid > 0 && base + (uint64_t)valid == n) {
    const uint32_t out = wout + wtotal;
    int32_t cur = start_in;
    if (startm) cur = (int32_t)(base + 31 - __clz(startm));
    uint32_t total = out;
    if (CHUNKED) {
      // the last chunk's last element, then the token lex_int_flag appends with the unshifted
      // start and end n (lexer.fut:161-169)
      const uint32_t v = (uint32_t)(n - 1);
      uint64_t cs, ce;
      chunk_span(p, (uint32_t)cur, v, cs, ce);
      uint64_t *ci = p.chunk_info + 4ull * (p.nchunks - 1);
      ci[0] = out;
      ci[1] = lab_end;
      ci[2] = cs;
      ci[3] = ce;
      if (lab_end != ign) {
        p.kinds[out] = lab_end;
        p.starts64[out] = (uint64_t)(cur - cur / (p.chunk + 1));
        p.ends64[out] = p.n_orig;
        total = out + 1;
      }
    } else if (lab_end != ign) {
      if (p.kinds8) p.kinds8[out] = (uint8_t)lab_end;
      else p.kinds[out] = lab_end;
      p.starts[out] = (uint32_t)cur;
      p.ends[out] = (uint32_t)n;
      total = out + 1;
    }
    if (BATCH) {
      p.doc_res[2 * tv.doc] = acc_end;
      p.doc_res[2 * tv.doc + 1] = total - (uint32_t)tv.tok_base;
    } else {
      p.result[1] = total;
      p.result[0] = acc_end;
    }
  }
}
#undef LEX_WALK

// Chunked mode: in lex_step every element that ends no token is scattered to prev_size - 1
// (lexer.fut:108-111, 131-132), the last token of the chunks before; with several such elements
// the last one (the chunk's last element) is the one kept, and a later chunk that adds no token
// overwrites the same slot again. Chunk j's element wins the slot iff chunk j + 1 moves the slot.
__global__ void lex_chunk_fixup(LexArgs p) {
  const uint32_t K = p.nchunks;
  for (uint32_t j = 1 + blockIdx.x * blockDim.x + threadIdx.x; j < K; j += gridDim.x * blockDim.x) {
    const uint64_t c = p.chunk_info[4ull * (j - 1)];
    if (c == 0) continue;  // slot -1: outside the array, dropped by scatter
    if (j + 1 < K && p.chunk_info[4ull * j] == c) continue;
    p.kinds[c - 1] = (uint32_t)p.chunk_info[4ull * j + 1];
    p.starts64[c - 1] = p.chunk_info[4ull * j + 2];
    p.ends64[c - 1] = p.chunk_info[4ull * j + 3];
  }
}

// Chunked mode: lay the chunks [j*chunk, min(j*chunk + chunk + 1, n)) end to end.
__global__ void lex_chunk_expand(const uint8_t *in, uint32_t chunk, uint64_t nv, uint8_t *out) {
  const uint32_t cp1 = chunk + 1;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nv;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t vv = (uint32_t)v;
    out[v] = in[vv - vv / cp1];
  }
}

// one word set from the device (no memset node, so that a captured graph replays it)
__global__ void set_u32(uint32_t *p, uint32_t v) { *p = v; }

__global__ void clear_u32(uint32_t *p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// ----------------------------------------------------------------------------------------------
// 64-ary min levels and the "previous smaller-or-equal" search (find_previous (<=),
// futhark/parser.fut:92-108, over a 64-ary tree instead of a binary one)

constexpr int MAX_LEVELS = 8;

struct Levels {
  const int32_t *L[MAX_LEVELS];
  uint64_t len[MAX_LEVELS];
  int nlev;
};

// Level 2 of two min hierarchies in one launch: every wavefront reduces 64-entry groups of
// level 1 of either hierarchy (coalesced).
__global__ void __launch_bounds__(256) min_level2(Levels la, Levels lb) {
  const int lane = threadIdx.x & (WAVE - 1);
  const uint64_t na = la.nlev > 2 ? la.len[2] : 0, nb = lb.nlev > 2 ? lb.len[2] : 0;
  const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x / WAVE);
  for (uint64_t c = blockIdx.x * (uint64_t)(blockDim.x / WAVE) + threadIdx.x / WAVE; c < na + nb;
       c += nw) {
    const bool first = c < na;
    const uint64_t k = first ? c : c - na;
    const int32_t *src = first ? la.L[1] : lb.L[1];
    const uint64_t slen = first ? la.len[1] : lb.len[1];
    const uint64_t i = k * WAVE + lane;
    const int32_t m = wave_min_i32(i < slen ? src[i] : INT32_MAX);
    if (lane == 0) const_cast<int32_t *>(first ? la.L[2] : lb.L[2])[k] = m;
  }
}

// Levels 3 .. of both hierarchies (small: level 3 has at most 1/4096 of level 1's entries), one
// block, one thread per 64-entry group.
__global__ void __launch_bounds__(1024) min_levels_top(Levels la, Levels lb) {
  // the two hierarchies side by side, one half of the workgroup each
  const int h = threadIdx.x >= blockDim.x / 2;
  const uint32_t t = threadIdx.x - (h ? blockDim.x / 2 : 0), nt = blockDim.x / 2;
  const Levels &lv = h ? lb : la;
  const int nlev = max(la.nlev, lb.nlev);
  for (int l = 3; l < nlev; ++l) {
    if (l < lv.nlev) {
      const int32_t *src = lv.L[l - 1];
      const uint64_t slen = lv.len[l - 1];
      for (uint64_t k = t; k < lv.len[l]; k += nt) {
        int32_t m = INT32_MAX;
#pragma unroll
        for (int q = 0; q < WAVE / 4; ++q) {
          const uint64_t i = k * WAVE + 4 * q;  // levels are padded past their length
          if (i < slen) {
            const int4 x = *reinterpret_cast<const int4 *>(src + i);
            m = min(m, x.x);
            if (i + 1 < slen) m = min(m, x.y);
            if (i + 2 < slen) m = min(m, x.z);
            if (i + 3 < slen) m = min(m, x.w);
          }
        }
        const_cast<int32_t *>(lv.L[l])[k] = m;
      }
    }
    __syncthreads();  // level l before level l + 1 (one block: workgroup scope suffices)
  }
}

// largest j in [lo, pos) with A[j] <= v, or -1; A is readable in whole 16-byte quads
__device__ __forceinline__ int64_t search_back(const int32_t *A, uint64_t lo, uint64_t pos,
                                               int32_t v) {
  if (pos <= lo) return -1;
  const int4 *A4 = reinterpret_cast<const int4 *>(A);
  int64_t q = (int64_t)((pos - 1) >> 2);
  const int64_t qlo = (int64_t)(lo >> 2);
  const int32_t big = INT32_MAX;
  while (q >= qlo) {
    int4 x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = (q - u >= qlo) ? A4[q - u] : make_int4(big, big, big, big);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t b = (q - u) * 4;
      if (b + 3 < (int64_t)pos && b + 3 >= (int64_t)lo && x[u].w <= v) return b + 3;
      if (b + 2 < (int64_t)pos && b + 2 >= (int64_t)lo && x[u].z <= v) return b + 2;
      if (b + 1 < (int64_t)pos && b + 1 >= (int64_t)lo && x[u].y <= v) return b + 1;
      if (b < (int64_t)pos && b >= (int64_t)lo && x[u].x <= v) return b;
    }
    q -= 4;
  }
  return -1;
}

// per-lane climb through levels 1.. from chunk `chunk0`, then descent (rarely needed)
__device__ int64_t prev_le_far(const Levels &lv, uint64_t chunk0, int32_t v) {
  if (lv.nlev <= 1) return -1;
  int lev = 1;
  uint64_t pos = chunk0;
  int64_t found = -1;
  while (true) {
    const uint64_t lo = (lev == lv.nlev - 1) ? 0 : ((pos >> 6) << 6);
    found = search_back(lv.L[lev], lo, pos, v);
    if (found >= 0) break;
    if (lev == lv.nlev - 1) return -1;
    pos >>= 6;
    ++lev;
  }
  while (lev > 0) {
    const uint64_t lo = (uint64_t)found * 64;
    const uint64_t hi = min(lo + 64, lv.len[lev - 1]);
    found = search_back(lv.L[lev - 1], lo, hi, v);
    --lev;
    if (found < 0) return -1;
  }
  return found;
}

__device__ __forceinline__ int hibit64(uint64_t m) { return 63 - __clzll((long long)m); }

// Binary lifting over a wave-level sparse table of minima: lane l holds x_l; returns for every lane
// the smallest p in [0, pos] such that x_p .. x_{pos-1} are all > v (p = pos when x_{pos-1} <= v).
// All 64 lanes must call this together.
//
// ds_bpermute selects the source lane by address bits 7:2 and reads no memory, so lane indexes
// are kept as byte addresses and a negative one merely reads some lane whose value goes unused.
__device__ __forceinline__ int lift_left(int32_t x, int pos, int32_t v) {
  const int lane = threadIdx.x & (WAVE - 1);
  // M[k] at lane l: min of x_l .. x_{l+2^k-1}. Windows running past lane 63 wrap around and hold
  // wrong minima, but no query below reads a window that ends past x_63.
  int32_t M[6];
  M[0] = x;
#pragma unroll
  for (int k = 1; k < 6; ++k) {
    const int s = 1 << (k - 1);
    M[k] = min(M[k - 1], __builtin_amdgcn_ds_bpermute(((lane + s) & (WAVE - 1)) << 2, M[k - 1]));
  }
  int p4 = pos << 2;
#pragma unroll
  for (int k = 5; k >= 0; --k) {
    const int src4 = p4 - (4 << k);
    // M[k] at lane src covers x_src .. x_{src+step-1} = x_{pos-step} .. x_{pos-1}
    const int32_t m = __builtin_amdgcn_ds_bpermute(src4, M[k]);
    if (src4 >= 0 && m > v) p4 = src4;
  }
  pos = p4 >> 2;
  // the steps add up to 63: from pos = 64 the loop stops at 1 without having looked at x_0
  const int32_t x0 = __builtin_amdgcn_readlane(x, 0);
  if (pos == 1 && x0 > v) pos = 0;
  return pos;
}

// Last j with y_j <= v among 64 values held in reverse (lane l holds y_{63-l}; INT32_MAX for
// missing values), or -1: the prefix minima of the reversed values fall monotonically with the
// lane, so the answer is 63 minus the number of prefix minima above v (a 6-step binary search).
// All 64 lanes must call this together; a lane that asks nothing passes v = INT32_MIN.
__device__ __forceinline__ int last_le_rev(int32_t yrev, int32_t v) {
  const int32_t pm = wave_incl_i32(yrev, INT32_MAX, OpMin());
  const int32_t lowest = __builtin_amdgcn_readlane(pm, WAVE - 1);
  int lo4 = 0;  // 4 * the number of prefix minima known to lie above v
#pragma unroll
  for (int step = WAVE / 2; step >= 1; step >>= 1)
    if (__builtin_amdgcn_ds_bpermute(lo4 + 4 * (step - 1), pm) > v) lo4 += 4 * step;
  return lowest <= v ? WAVE - 1 - (lo4 >> 2) : -1;
}

// The reversed inputs of the search for chunk `chunk`: the previous chunk's elements (lane l:
// element 64*chunk - 1 - l) and the minima of the 64 chunks before that (lane l: chunk - 2 - l).
struct SearchRev {
  int32_t prev, win;
};

__device__ __forceinline__ SearchRev search_rev(const Levels &lv, uint64_t chunk, int lane) {
  SearchRev r{INT32_MAX, INT32_MAX};
  if (chunk >= 1) r.prev = lv.L[0][chunk * WAVE - 1 - lane];
  if (lv.nlev >= 2 && (int64_t)chunk - 2 - lane >= 0) r.win = lv.L[1][chunk - 2 - lane];
  return r;
}

// Wave-cooperative previous smaller-or-equal for the 64 consecutive elements of chunk `chunk`
// (lane l holds element chunk*64+l with value v; lanes past the end hold INT32_MAX): binary
// lifting inside the chunk; then the previous chunk (every chunk before the last one is full);
// then a search over the minima of the 64 chunks before that picks the chunk holding the answer,
// and each distinct such chunk is loaded once (coalesced) and searched; a per-lane climb of the
// min hierarchy handles the rest. All 64 lanes must call this together.
// p0 = lift_left(v, lane, v), computed by the caller (two chunks' lifts overlap there).
__device__ int64_t prev_le_wave(const Levels &lv, uint64_t chunk, int lane, int32_t v,
                                const SearchRev &rv, bool need, int p0) {
  const int64_t cbase = (int64_t)(chunk * WAVE);
  int64_t res = -1;
  bool done = !need;
  if (!done && p0 > 0) {
    res = cbase + p0 - 1;
    done = true;
  }
  if (__ballot(!done) == 0ull || chunk == 0) return res;
  const int j1 = last_le_rev(rv.prev, done ? INT32_MIN : v);
  if (!done && j1 >= 0) {
    res = cbase - WAVE + j1;
    done = true;
  }
  if (__ballot(!done) == 0ull || chunk == 1) return res;
  // window: chunks chunk-65 .. chunk-2 (window index j is chunk chunk-65+j)
  const int jw = last_le_rev(rv.win, done ? INT32_MIN : v);
  int64_t cb = (!done && jw >= 0) ? (int64_t)chunk - (WAVE + 1) + jw : -1;
  uint64_t pend = __ballot(cb >= 0);
  while (pend) {
    const int t = __ffsll((long long)pend) - 1;
    const int64_t c = (int64_t)(uint32_t)__builtin_amdgcn_readlane((int)(cb & 0xFFFFFFFF), t) |
                      ((int64_t)__builtin_amdgcn_readlane((int)(cb >> 32), t) << 32);
    const int32_t yrev = lv.L[0][c * WAVE + (WAVE - 1) - lane];
    const bool mine = cb == c;
    const int j2 = last_le_rev(yrev, mine ? v : INT32_MIN);
    if (mine) {
      res = c * WAVE + j2;
      done = true;
      cb = -1;
    }
    pend = __ballot(cb >= 0);
  }
  if (!done)
    res = prev_le_far(lv, chunk > (uint64_t)(WAVE + 1) ? chunk - (WAVE + 1) : 0, v);
  return res;
}

// ----------------------------------------------------------------------------------------------
// parser kernels
//
// The LLP parser of futhark/parser.fut:235-336 over the pair table of the grammar. The
// reference builds the bracket stream and the production stream by segmented copies of the pair
// entries (segmented_copy, :216-233), then scans them element by element: bracket depths
// (depths, :153-166), s = exscan(arity - 1) for the parents (:286-296) and the rank of every
// terminal production (terminal_offsets, :305-314). Here every pair entry carries its totals
// (brackets, productions, depth change, lowest depth prefix, sum of arity - 1, terminal
// productions) and every bracket / production of an entry its value relative to the entry start,
// so one scan over the T + 2 positions gives every element its depth, s and rank:
//   parse_keys2   : key -> entry of every position, per-tile totals          (reads the tokens)
//   parse_pscan   : exclusive scan of the tile totals, depth validity        (one block)
//   parse_expand2 : per tile, the positions' prefixes, then every bracket / production element
//                   written directly (depth + symbol; s, node and span), with 64-element minima
//   bracket_match : brackets_matches (:174-184) by previous smaller-or-equal depth
//   parse_parents : parents (:286-296) by previous smaller-or-equal s

constexpr int PK_THREADS = 256;
constexpr int PK_IPT = 4;
constexpr int PK_TILE = PK_THREADS * PK_IPT;  // positions per tile
constexpr int PK_MAXW = 16;                   // q + k at most
constexpr int PK_IDX8_MAX = 52 * 1024;        // dense u8 pair index kept in LDS up to this size
constexpr int PS_THREADS = 1024;              // parse_pscan block
constexpr uint16_t PK_NO_ENTRY = 0xFFFF;      // position whose key has no pair entry

// One LLP pair entry (16 bytes); mind = lowest inclusive depth prefix relative to the entry
// start, INT16_MAX without brackets.
struct EntInfo {
  uint16_t nb, np;
  int16_t dd, mind;
  int16_t ds;
  uint16_t nt;
  uint16_t sst, pst;  // first bracket in stackw / first production in prodw
};

// Totals of a range of positions. All fields add, except mind, which is the lowest depth prefix
// inside the range relative to its start (INT32_MAX when the range has no bracket).
struct PTup {
  int32_t nb, np, dd, mind, ds, nt;
};

__device__ __forceinline__ PTup ptup_id() { return PTup{0, 0, 0, INT32_MAX, 0, 0}; }

__device__ __forceinline__ PTup ptup_op(const PTup &a, const PTup &b) {
  PTup r;
  r.nb = a.nb + b.nb;
  r.np = a.np + b.np;
  r.dd = a.dd + b.dd;
  r.mind = b.mind == INT32_MAX ? a.mind : min(a.mind, a.dd + b.mind);
  r.ds = a.ds + b.ds;
  r.nt = a.nt + b.nt;
  return r;
}

__device__ __forceinline__ PTup ptup_of(const EntInfo &e) {
  return PTup{e.nb, e.np, e.dd, e.nb ? (int32_t)e.mind : INT32_MAX, e.ds, e.nt};
}

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ PTup ptup_dpp(const PTup &v) {
  return PTup{dpp_i32<CTRL, ROW_MASK>(0, v.nb),       dpp_i32<CTRL, ROW_MASK>(0, v.np),
              dpp_i32<CTRL, ROW_MASK>(0, v.dd),       dpp_i32<CTRL, ROW_MASK>(INT32_MAX, v.mind),
              dpp_i32<CTRL, ROW_MASK>(0, v.ds),       dpp_i32<CTRL, ROW_MASK>(0, v.nt)};
}

// inclusive scan across the wavefront, in lane order (DPP rows, then row broadcasts)
__device__ __forceinline__ PTup ptup_wave_incl(PTup v, int lane) {
  (void)lane;
  v = ptup_op(ptup_dpp<DPP_ROW_SHR1>(v), v);
  v = ptup_op(ptup_dpp<DPP_ROW_SHR2>(v), v);
  v = ptup_op(ptup_dpp<DPP_ROW_SHR4>(v), v);
  v = ptup_op(ptup_dpp<DPP_ROW_SHR8>(v), v);
  v = ptup_op(ptup_dpp<DPP_ROW_BCAST15, 0xA>(v), v);
  v = ptup_op(ptup_dpp<DPP_ROW_BCAST31, 0xC>(v), v);
  return v;
}

// Block-wide exclusive scan in thread order; returns the exclusive prefix of this thread and the
// block total. s_w: LDS of at least blockDim.x / WAVE tuples.
__device__ __forceinline__ PTup ptup_block_excl(PTup v, PTup *s_w, PTup &total) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  const int nw = blockDim.x / WAVE;
  const PTup incl = ptup_wave_incl(v, lane);
  if (lane == WAVE - 1) s_w[w] = incl;
  __syncthreads();
  PTup pre = ptup_id();
  for (int i = 0; i < w; ++i) pre = ptup_op(pre, s_w[i]);
  total = pre;
  for (int i = w; i < nw; ++i) total = ptup_op(total, s_w[i]);
  const PTup ex = ptup_dpp<DPP_WAVE_SHR1>(incl);  // lane 0: identity
  __syncthreads();
  return ptup_op(pre, ex);
}

// Adjacent cancellation. A position's trailing left brackets and the next position's leading right
// brackets follow each other directly in the bracket stream (brackets_matches, parser.fut:174-184),
// so the last left of the one and the first right of the other match (the previous smaller-or-equal
// depth of that right bracket is that left bracket), then the next pair inward, and so on:
// min(trailing lefts, leading rights) pairs. Their symbols are checked when the stream is
// expanded and they are never written; the remaining brackets keep their depth values and their
// matches, and the depth check still sees every bracket (the entries' totals are not changed).
// Pairs cancel inside one wavefront's 256 positions (the neighbour comes through a DPP wave shift).
// tlr[i]: trailing lefts | leading rights << 8 (| brackets between << 16) of the entry of this
// lane's position i (0 without one); cr[i]: pairs position i cancels with position i + 1; cl0: pairs the first position cancels
// with the previous lane's last. All lanes of the wavefront take part.
__device__ __forceinline__ void pk_cancel(const uint32_t (&tlr)[PK_IPT], int32_t (&cr)[PK_IPT],
                                          int32_t &cl0) {
#pragma unroll
  for (int i = 0; i + 1 < PK_IPT; ++i) cr[i] = (int32_t)min(tlr[i] & 0xFFu, (tlr[i + 1] >> 8) & 0xFFu);
  const uint32_t nx = (uint32_t)dpp_i32<DPP_WAVE_SHL1>(0, (int32_t)tlr[0]);  // lane 63: none
  cr[PK_IPT - 1] = (int32_t)min(tlr[PK_IPT - 1] & 0xFFu, (nx >> 8) & 0xFFu);
  cl0 = dpp_i32<DPP_WAVE_SHR1>(0, cr[PK_IPT - 1]);  // lane 0: none
}

// brackets the cancellations above take away from this lane's positions
__device__ __forceinline__ int32_t pk_cancelled(const int32_t (&cr)[PK_IPT], int32_t cl0) {
  int32_t s = cl0 + cr[PK_IPT - 1];
#pragma unroll
  for (int i = 0; i + 1 < PK_IPT; ++i) s += 2 * cr[i];
  return s;
}

// Second level. After the adjacent pairs are gone, a position's remaining trailing lefts and the
// remaining leading rights of the next position that still has brackets (the positions between
// have none left) follow each other directly as well: min(lefts, rights) more pairs. Inside the
// wavefront: the next such position in another lane comes through a ballot and a permute.
// Per position: n1 brackets left after the adjacent cancellation, r1 leading rights and l1
// trailing lefts among them, f1 the stackw index of its first remaining bracket.
// Out: cr2[i] pairs position i cancels with its next position that has brackets left, nf[i] that
// position's f1, cl2[i] the pairs position i cancels with its previous one. All lanes of the
// wavefront take part. (parse_keys2 counts the same pairs with a leaner copy of the first half.)
__device__ __forceinline__ void pk_cancel2(const int32_t (&n1)[PK_IPT],
                                           const int32_t (&r1)[PK_IPT],
                                           const int32_t (&l1)[PK_IPT],
                                           const uint32_t (&f1)[PK_IPT], int lane,
                                           int32_t (&cr2)[PK_IPT], uint32_t (&nf)[PK_IPT],
                                           int32_t (&cl2)[PK_IPT]) {
  // the lane's first position with brackets left
  int32_t fr = 0;
  uint32_t ff = 0;
  bool any = false;
#pragma unroll
  for (int i = PK_IPT - 1; i >= 0; --i) {
    if (n1[i] > 0) {
      fr = r1[i];
      ff = f1[i];
      any = true;
    }
  }
  const uint64_t mask = __ballot(any);
  const uint64_t above = lane == WAVE - 1 ? 0ull : (mask >> (lane + 1));
  const int nl = above ? lane + 1 + (int)__builtin_ctzll(above) : lane;
  const int32_t xr_all = __shfl(fr, nl, WAVE);
  const uint32_t xf_all = (uint32_t)__shfl((int)ff, nl, WAVE);
  const int32_t xr = above ? xr_all : 0;
#pragma unroll
  for (int i = 0; i < PK_IPT; ++i) {
    int32_t nr = xr;  // leading rights left at the next position with brackets
    uint32_t nfi = xf_all;
#pragma unroll
    for (int k = PK_IPT - 1; k > i; --k) {
      if (n1[k] > 0) {
        nr = r1[k];
        nfi = f1[k];
      }
    }
    cr2[i] = min(l1[i], nr);
    nf[i] = nfi;
  }
  // the lane's last position with brackets left: the pairs it cancels forward
  int32_t lc = 0;
#pragma unroll
  for (int i = 0; i < PK_IPT; ++i)
    if (n1[i] > 0) lc = cr2[i];
  const uint64_t below = mask & ((1ull << lane) - 1ull);
  const int pl = below ? 63 - (int)__builtin_clzll(below) : lane;
  const int32_t xc_all = __shfl(lc, pl, WAVE);
  const int32_t xc = below ? xc_all : 0;
#pragma unroll
  for (int k = 0; k < PK_IPT; ++k) {
    int32_t pc = xc;
#pragma unroll
    for (int i = 0; i < k; ++i)
      if (n1[i] > 0) pc = cr2[i];
    cl2[k] = n1[k] > 0 ? pc : 0;
  }
}

struct PArgs {
  const uint32_t *kinds;
  const uint8_t *kinds8;  // when set: the token kinds as bytes (straight from the lexer)
  uint64_t T;  // tokens
  uint64_t P;  // positions: T + 2 (start and end terminal added, to_keys parser.fut:235-240)
  // single document straight after the lexer: the token count on the device (T, P and the tile
  // count above are then upper bounds that size the grids)
  const uint32_t *T_dev;
  uint32_t q, k, nT, start_t, end_t, empty_t;
  const uint8_t *index8;    // dense pair index, entry or 0xFF (when E < 255)
  const uint32_t *index32;  // dense pair index, entry or 0xFFFFFFFF
  uint32_t index_len;
  // hashed pair table (index32 NULL): dense key index per slot, entry id per slot
  const uint64_t *hkeys;
  const uint32_t *hvals;
  uint32_t hbits, hprobe;   // log2 of the slots; the longest probe sequence of any stored key
  const EntInfo *einfo;
  uint32_t E;
  const uint32_t *stackw;  // bracket symbol (| 0x8000 left) | depth value relative to the entry << 16
  const uint2 *prodw;      // x: production | (terminal + 1) << 16; y: s relative | rank relative << 16
  const uint32_t *ecan;    // per entry: trailing lefts | leading rights << 8 | brackets between << 16
  uint32_t nstack, nprod;  // lengths of stackw / prodw
  uint16_t *ent;           // [ntiles * PK_TILE] pair entry of every position
  PTup *tiles;             // [ntiles] tile totals
  PTup *pre;               // [ntiles] exclusive prefixes (parse_pscan)
  PTup *total;             // [1]
  uint32_t *flags;         // [0] reject
  uint64_t ntiles;
  int subst;               // single document: terminal productions pair up with the tokens
  // batched mode: every document's positions start a new tile; element offsets run through the
  // documents, depths, s and ranks restart in every document
  struct PDoc *pdocs;        // [ndocs]
  const uint32_t *ptile_doc; // [tiles] document of every position tile
  const uint64_t *ntiles_dev;
  uint32_t ndocs;
  uint32_t *doc_flags;       // [ndocs] bit 0: missing pair, bit 1: bracket mismatch
  uint32_t *chunk_doc;       // [N / 64] document of the first production of every 64-chunk
  uint32_t *chunk_base;      // [N / 64] first node of that document; bit 31: a later document
                             // starts inside the chunk
  uint64_t expand_tile0;     // the striding parse_expand2's first tile
  // batched mode: capacities of the bracket and production scratch, sized before the launch; a
  // batch whose totals exceed them (or 2^31) writes no elements (the host then sizes them again)
  uint64_t bcap, ncap;
};

// Batched mode: the brackets / productions the element kernels process: the grand totals, or
// none when they exceed the scratch capacities (then nothing is written).
__device__ __forceinline__ bool batch_elems_fit(const PArgs &a) {
  const PTup t = *a.total;
  return a.flags[1] == 0u && t.nb >= 0 && t.np >= 0 && (uint64_t)t.nb <= a.bcap &&
         (uint64_t)t.np <= a.ncap;
}


struct PDoc {
  uint64_t tok_base;    // first token slot of the document (batched lexer layout)
  uint64_t br_off;      // first bracket / production element (global)
  uint64_t node_off;
  uint64_t B, N;        // brackets, productions
  uint32_t T;           // tokens
  uint32_t lex_ok;
  uint32_t ptile_base;  // first position tile
  uint32_t ptiles;
  uint32_t status;      // 1 while accepted so far
  uint32_t subst;
};

// The positions a tile covers: the document's tokens, its position count, the tile's first
// position inside the document.
struct PosView {
  const uint32_t *kinds;
  const uint8_t *kinds8;
  uint64_t P;
  int64_t p0;
  uint64_t tok_base;
  uint32_t doc;
  bool reset;
  int subst;
};

template <bool BATCH>
__device__ __forceinline__ PosView pos_view(const PArgs &a, uint64_t tile) {
  PosView v;
  if (!BATCH) {
    v.kinds = a.kinds;
    v.kinds8 = a.kinds8;
    v.P = a.T_dev ? std::min<uint64_t>(*a.T_dev, a.T) + 2 : a.P;
    v.p0 = (int64_t)(tile * PK_TILE);
    v.tok_base = 0;
    v.doc = 0;
    v.reset = tile == 0;
    v.subst = a.subst;
  } else {
    const uint32_t d = a.ptile_doc[tile] & 0x7FFFFFFFu;
    const PDoc &doc = a.pdocs[d];
    const uint32_t lt = (uint32_t)tile - doc.ptile_base;
    v.kinds = a.kinds + doc.tok_base;
    v.kinds8 = a.kinds8 ? a.kinds8 + doc.tok_base : nullptr;
    v.P = (uint64_t)doc.T + 2;
    v.p0 = (int64_t)lt * PK_TILE;
    v.tok_base = doc.tok_base;
    v.doc = d;
    v.reset = lt == 0;
    v.subst = (int)doc.subst;
  }
  return v;
}

// position tiles to process: the device count in batched mode or after the lexer, else a.ntiles
template <bool BATCH>
__device__ __forceinline__ uint64_t pk_ntiles(const PArgs &a) {
  if (BATCH) return *a.ntiles_dev;
  if (a.T_dev) return (std::min<uint64_t>(*a.T_dev, a.T) + 2 + PK_TILE - 1) / PK_TILE;
  return a.ntiles;
}

// symbol of position x of [start] ++ kinds ++ [end], the empty terminal outside (get_key,
// parser.fut:117-132); a kind the table does not know rejects
__device__ __forceinline__ uint32_t pk_symbol(const PArgs &a, const PosView &pv, int64_t x) {
  if (x < 0 || x >= (int64_t)pv.P) return a.empty_t;
  if (x == 0) return a.start_t;
  if (x == (int64_t)pv.P - 1) return a.end_t;
  return pv.kinds8 ? (uint32_t)pv.kinds8[x - 1] : pv.kinds[x - 1];
}

// keys and pair lookups (keys / lookup / to_keys, parser.fut:110-151, 235-240); per-tile totals.
// Persistent, PKK_TILES tiles per block and iteration, one wavefront group of PK_THREADS threads per
// tile; the u8 pair index is loaded into LDS once per block.
constexpr int PKK_THREADS = 1024;
constexpr int PKK_TILES = PKK_THREADS / PK_THREADS;

// Entry id of the key window with dense index di in the hashed pair table, or 0xFFFFFFFF: linear
// probing from the multiplicative-hash home slot, at most hprobe slots (every stored key lies
// within them, counted when the tables were built), stopping at a free slot.
__device__ __forceinline__ uint32_t pk_hash_lookup(const PArgs &a, uint64_t di) {
  const uint64_t mask = (1ull << a.hbits) - 1;
  uint64_t h = (di * 0x9E3779B97F4A7C15ull) >> (64 - a.hbits);
  for (uint32_t i = 0; i < a.hprobe; ++i) {
    const uint64_t key = a.hkeys[h];
    if (key == di) return a.hvals[h];
    if (key == ~0ull) break;
    h = (h + 1) & mask;
  }
  return 0xFFFFFFFFu;
}

// PosView of tile `tile` (an empty view past the tile count) and the symbols of the key windows of
// this thread's PK_IPT positions in it.
template <bool BATCH, int KW>
__device__ __forceinline__ void keys_window(const PArgs &a, uint64_t tile, uint64_t ntiles,
                                            uint32_t W, PosView &pv,
                                            uint32_t (&sym)[PK_IPT + KW - 1], bool &raw,
                                            uint32_t &off) {
  if (tile < ntiles) {
    pv = pos_view<BATCH>(a, tile);
  } else {
    pv = PosView{a.kinds, a.kinds8, 0, 0, 0, 0, false, 0};  // every field set: the view stays in registers
  }
  const int64_t lp0 = pv.p0 + (int64_t)(threadIdx.x % PK_THREADS) * PK_IPT;  // in the document
  // A tile away from both ends of byte kinds (uniform over its threads): a thread's 7 symbols
  // are 7 consecutive bytes, read as the three aligned words around them (raw; keys_symbols
  // takes the bytes out). 16 positions of margin keep the words inside the kinds.
  raw = KW == 4 && tile < ntiles && pv.kinds8 != nullptr && pv.p0 >= (int64_t)a.q + 4 &&
        pv.p0 + PK_TILE + 16 <= (int64_t)pv.P;
  if (raw) {
    const uintptr_t addr = (uintptr_t)(pv.kinds8 + (lp0 - (int64_t)a.q - 1));
    off = (uint32_t)(addr & 3);
    const uint32_t *wp = reinterpret_cast<const uint32_t *>(addr - off);
    sym[0] = wp[0];
    sym[1] = wp[1];
    sym[2] = wp[2];
#pragma unroll
    for (int j = 3; j < PK_IPT + KW - 1; ++j) sym[j] = 0u;
    return;
  }
  off = 0;
#pragma unroll
  for (int j = 0; j < PK_IPT + KW - 1; ++j)
    sym[j] = j < PK_IPT + (int)W - 1 ? pk_symbol(a, pv, lp0 - (int64_t)a.q + j) : 0u;
}

// The symbols of a window keys_window read raw: bytes off .. off + 6 of its three words.
template <int KW>
__device__ __forceinline__ void keys_symbols(const uint32_t (&raw)[PK_IPT + KW - 1], uint32_t off,
                                             uint32_t (&sym)[PK_IPT + KW - 1]) {
  const uint32_t lo = __builtin_amdgcn_alignbyte(raw[1], raw[0], off);
  const uint32_t hi = __builtin_amdgcn_alignbyte(raw[2], raw[1], off);
#pragma unroll
  for (int j = 0; j < PK_IPT + KW - 1; ++j)
    sym[j] = j < 4 ? (lo >> (8 * j)) & 0xFFu : (j < 8 ? (hi >> (8 * (j - 4))) & 0xFFu : 0u);
}

// KW: the largest key window q + k the instance handles (4 or PK_MAXW); the next iteration's
// symbols are loaded while the current ones are processed.
template <bool IDX8, bool BATCH, int KW>
// Key windows of at most 4 terminals: two 1024-thread blocks per CU, so 8 waves per SIMD (at most
// 64 VGPRs).
__global__ void __launch_bounds__(PKK_THREADS) __attribute__((amdgpu_waves_per_eu(KW == 4 ? 8 : 1)))
parse_keys2(PArgs a) {
  extern __shared__ __align__(16) uint8_t pk_smem[];
  __shared__ PTup s_w[PKK_THREADS / WAVE];
  __shared__ EntInfo s_ei[256];
  __shared__ uint32_t s_can[256];
  if (IDX8) {
    for (uint32_t i = threadIdx.x; i < a.index_len; i += blockDim.x) pk_smem[i] = a.index8[i];
    for (uint32_t i = threadIdx.x; i < a.E; i += blockDim.x) {
      s_ei[i] = a.einfo[i];
      s_can[i] = a.ecan[i];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  const uint32_t W = a.q + a.k;
  const uint64_t ntiles = pk_ntiles<BATCH>(a);
  const uint64_t nsuper = (ntiles + PKK_TILES - 1) / PKK_TILES;
  bool miss = false;
  PosView npv;
  uint32_t nsym[PK_IPT + KW - 1];
  bool nraw = false;
  uint32_t noff = 0;
  if ((uint64_t)blockIdx.x < nsuper)
    keys_window<BATCH, KW>(a, blockIdx.x * (uint64_t)PKK_TILES + threadIdx.x / PK_THREADS, ntiles,
                           W, npv, nsym, nraw, noff);
  for (uint64_t sup = blockIdx.x; sup < nsuper; sup += gridDim.x) {
    const uint64_t my_tile = sup * PKK_TILES + threadIdx.x / PK_THREADS;
    const PosView pv = npv;
    uint32_t sym[PK_IPT + KW - 1];
    if (nraw) {
      keys_symbols<KW>(nsym, noff, sym);
    } else {
#pragma unroll
      for (int j = 0; j < PK_IPT + KW - 1; ++j) sym[j] = nsym[j];
    }
    if (sup + gridDim.x < nsuper)
      keys_window<BATCH, KW>(a, (sup + gridDim.x) * PKK_TILES + threadIdx.x / PK_THREADS, ntiles,
                             W, npv, nsym, nraw, noff);
    const int64_t lp0 = pv.p0 + (int64_t)(threadIdx.x % PK_THREADS) * PK_IPT;  // in the document
    PTup acc = ptup_id();
    uint32_t idp[PK_IPT / 2] = {0u, 0u};  // entry of every position, two u16 to a word
    uint32_t tlr[PK_IPT];
    bool wmiss = false;
#pragma unroll
    for (int i = 0; i < PK_IPT; ++i) {
      const int64_t p = lp0 + i;
      tlr[i] = 0;
      if (p >= (int64_t)pv.P) {  // past the document: expands to nothing (parse_expand2)
        idp[i >> 1] |= (uint32_t)PK_NO_ENTRY << (16 * (i & 1));
        continue;
      }
      uint64_t di = 0;
      bool unknown = false;
#pragma unroll
      for (int j = 0; j < KW; ++j) {
        if (j < (int)W) {
          const uint32_t t = sym[i + j];
          unknown |= t >= a.nT;
          di = di * a.nT + (t < a.nT ? t : 0u);
        }
      }
      uint32_t id;
      if (IDX8) {
        id = unknown ? 0xFFu : pk_smem[di];
        if (id == 0xFFu) id = 0xFFFFFFFFu;
      } else {
        id = unknown ? 0xFFFFFFFFu : (a.index32 ? a.index32[di] : pk_hash_lookup(a, di));
      }
      if (id == 0xFFFFFFFFu) {
        wmiss = true;
        idp[i >> 1] |= (uint32_t)PK_NO_ENTRY << (16 * (i & 1));  // expands to nothing: rejected anyway
        continue;
      }
      idp[i >> 1] |= id << (16 * (i & 1));
      tlr[i] = IDX8 ? s_can[id] : a.ecan[id];
      acc = ptup_op(acc, ptup_of(IDX8 ? s_ei[id] : a.einfo[id]));
    }
    {
      int32_t cr[PK_IPT], cl0;
      pk_cancel(tlr, cr, cl0);
      acc.nb -= pk_cancelled(cr, cl0);
      // second level (as pk_cancel2; only the pair count is needed here): per position the leading
      // rights and trailing lefts left after the first level, and whether any bracket is left,
      // packed into one word so that the tile's registers stay within 8 waves per SIMD
      uint32_t lv[PK_IPT];
#pragma unroll
      for (int i = 0; i < PK_IPT; ++i) {
        const int32_t c_l = i > 0 ? cr[i - 1] : cl0;
        const uint32_t r1 = ((tlr[i] >> 8) & 0xFFu) - (uint32_t)c_l;
        const uint32_t l1 = (tlr[i] & 0xFFu) - (uint32_t)cr[i];
        lv[i] = r1 | (l1 << 8) | ((r1 | l1 | (tlr[i] >> 16)) != 0u ? 0x10000u : 0u);
      }
      uint32_t fr = 0;
#pragma unroll
      for (int i = PK_IPT - 1; i >= 0; --i)
        if (lv[i] >> 16) fr = lv[i] & 0xFFu;
      const uint64_t mask = __ballot((lv[0] | lv[1] | lv[2] | lv[3]) >> 16);
      const uint64_t above = lane == WAVE - 1 ? 0ull : (mask >> (lane + 1));
      const int nl = above ? lane + 1 + (int)__builtin_ctzll(above) : lane;
      const uint32_t xr_all = (uint32_t)__shfl((int)fr, nl, WAVE);
      int32_t c2 = 0;
#pragma unroll
      for (int i = 0; i < PK_IPT; ++i) {
        uint32_t nr = above ? xr_all : 0u;
#pragma unroll
        for (int k = PK_IPT - 1; k > i; --k)
          if (lv[k] >> 16) nr = lv[k] & 0xFFu;
        c2 += (int32_t)min((lv[i] >> 8) & 0xFFu, nr);
      }
      acc.nb -= 2 * c2;
    }
    if (BATCH) {
      if (__ballot(wmiss) != 0ull && lane == 0) atomicOr(a.doc_flags + pv.doc, 1u);
    } else {
      miss |= wmiss;
    }
    // slots are tile-major: tile * PK_TILE + position in the tile; every slot of a tile is written,
    // those past the document's positions with PK_NO_ENTRY, so that parse_expand2 tells them
    // without the document's bounds
    const uint64_t pbase = sup * PKK_TILES * PK_TILE + (uint64_t)threadIdx.x * PK_IPT;
    if (my_tile < ntiles) *reinterpret_cast<uint2 *>(a.ent + pbase) = make_uint2(idp[0], idp[1]);
    const PTup incl = ptup_wave_incl(acc, lane);
    if (lane == WAVE - 1) s_w[w] = incl;
    __syncthreads();
    if (threadIdx.x < PKK_TILES) {
      const uint64_t tile = sup * PKK_TILES + threadIdx.x;
      if (tile < ntiles) {
        const int wpt = PK_THREADS / WAVE;
        PTup t = s_w[threadIdx.x * wpt];
        for (int i = 1; i < wpt; ++i) t = ptup_op(t, s_w[threadIdx.x * wpt + i]);
        a.tiles[tile] = t;
      }
    }
    __syncthreads();
  }
  if (!BATCH && __ballot(miss) != 0ull && lane == 0) atomicOr(a.flags, 1u);
}

// Segmented tile totals: element counts (nb, np) run through all tiles, the other fields restart
// at a tile that starts a document (batched mode).
struct SegTup {
  PTup t;
  int f;
};

__device__ __forceinline__ SegTup seg_op(const SegTup &a, const SegTup &b) {
  SegTup r;
  if (b.f) {
    r.t = b.t;
    r.t.nb += a.t.nb;
    r.t.np += a.t.np;
    r.f = 1;
  } else {
    r.t = ptup_op(a.t, b.t);
    r.f = a.f;
  }
  return r;
}

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ SegTup seg_dpp(const SegTup &v) {
  return SegTup{ptup_dpp<CTRL, ROW_MASK>(v.t), dpp_i32<CTRL, ROW_MASK>(0, v.f)};
}

// Tile scan into `pre` (exclusive) and the grand total, in three kernels: block totals of 1024
// tiles, one block over the block totals, then every block again with its prefix. In a tile that
// starts a document the prefix keeps only the element counts. The grand total decides the depth
// check of depths (parser.fut:153-166) of a single document: no prefix below zero, and zero at
// the end.
constexpr int PSB = 1024;  // tiles per scan block

template <bool BATCH>
__device__ __forceinline__ int tile_reset(const PArgs &a, uint64_t t) {
  if (!BATCH) return 0;
  return (a.ptile_doc[t] >> 31) != 0 ? 1 : 0;
}

// block-wide segmented scan in thread order: exclusive prefix of this thread, block total
__device__ __forceinline__ SegTup seg_block_excl(SegTup v, SegTup *s_w, SegTup &total) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  const int nw = blockDim.x / WAVE;
  SegTup incl = v;
  incl = seg_op(seg_dpp<DPP_ROW_SHR1>(incl), incl);
  incl = seg_op(seg_dpp<DPP_ROW_SHR2>(incl), incl);
  incl = seg_op(seg_dpp<DPP_ROW_SHR4>(incl), incl);
  incl = seg_op(seg_dpp<DPP_ROW_SHR8>(incl), incl);
  incl = seg_op(seg_dpp<DPP_ROW_BCAST15, 0xA>(incl), incl);
  incl = seg_op(seg_dpp<DPP_ROW_BCAST31, 0xC>(incl), incl);
  if (lane == WAVE - 1) s_w[w] = incl;
  __syncthreads();
  SegTup pre{ptup_id(), 0};
  for (int i = 0; i < w; ++i) pre = seg_op(pre, s_w[i]);
  total = pre;
  for (int i = w; i < nw; ++i) total = seg_op(total, s_w[i]);
  const SegTup ex = seg_dpp<DPP_WAVE_SHR1>(incl);  // lane 0: identity
  __syncthreads();
  return seg_op(pre, ex);
}

template <bool BATCH>
__global__ void __launch_bounds__(PSB) pscan_reduce(PArgs a, SegTup *btot) {
  __shared__ SegTup s_w[PSB / WAVE];
  const uint64_t ntiles = pk_ntiles<BATCH>(a);
  const uint64_t t = (uint64_t)blockIdx.x * PSB + threadIdx.x;
  if ((uint64_t)blockIdx.x * PSB >= ntiles) return;  // whole block beyond the tiles
  const SegTup v = t < ntiles ? SegTup{a.tiles[t], tile_reset<BATCH>(a, t)} : SegTup{ptup_id(), 0};
  SegTup total;
  (void)seg_block_excl(v, s_w, total);
  if (threadIdx.x == 0) btot[blockIdx.x] = total;
}

__global__ void __launch_bounds__(PSB) pscan_top(PArgs a, SegTup *btot, uint64_t nblocks_max) {
  __shared__ SegTup s_w[PSB / WAVE];
  __shared__ SegTup s_carry;
  __shared__ unsigned long long s_elems[2];  // brackets, productions over all blocks (64-bit)
  const uint64_t ntiles = a.ntiles_dev ? *a.ntiles_dev : pk_ntiles<false>(a);
  const uint64_t nb = std::min<uint64_t>(nblocks_max, (ntiles + PSB - 1) / PSB);
  if (threadIdx.x == 0) {
    s_carry = SegTup{ptup_id(), 0};
    s_elems[0] = s_elems[1] = 0ull;
  }
  __syncthreads();
  unsigned long long mine_b = 0ull, mine_p = 0ull;
  for (uint64_t off = 0; off < nb; off += PSB) {
    const uint64_t i = off + threadIdx.x;
    const SegTup v = i < nb ? btot[i] : SegTup{ptup_id(), 0};
    mine_b += (uint32_t)v.t.nb;
    mine_p += (uint32_t)v.t.np;
    SegTup total;
    const SegTup ex = seg_block_excl(v, s_w, total);
    const SegTup carry = s_carry;
    if (i < nb) btot[i] = seg_op(carry, ex);
    __syncthreads();
    if (threadIdx.x == 0) s_carry = seg_op(carry, total);
    __syncthreads();
  }
  // element offsets are 32-bit: flag a grand total of brackets or productions of 2^31 or more
  // (the int32 sums above would have wrapped). Every block total is below 2^31 (at most
  // PSB * PK_TILE positions of fewer than 2048 elements each, checked on the host). One LDS
  // atomic per wavefront (one per thread serialised on the two words: 14 -> 6 us).
#pragma unroll
  for (int d = WAVE / 2; d >= 1; d >>= 1) {
    mine_b += __shfl_xor(mine_b, d, WAVE);
    mine_p += __shfl_xor(mine_p, d, WAVE);
  }
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    atomicAdd(&s_elems[0], mine_b);
    atomicAdd(&s_elems[1], mine_p);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    *a.total = s_carry.t;
    if (s_elems[0] >= (1ull << 31) || s_elems[1] >= (1ull << 31)) a.flags[1] = 1u;
  }
}

template <bool BATCH>
__global__ void __launch_bounds__(PSB) pscan_down(PArgs a, const SegTup *bpre) {
  __shared__ SegTup s_w[PSB / WAVE];
  const uint64_t ntiles = pk_ntiles<BATCH>(a);
  const uint64_t t = (uint64_t)blockIdx.x * PSB + threadIdx.x;
  if ((uint64_t)blockIdx.x * PSB >= ntiles) return;
  const SegTup v = t < ntiles ? SegTup{a.tiles[t], tile_reset<BATCH>(a, t)} : SegTup{ptup_id(), 0};
  SegTup total;
  const SegTup ex = seg_op(bpre[blockIdx.x], seg_block_excl(v, s_w, total));
  if (t < ntiles) {
    PTup out = ex.t;
    if (v.f) out = PTup{ex.t.nb, ex.t.np, 0, INT32_MAX, 0, 0};
    a.pre[t] = out;
  }
}

// Batched parser set-up: per document, positions T + 2 when the lexer accepted it, in tiles of
// PK_TILE; the first tile of every document (exclusive sum) and the tile count.
__global__ void __launch_bounds__(PS_THREADS) pb_setup(PArgs a, const LexDoc *ldocs,
                                                        const uint32_t *doc_res,
                                                        uint64_t *ntiles_out) {
  __shared__ uint32_t s_w[PS_THREADS / WAVE];
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  const uint32_t D = a.ndocs;
  const uint32_t run = (D + PS_THREADS - 1) / PS_THREADS;
  const uint32_t lo = std::min<uint32_t>(D, threadIdx.x * run);
  const uint32_t hi = std::min<uint32_t>(D, lo + run);
  uint32_t acc = 0;
  for (uint32_t d = lo; d < hi; ++d) {
    const uint32_t ok = doc_res[2 * d];
    const uint32_t T = doc_res[2 * d + 1];
    acc += ok ? (uint32_t)((T + 2 + PK_TILE - 1) / PK_TILE) : 0u;
  }
  const uint32_t incl = (uint32_t)wave_incl_i32((int32_t)acc, 0, OpAdd());
  if (lane == WAVE - 1) s_w[w] = incl;
  __syncthreads();
  uint32_t base = incl - acc;
  for (int i = 0; i < w; ++i) base += s_w[i];
  for (uint32_t d = lo; d < hi; ++d) {
    const uint32_t ok = doc_res[2 * d];
    const uint32_t T = doc_res[2 * d + 1];
    PDoc pd;
    pd.tok_base = ldocs[d].tok_base;
    pd.br_off = pd.node_off = pd.B = pd.N = 0;
    pd.T = T;
    pd.lex_ok = ok;
    pd.ptile_base = base;
    pd.ptiles = ok ? (uint32_t)((T + 2 + PK_TILE - 1) / PK_TILE) : 0u;
    pd.status = ok;
    pd.subst = 0;
    a.pdocs[d] = pd;
    a.doc_flags[d] = 0;
    base += pd.ptiles;
  }
  if (threadIdx.x == PS_THREADS - 1) *ntiles_out = base;
}

// Batched parser: the document of every position tile.
__global__ void pb_tile_docs(PArgs a) {
  const PDoc &doc = a.pdocs[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < doc.ptiles; i += blockDim.x)
    const_cast<uint32_t *>(a.ptile_doc)[doc.ptile_base + i] = blockIdx.x | (i == 0 ? 0x80000000u : 0u);
}

// Batched parser: per document element ranges, the depth check (depths, parser.fut:153-166) and
// whether the terminal productions pair up with the tokens (safe_zip, :300-303).
__global__ void pb_docs(PArgs a) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= a.ndocs) return;
  PDoc pd = a.pdocs[d];
  if (pd.ptiles == 0) {
    // no positions: the offsets where the document's elements would start keep them monotone
    const uint64_t ntiles = *a.ntiles_dev;
    const PTup at = pd.ptile_base < ntiles ? a.pre[pd.ptile_base] : *a.total;
    pd.br_off = (uint32_t)at.nb;
    pd.node_off = (uint32_t)at.np;
    pd.status = 0;
    a.pdocs[d] = pd;
    return;
  }
  const uint32_t f = pd.ptile_base, l = f + pd.ptiles - 1;
  const PTup first = a.pre[f];
  PTup inc = ptup_op(a.pre[l], a.tiles[l]);
  pd.br_off = (uint32_t)first.nb;
  pd.node_off = (uint32_t)first.np;
  pd.B = (uint32_t)(inc.nb - first.nb);
  pd.N = (uint32_t)(inc.np - first.np);
  bool ok = pd.lex_ok && (a.doc_flags[d] & 1u) == 0;
  if (inc.mind < 0 || inc.dd != 0) ok = false;  // B counts the brackets left after cancellation
  pd.status = ok ? 1u : 0u;
  pd.subst = (uint32_t)inc.nt == pd.T ? 1u : 0u;
  a.pdocs[d] = pd;
}

constexpr int PX_STAGE = 256;  // element slots per wavefront and round

// LDS slot of position record q of a wavefront (one padding record after every 16)
__device__ __forceinline__ int px_pad(int q) { return q + (q >> 4); }
constexpr int PX_TOK = 320;    // token spans per wavefront staged in LDS (32-bit spans)

// dynamic LDS of parse_expand2: stackw (u32, padded to 8 bytes), prodw (uint2), entry infos at a
// 16-byte boundary, then the entries' cancellation counts (u32)
__host__ __device__ inline size_t px_einfo_offset(size_t nstack, size_t nprod) {
  return (4 * ((nstack + 1) & ~(size_t)1) + 8 * nprod + 15) & ~(size_t)15;
}
__host__ __device__ inline size_t px_lds_bytes(size_t nstack, size_t nprod, size_t E) {
  return px_einfo_offset(nstack, nprod) + (sizeof(EntInfo) + sizeof(uint32_t)) * E;
}
constexpr size_t PX_TABLE_LDS_MAX = 32 << 10;  // stackw + prodw + entry infos, copied to LDS per block (64 KiB with the static LDS)

// Per tile: the positions' prefixes (block scan), then the brackets and productions of every
// wavefront's positions, written coalesced: each round covers a 64-aligned window of PX_STAGE
// element slots; every position whose elements begin inside the window marks its first slot in
// LDS, and a running max-scan over the marks gives every slot its position (a load-balanced
// expansion: one lane per element, whatever the positions' element counts). Chunks of 64 elements
// shared with a neighbouring wavefront take their minimum through atomicMin on an array cleared to
// INT32_MAX.
// Five workgroups of 4 waves per CU (84 VGPRs): without the hint the compiler settles at 106
// VGPRs (4 waves per SIMD) for the batched and device-sized forms.
template <typename S, bool BATCH, bool STRIDE = false>
__global__ void __launch_bounds__(PK_THREADS) __attribute__((amdgpu_waves_per_eu(STRIDE ? 4 : 5)))
parse_expand2(PArgs a, const S *starts, const S *ends,
                                                            int32_t *depth, uint16_t *br,
                                                            int32_t *dl1, int32_t *sx, int32_t *sl1,
                                                            uint32_t *node, S *span,
                                                            int write_nodes) {
  __shared__ PTup s_w[PK_THREADS / WAVE];
  // per position of the wavefront, position q at q + q / 16: the 16-byte records a lane writes
  // (4 consecutive positions) then start 5 records apart from its neighbour's, so the lanes' stores
  // spread over all banks (unpadded they were 64 bytes apart: 16 lanes on each bank group)
  __shared__ uint4 s_pos[PK_THREADS / WAVE][WAVE * PK_IPT + WAVE * PK_IPT / 16];
  __shared__ __align__(16) uint16_t s_mark[PK_THREADS / WAVE][PX_STAGE];    // 1 + position starting at a slot
  // the spans of the tokens the wavefront's terminal productions take (a contiguous rank range)
  constexpr int NTOK = sizeof(S) == 4 ? PX_TOK : 1;
  __shared__ uint2 s_tok[PK_THREADS / WAVE][NTOK];
  extern __shared__ __align__(16) uint8_t px_smem[];  // stackw, prodw, then the entry infos
  uint32_t *s_stackw = reinterpret_cast<uint32_t *>(px_smem);
  uint2 *s_prodw = reinterpret_cast<uint2 *>(px_smem + 4 * ((a.nstack + 1) & ~1u));
  EntInfo *s_einfo = reinterpret_cast<EntInfo *>(px_smem + px_einfo_offset(a.nstack, a.nprod));
  uint32_t *s_can = reinterpret_cast<uint32_t *>(s_einfo + a.E);
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  // One block per position tile. The batched tile count is known on the device only: the grid
  // covers an estimate of it, and blocks past the count return at once; STRIDE: the tiles from
  // a.expand_tile0 on that the estimate missed, by a grid of a few blocks per CU striding over
  // them (none in the rule: the estimate grows to what calls needed). A batch that does not fit
  // its element scratch writes nothing.
  const uint64_t t_first = STRIDE ? a.expand_tile0 + blockIdx.x : (uint64_t)blockIdx.x;
  const uint64_t t_end = STRIDE ? (batch_elems_fit(a) ? *a.ntiles_dev : 0) : t_first + 1;
  for (uint64_t tile = t_first; tile < t_end; tile += STRIDE ? gridDim.x : 1) {
  if (BATCH && !STRIDE) {
    if (tile >= *a.ntiles_dev || !batch_elems_fit(a)) return;
  }
  const uint64_t pbase = tile * PK_TILE + (uint64_t)threadIdx.x * PK_IPT;  // slot
  // the lane's four entries: the entry array holds every tile whole, with PK_NO_ENTRY past the
  // document's positions (parse_keys2), so the entries alone tell which positions are there, and
  // the document's descriptor (and the device token count) are needed only after the block scan
  const uint2 packed = *reinterpret_cast<const uint2 *>(a.ent + pbase);
  const PosView pv = pos_view<BATCH>(a, tile);
  const int subst = pv.subst;
  // batched: the first node of the tile's document and of the next one, for the chunk words
  uint32_t doc_node0 = 0, next_node0 = 0xFFFFFFFFu;
  if (BATCH) {
    doc_node0 = (uint32_t)a.pdocs[pv.doc].node_off;
    if (pv.doc + 1 < a.ndocs) next_node0 = (uint32_t)a.pdocs[pv.doc + 1].node_off;
  }
  const uint16_t ids[PK_IPT] = {(uint16_t)packed.x, (uint16_t)(packed.x >> 16), (uint16_t)packed.y,
                                (uint16_t)(packed.y >> 16)};
  // the tables' copy into LDS (once per block) overlaps the loads of the entries above
  if (tile == t_first) {
    for (uint32_t i = threadIdx.x; i < a.nstack; i += blockDim.x) s_stackw[i] = a.stackw[i];
    for (uint32_t i = threadIdx.x; i < a.nprod; i += blockDim.x) s_prodw[i] = a.prodw[i];
    for (uint32_t i = threadIdx.x; i < a.E; i += blockDim.x) {
      s_einfo[i] = a.einfo[i];
      s_can[i] = a.ecan[i];
    }
  }
  const PTup pre = a.pre[tile];
  __syncthreads();
  EntInfo inf[PK_IPT];
  uint32_t tlr[PK_IPT];
  PTup acc = ptup_id();
#pragma unroll
  for (int i = 0; i < PK_IPT; ++i) {
    const bool has = ids[i] != PK_NO_ENTRY;
    inf[i] = has ? s_einfo[ids[i]] : EntInfo{0, 0, 0, INT16_MAX, 0, 0, 0, 0};
    tlr[i] = has ? s_can[ids[i]] : 0u;
    acc = ptup_op(acc, ptup_of(inf[i]));
  }
  // both levels of cancellation, as in parse_keys2: the brackets every position writes, from its
  // first remaining one (skip[i] leading rights skipped)
  int32_t cr[PK_IPT], cl[PK_IPT], nbw[PK_IPT], skip[PK_IPT];
  int32_t r1[PK_IPT], l1[PK_IPT], cr2[PK_IPT], cl2[PK_IPT];
  uint32_t f1[PK_IPT], nf[PK_IPT];
  pk_cancel(tlr, cr, cl[0]);
  acc.nb -= pk_cancelled(cr, cl[0]);
#pragma unroll
  for (int i = 0; i < PK_IPT; ++i) {
    if (i > 0) cl[i] = cr[i - 1];
    nbw[i] = (int32_t)inf[i].nb - cl[i] - cr[i];
    r1[i] = (int32_t)((tlr[i] >> 8) & 0xFFu) - cl[i];
    l1[i] = (int32_t)(tlr[i] & 0xFFu) - cr[i];
    f1[i] = (uint32_t)inf[i].sst + (uint32_t)cl[i];
  }
  pk_cancel2(nbw, r1, l1, f1, lane, cr2, nf, cl2);
#pragma unroll
  for (int i = 0; i < PK_IPT; ++i) {
    // the lane's own brackets (its positions' offsets come from a scan over the lanes; the tile
    // total equals parse_keys2's, which counts both sides of a pair where the left one sits)
    acc.nb -= cl2[i] + cr2[i];
    nbw[i] -= cl2[i] + cr2[i];
    skip[i] = cl[i] + cl2[i];
  }
  // the cancelled pairs must match: the j-th trailing left of a position and the j-th leading
  // right of the next one; at the second level, the j-th remaining trailing left and the j-th
  // remaining leading right of the next position with brackets left
  {
    uint32_t nsst[PK_IPT];  // first bracket of the next position's entry
#pragma unroll
    for (int i = 0; i + 1 < PK_IPT; ++i) nsst[i] = inf[i + 1].sst;
    nsst[PK_IPT - 1] = (uint32_t)dpp_i32<DPP_WAVE_SHL1>(0, (int32_t)inf[0].sst);
    bool mism = false;
#pragma unroll
    for (int i = 0; i < PK_IPT; ++i) {
      const uint32_t last = (uint32_t)inf[i].sst + inf[i].nb - 1u;
      for (int32_t j = 0; j < cr[i]; ++j)
        mism |= ((s_stackw[last - (uint32_t)j] ^ s_stackw[nsst[i] + (uint32_t)j]) & 0x7FFFu) != 0u;
      const uint32_t last2 = last - (uint32_t)cr[i];
      for (int32_t j = 0; j < cr2[i]; ++j)
        mism |= ((s_stackw[last2 - (uint32_t)j] ^ s_stackw[nf[i] + (uint32_t)j]) & 0x7FFFu) != 0u;
    }
    if (__ballot(mism) != 0ull && lane == 0) {
      if (BATCH) atomicOr(a.doc_flags + pv.doc, 2u);
      else atomicOr(a.flags, 1u);
    }
  }
  PTup total;
  const PTup ex = ptup_block_excl(acc, s_w, total);  // also orders the table copies above
  // element prefixes of this lane's positions (global)
  int32_t ob[PK_IPT], op[PK_IPT];
  uint4 *const pos = s_pos[w];
  uint16_t *const mark = s_mark[w];
  {
    int32_t b = pre.nb + ex.nb, q = pre.np + ex.np, d = pre.dd + ex.dd;
#pragma unroll
    for (int i = 0; i < PK_IPT; ++i) {
      ob[i] = b;
      op[i] = q;
      pos[px_pad(lane * PK_IPT + i)] = make_uint4((uint32_t)b, (uint32_t)d, (uint32_t)(inf[i].sst + skip[i]), 0u);
      b += nbw[i];
      q += inf[i].np;
      d += inf[i].dd;
    }
  }
  const int32_t lb_hi = ob[PK_IPT - 1] + nbw[PK_IPT - 1];  // this lane's brackets: [ob[0], lb_hi)
  const int32_t lp_hi = op[PK_IPT - 1] + inf[PK_IPT - 1].np;
  // ranks of the wavefront's terminal productions: [rw_lo, rw_hi); their spans go to LDS now, so
  // that the production rounds below read no global memory
  int32_t rw_lo = 0;
  bool tok_lds = false;
  if (sizeof(S) == 4 && write_nodes && subst) {
    int32_t rl = pre.nt + ex.nt, rh = rl;
#pragma unroll
    for (int i = 0; i < PK_IPT; ++i) rh += inf[i].nt;
    rw_lo = __shfl(rl, 0, WAVE);
    const int32_t cnt = __shfl(rh, WAVE - 1, WAVE) - rw_lo;
    tok_lds = cnt <= PX_TOK;
    if (tok_lds) {
      uint32_t ts[PX_TOK / WAVE], te[PX_TOK / WAVE];
#pragma unroll
      for (int u = 0; u < PX_TOK / WAVE; ++u) {
        const int32_t t = u * WAVE + lane;
        ts[u] = te[u] = 0;
        if (t < cnt) {
          const uint64_t gk = pv.tok_base + (uint32_t)(rw_lo + t);
          ts[u] = (uint32_t)starts[gk];
          te[u] = (uint32_t)ends[gk];
        }
      }
#pragma unroll
      for (int u = 0; u < PX_TOK / WAVE; ++u) s_tok[w][u * WAVE + lane] = make_uint2(ts[u], te[u]);
    }
  }

  // brackets: depth value (depths, parser.fut:153-166) and symbol
  {
    const int32_t wlo = __shfl(ob[0], 0, WAVE), whi = __shfl(lb_hi, WAVE - 1, WAVE);
    int32_t carry = 0;
    for (int32_t r0w = wlo & ~(WAVE - 1); r0w < whi; r0w += PX_STAGE) {
      reinterpret_cast<uint2 *>(mark)[lane] = make_uint2(0u, 0u);
      wave_lds_sync();
#pragma unroll
      for (int i = 0; i < PK_IPT; ++i)
        if (nbw[i] > 0 && ob[i] >= r0w && ob[i] < r0w + PX_STAGE)
          mark[ob[i] - r0w] = (uint16_t)(lane * PK_IPT + i + 1);
      wave_lds_sync();
#pragma unroll
      for (int j = 0; j < PX_STAGE / WAVE; ++j) {
        const int32_t cb = r0w + j * WAVE;
        if (cb >= whi) break;
        const int32_t g = cb + lane;
        const int32_t o = max(wave_incl_i32((int32_t)mark[j * WAVE + lane], 0, OpMax()), carry);
        carry = __builtin_amdgcn_readlane(o, WAVE - 1);
        const bool act = g >= wlo && g < whi;
        int32_t dv = INT32_MAX;
        if (act) {
          const uint4 p = pos[px_pad(o - 1)];
          const uint32_t sw = s_stackw[p.z + (uint32_t)(g - (int32_t)p.x)];
          dv = (int32_t)p.y + (int32_t)(int16_t)(sw >> 16);
          __builtin_nontemporal_store(dv, depth + g);  // streamed: see the productions below
          __builtin_nontemporal_store((uint16_t)sw, br + g);
        }
        const int32_t m = wave_min_i32(dv);
        if (lane == 0) {
          if (cb >= wlo && cb + WAVE <= whi) dl1[cb / WAVE] = m;
          else atomicMin(dl1 + cb / WAVE, m);
        }
      }
      wave_lds_sync();
    }
  }
  {
    int32_t q = pre.np + ex.np, sv = pre.ds + ex.ds, r = pre.nt + ex.nt;
#pragma unroll
    for (int i = 0; i < PK_IPT; ++i) {
      pos[px_pad(lane * PK_IPT + i)] = make_uint4((uint32_t)q, (uint32_t)sv, (uint32_t)r, inf[i].pst);
      q += inf[i].np;
      sv += inf[i].ds;
      r += inf[i].nt;
    }
  }
  // productions: s (parents, parser.fut:286-296), node and span (terminal_offsets, :305-332)
  {
    const int32_t wlo = __shfl(op[0], 0, WAVE), whi = __shfl(lp_hi, WAVE - 1, WAVE);
    const bool root_wave = pv.reset && w == 0;  // the wavefront holding the document's first node
    int32_t carry = 0;
    for (int32_t r0w = wlo & ~(WAVE - 1); r0w < whi; r0w += PX_STAGE) {
      reinterpret_cast<uint2 *>(mark)[lane] = make_uint2(0u, 0u);
      wave_lds_sync();
#pragma unroll
      for (int i = 0; i < PK_IPT; ++i)
        if (inf[i].np && op[i] >= r0w && op[i] < r0w + PX_STAGE)
          mark[op[i] - r0w] = (uint16_t)(lane * PK_IPT + i + 1);
      wave_lds_sync();
      // every chunk's values and span loads first, then the stores: the loads of later chunks
      // do not wait behind the stores of earlier ones
      constexpr int NJ = PX_STAGE / WAVE;
      int32_t svs[NJ];
      uint32_t nds[NJ];
      S v0[NJ], v1[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int32_t cb = r0w + j * WAVE;
        const int32_t g = cb + lane;
        const int32_t o = max(wave_incl_i32((int32_t)mark[j * WAVE + lane], 0, OpMax()), carry);
        carry = __builtin_amdgcn_readlane(o, WAVE - 1);
        svs[j] = INT32_MAX;
        nds[j] = 0;
        v0[j] = 0;
        v1[j] = 0;
        if (g >= wlo && g < whi) {
          const uint4 p = pos[px_pad(o - 1)];
          const uint2 pw = s_prodw[p.w + (uint32_t)(g - (int32_t)p.x)];
          const int32_t t = subst ? (int32_t)(pw.x >> 16) - 1 : -1;
          // the document's root (its first production) gets s = -1: every other node has s >= 0 in
          // a valid tree, so the root still ends every node's search, and the search kernel
          // tells it by its value, with no document offsets
          svs[j] = root_wave && g == wlo ? -1 : (int32_t)p.y + (int32_t)(int16_t)(pw.y & 0xFFFFu);
          nds[j] = t >= 0 ? (ALPACC_NODE_TERMINAL | (uint32_t)t) : (pw.x & 0xFFFFu);
          if (write_nodes && t >= 0) {
            const int32_t rk = (int32_t)p.z + (int32_t)(pw.y >> 16);
            if (tok_lds) {
              const uint2 sp = s_tok[w][rk - rw_lo];
              v0[j] = (S)sp.x;
              v1[j] = (S)sp.y;
            } else {
              const uint64_t gk = pv.tok_base + (uint32_t)rk;
              v0[j] = starts[gk];
              v1[j] = ends[gk];
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int32_t cb = r0w + j * WAVE;
        if (cb >= whi) break;
        const int32_t g = cb + lane;
        if (g >= wlo && g < whi) {
          // every element array is streamed out with non-temporal stores: 1.9 GB per 256 MiB of
          // JSON passes through the caches before the searches could reuse any of it (measured:
          // 0.77 -> 0.67 ms here, the searches slightly faster too)
          __builtin_nontemporal_store(svs[j], sx + g);
          if (write_nodes) {
            __builtin_nontemporal_store(nds[j], node + g);
            if (sizeof(S) == 4) {
              __builtin_nontemporal_store((uint64_t)(uint32_t)v0[j] | ((uint64_t)(uint32_t)v1[j] << 32),
                                          reinterpret_cast<uint64_t *>(span) + g);
            } else {
              span[2 * (uint64_t)g] = v0[j];
              span[2 * (uint64_t)g + 1] = v1[j];
            }
          }
        }
        const int32_t m = wave_min_i32(svs[j]);
        if (lane == 0) {
          if (cb >= wlo && cb + WAVE <= whi) sl1[cb / WAVE] = m;
          else atomicMin(sl1 + cb / WAVE, m);
          if (BATCH && cb >= wlo) {
            a.chunk_doc[cb / WAVE] = pv.doc;
            a.chunk_base[cb / WAVE] =
                doc_node0 | ((uint64_t)next_node0 < (uint64_t)cb + WAVE ? 0x80000000u : 0u);
          }
        }
      }
      wave_lds_sync();
    }
  }
  if (STRIDE) __syncthreads();  // the block's waves start the next tile together
  }  // tiles
}

// chunk minima of both element streams start at INT32_MAX (chunks shared by two wavefronts take
// atomicMin)
__global__ void elem_init(int32_t *dl1, uint64_t n1, int32_t *sl1, uint64_t n2) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  for (uint64_t i = t; i < n1; i += stride) dl1[i] = INT32_MAX;
  for (uint64_t i = t; i < n2; i += stride) sl1[i] = INT32_MAX;
}

// brackets_matches (parser.fut:174-184): every right bracket equals (without the left bit) the
// bracket at its previous smaller-or-equal depth. One wavefront per 64 brackets; a match inside
// the wavefront's own chunk is read from a register.
template <bool BATCH>
__device__ void bracket_match_run(const uint16_t *br, uint64_t B, const Levels &lv, uint32_t *flags,
                                  const PArgs &a, uint64_t blk, uint64_t nblk) {
  const int lane = threadIdx.x & (WAVE - 1);
  if (BATCH) B = batch_elems_fit(a) ? std::min<uint64_t>(B, (uint32_t)a.total->nb) : 0;
  const uint64_t nch = (B + WAVE - 1) / WAVE;
  const uint64_t wstride = nblk * (blockDim.x / WAVE);
  bool bad = false;
  // two chunks per iteration: their in-chunk lifts (dependent permute chains) overlap
  for (uint64_t c0 = blk * (blockDim.x / WAVE) + threadIdx.x / WAVE; c0 < nch; c0 += 2 * wstride) {
    const uint64_t cc[2] = {c0, c0 + wstride};
    uint16_t b[2];
    int32_t v[2];
    SearchRev rv[2];
    int p0[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint64_t i = cc[h] * WAVE + lane;
      const bool valid = cc[h] < nch && i < B;
      b[h] = valid ? br[i] : (uint16_t)0x8000u;
      v[h] = valid ? lv.L[0][i] : INT32_MAX;
      rv[h] = cc[h] < nch ? search_rev(lv, cc[h], lane) : SearchRev{INT32_MAX, INT32_MAX};
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) p0[h] = lift_left(v[h], lane, v[h]);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint64_t c = cc[h];
      if (c >= nch) break;  // uniform in the wave
      const uint64_t i = c * WAVE + lane;
      const bool right = i < B && !(b[h] & 0x8000u);
      const int64_t j = prev_le_wave(lv, c, lane, v[h], rv[h], right, p0[h]);
      const int64_t jl = j - (int64_t)(c * WAVE);
      const uint16_t bin = (uint16_t)__shfl((int)b[h], (int)(jl & (WAVE - 1)), WAVE);
      bool mism = false;
      if (right) {
        if (j < 0) {
          mism = true;
        } else {
          const uint16_t bj = (jl >= 0 && jl < WAVE) ? bin : br[j];
          mism = (bj & 0x7FFFu) != (b[h] & 0x7FFFu);
        }
      }
      if (BATCH && mism) {
        // the document holding bracket i: the last d with br_off <= i (offsets are monotone)
        uint32_t lo = 0, hi = a.ndocs;
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) / 2;
          if (a.pdocs[mid].br_off <= i) lo = mid;
          else hi = mid;
        }
        atomicOr(a.doc_flags + lo, 2u);
      }
      bad |= mism;
    }
  }
  if (!BATCH && __ballot(bad) != 0ull && lane == 0) atomicOr(flags, 1u);
}

// parents (parser.fut:286-296): previous smaller-or-equal s, the root its own parent.
// Batched: parents are relative to the document's first node, every document's first node is
// its own parent.
template <bool BATCH>
__device__ void parents_run(uint64_t N, const Levels &lv, uint32_t *parent, const PArgs &a,
                            uint64_t blk, uint64_t nblk) {
  const int lane = threadIdx.x & (WAVE - 1);
  if (BATCH) N = batch_elems_fit(a) ? std::min<uint64_t>(N, (uint32_t)a.total->np) : 0;
  const uint64_t nch = (N + WAVE - 1) / WAVE;
  const uint64_t wstride = nblk * (blockDim.x / WAVE);
  // one chunk per iteration: this search is bound by instruction issue, and a second chunk in
  // flight (as in bracket_match) costs more in occupancy than it gains in overlap
  for (uint64_t c = blk * (blockDim.x / WAVE) + threadIdx.x / WAVE; c < nch; c += wstride) {
    const uint64_t i = c * WAVE + lane;
    const bool valid = i < N;
    // batched: the chunk's word (the first node of its first node's document), loaded now and
    // used after the search, so that its latency hides behind the search's
    uint32_t cw = 0;
    if (BATCH) cw = a.chunk_base[c];
    const int32_t v = valid ? lv.L[0][i] : INT32_MAX;
    const SearchRev rv = search_rev(lv, c, lane);
    // a document's root has s = -1 (every other node s >= 0): it searches nothing and is its own
    // parent
    const int64_t j = prev_le_wave(lv, c, lane, v, rv, valid && v >= 0, lift_left(v, lane, v));
    if (valid) {
      uint64_t base = 0;
      if (BATCH) {
        base = cw & 0x7FFFFFFFu;
        if (cw >> 31) {  // a later document starts inside the chunk: forward from the chunk's
          uint32_t d = a.chunk_doc[c];
          while (d + 1 < a.ndocs && a.pdocs[d + 1].node_off <= i) ++d;
          base = a.pdocs[d].node_off;
        }
      }
      __builtin_nontemporal_store(v < 0 ? 0u : (uint32_t)((uint64_t)j - base), parent + i);
    }
  }
}

inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 1u << 16) {
  uint64_t g = (n + block - 1) / block;
  if (g == 0) g = 1;
  return (unsigned)std::min<uint64_t>(g, cap);
}

// lex_reduce / lex_emit with the table placement chosen at context creation
template <bool BATCH>
void launch_lex_reduce(int tab, unsigned blocks, size_t lds, const LexArgs &a, hipStream_t st) {
  if (tab == 2) lex_reduce<2, BATCH><<<blocks, LEX_BLOCK, lds, st>>>(a);
  else if (tab == 1) lex_reduce<1, BATCH><<<blocks, LEX_BLOCK, lds, st>>>(a);
  else lex_reduce<0, BATCH><<<blocks, LEX_BLOCK, lds, st>>>(a);
}

template <bool CHUNKED, bool BATCH>
void launch_lex_emit(bool dfa_lds, unsigned grid, const LexArgs &a, hipStream_t st) {
  if (dfa_lds) lex_emit<CHUNKED, BATCH, true><<<grid, LEX_BLOCK, 0, st>>>(a);
  else lex_emit<CHUNKED, BATCH, false><<<grid, LEX_BLOCK, 0, st>>>(a);
}

// BS: workgroup size (one 64-element chunk per wavefront and iteration, whatever the size)
template <bool BATCH, int BS>
__global__ void __launch_bounds__(BS) bracket_match(const uint16_t *br, uint64_t B, Levels lv,
                                                    uint32_t *flags, PArgs a) {
  bracket_match_run<BATCH>(br, B, lv, flags, a, blockIdx.x, gridDim.x);
}

template <bool BATCH, int BS>
__global__ void __launch_bounds__(BS) parse_parents(uint64_t N, Levels lv, uint32_t *parent,
                                                    PArgs a) {
  parents_run<BATCH>(N, lv, parent, a, blockIdx.x, gridDim.x);
}

// The search kernels with workgroups of `bs` threads (256, 512 or 1024) over n elements, at most
// `cap` workgroups of 256 threads' worth.
template <bool BATCH>
void launch_bracket_match(unsigned bs, unsigned cap, const uint16_t *br, uint64_t B,
                          const Levels &lv, uint32_t *flags, const PArgs &a, hipStream_t st) {
  const unsigned g = std::max(1u, grid_for(B, 256, cap) / (bs / 256));
  if (bs == 1024) bracket_match<BATCH, 1024><<<g, 1024, 0, st>>>(br, B, lv, flags, a);
  else if (bs == 512) bracket_match<BATCH, 512><<<g, 512, 0, st>>>(br, B, lv, flags, a);
  else bracket_match<BATCH, 256><<<g, 256, 0, st>>>(br, B, lv, flags, a);
}

template <bool BATCH>
void launch_parse_parents(unsigned bs, unsigned cap, uint64_t N, const Levels &lv,
                          uint32_t *parent, const PArgs &a, hipStream_t st) {
  const unsigned g = std::max(1u, grid_for(N, 256, cap) / (bs / 256));
  if (bs == 1024) parse_parents<BATCH, 1024><<<g, 1024, 0, st>>>(N, lv, parent, a);
  else if (bs == 512) parse_parents<BATCH, 512><<<g, 512, 0, st>>>(N, lv, parent, a);
  else parse_parents<BATCH, 256><<<g, 256, 0, st>>>(N, lv, parent, a);
}

}  // namespace

// ----------------------------------------------------------------------------------------------
// host side

static std::string g_create_error;

enum Stage {
  S_LEX_REDUCE = 0,
  S_LEX_SCAN,
  S_LEX_EMIT,
  S_PARSE,          // the whole parser stage
  S_PARSE_KEYS,     // parse_keys2 + the tile scan
  S_PARSE_EXPAND,   // parse_expand2
  S_BRACKET_MATCH,  // bracket_match
  S_PARSE_PARENTS,  // parse_parents
  NSTAGE
};
static const char *kStageNames[NSTAGE] = {"lex_reduce",    "lex_scan",     "lex_emit",
                                          "parse",         "parse_keys",   "parse_expand",
                                          "bracket_match", "parse_parents"};

struct alpacc_ctx {
  int device = 0;
  int cus = 0;
  std::string err;
  int launches = 0;
  // testing hook: raise the lexer's look-back time-out flag after it is cleared
  bool force_timeout = false;
  // lexer tables
  bool has_lexer = false;
  uint32_t e = 0, C = 0, m1 = 0, init = 0, dead = 0;
  int32_t ignore = -1;
  uint8_t *d_cls = nullptr;
  uint16_t *d_step = nullptr, *d_comp = nullptr;
  uint32_t *d_walk = nullptr;
  uint8_t *d_apply = nullptr;
  int lex_tab = 0;       // lex_reduce's table placement (see its TAB)
  bool dfa_lds = true;   // lex_emit's DFA table in LDS
  size_t lex_lds = 0;    // lex_reduce's dynamic LDS
  int lex_tables_lds = 1;  // tuning: 0 keeps every lexer table in L2
  // parser tables
  bool has_parser = false;
  uint32_t q = 0, k = 0, nT = 0, start_t = 0, end_t = 0, empty_t = 0, E = 0;
  uint32_t *d_pair_index = nullptr;  // dense index, u32 entries
  uint8_t *d_index8 = nullptr;       // the same as u8 (E < 255 and small enough for LDS)
  size_t index_len = 0;
  uint64_t *d_hkeys = nullptr;       // the hashed pair table (no dense index)
  uint32_t *d_hvals = nullptr;
  uint32_t hbits = 0, hprobe = 0;
  EntInfo *d_einfo = nullptr;
  uint32_t *d_stackw = nullptr;
  uint2 *d_prodw = nullptr;
  uint32_t *d_ecan = nullptr;  // per entry: trailing lefts | leading rights << 8 | brackets between << 16
  uint32_t max_nb = 0, max_np = 0;
  size_t nstackw = 0, nprodw = 0;
  // scratch (grow-only)
  std::vector<std::pair<void *, size_t>> scratch;
  uint64_t table_bytes = 0;  // device bytes of the uploaded tables
  void *noalloc_scratch = nullptr;  // device address of this device's allocation-free scratch
  // pinned host staging (grow-only) of the batched calls: descriptors in, results out
  uint8_t *h_stage = nullptr;
  size_t h_stage_len = 0;
  // batched lex + parse: HIP graph of the last call's queue, replayed while the plan repeats
  int graph_mode = 1;                 // 0 never, 1 from the second identical plan on, 2 always
  hipStream_t cap_stream = nullptr;   // capture stream (default priority)
  hipGraphExec_t gexec = nullptr;
  std::vector<uint8_t> gplan;         // the plan gexec was captured from
  std::vector<uint8_t> last_plan;     // the previous call's plan (graph_mode 1)
  uint64_t graph_launches = 0, graph_captures = 0;
  // brackets per input byte the batched scratch is sized for (grows with what calls need)
  double b_per_byte = 0.25;
  // key positions per input byte the batched expansion's grid is sized for (grows likewise)
  double p_per_byte = 0.25;
  // launch shapes (alpacc_set_tuning): persistent lex_reduce / parse_keys2 blocks per CU, and the
  // largest grid of the search kernels (16384 workgroups of 256 threads: 8 per SIMD's worth of
  // waves several times over; larger grids measured slower on a 256 MiB JSON document)
  uint32_t lex_bpc = 2, keys_bpc = 2, search_cap = 1u << 14, search_bs = 256;
  uint32_t *h_small = nullptr;  // pinned host words for results
  // profiling: HIP events on the call's stream around each stage
  bool prof = false;
  hipEvent_t ev[NSTAGE][2] = {};
  bool rec[NSTAGE] = {};
  double prof_ms[NSTAGE] = {};
  uint64_t prof_calls[NSTAGE] = {};

  void prof_begin(int s, hipStream_t st) {
    if (prof) (void)hipEventRecord(ev[s][0], st);
  }
  void prof_end(int s, hipStream_t st) {
    if (prof) {
      (void)hipEventRecord(ev[s][1], st);
      rec[s] = true;
    }
  }
  void prof_drop(int s) {
    rec[s] = false;
    if (s == S_PARSE)
      for (int i = S_PARSE_KEYS; i <= S_PARSE_PARENTS; ++i) rec[i] = false;
  }
  // element counts of the last call: tokens, positions, brackets, productions
  uint64_t last_counts[4] = {0, 0, 0, 0};
  void prof_collect() {
    for (int s = 0; s < NSTAGE; ++s) {
      float ms = 0.f;
      if (rec[s] && hipEventElapsedTime(&ms, ev[s][0], ev[s][1]) == hipSuccess) {
        prof_ms[s] += ms;
        prof_calls[s]++;
      }
      rec[s] = false;
    }
  }

  // pinned staging of at least `bytes`; growing it first waits for the stream
  uint8_t *stage(size_t bytes, hipStream_t st) {
    if (h_stage_len < bytes) {
      (void)hipStreamSynchronize(st);
      if (h_stage) (void)hipHostFree(h_stage);
      h_stage = nullptr;
      h_stage_len = 0;
      const size_t want = std::max<size_t>((bytes + 4095) & ~(size_t)4095, 64 << 10);
      if (hipHostMalloc((void **)&h_stage, want, hipHostMallocDefault) != hipSuccess) {
        h_stage = nullptr;
        return nullptr;
      }
      h_stage_len = want;
    }
    return h_stage;
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    gplan.clear();
  }

  void *buf(int id, size_t bytes) {
    if ((int)scratch.size() <= id) scratch.resize(id + 1, {nullptr, 0});
    auto &s = scratch[id];
    if (s.second < bytes) {
      if (s.first) (void)hipFree(s.first);
      s.first = nullptr;
      s.second = 0;
      size_t want = std::max<size_t>(bytes, 256);
      if (hipMalloc(&s.first, want) != hipSuccess) return nullptr;
      s.second = want;
    }
    return s.first;
  }
};

namespace {

enum BufId {
  B_SMALL = 0,
  B_LANE_EXCL,
  B_TILE_AGG,
  B_TILE_EXCL,
  B_BLOCK_AGG,
  B_TILE_STATE,
  B_ST_COUNT,
  B_TOK_KIND,
  B_TOK_START,
  B_TOK_END,
  B_ENT,
  B_PTILES,
  B_BR,
  B_DEPTH,
  B_LEVELS,
  B_LEVELS_S,
  B_SX,
  B_DL1,
  B_SL1,
  B_VIRT,
  B_CHUNK_INFO,
  B_TOK_START64,
  B_TOK_END64,
  B_DOCS,
  B_TILE_DOC,
  B_DOC_RES,
  B_PPRE,
  B_BTOT,
  B_PDOCS,
  B_DOC_FLAGS,
  B_PTILE_DOC,
  B_RESULTS,
  B_CHUNK_DOC,
  B_TILE_REC,
  B_CHUNKS,
  B_CHUNK_BASE,
};

template <typename T>
T *upload(const T *h, size_t count, std::string &err) {
  if (count == 0) count = 1;
  T *d = nullptr;
  if (hipMalloc(&d, count * sizeof(T)) != hipSuccess) {
    err = "hipMalloc failed for a table";
    return nullptr;
  }
  if (h && hipMemcpy(d, h, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
    err = "hipMemcpy failed for a table";
    return nullptr;
  }
  return d;
}

#define CK(expr)                                                        \
  do {                                                                  \
    hipError_t e_ = (expr);                                             \
    if (e_ != hipSuccess) {                                             \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);      \
      return ALPACC_ERR_HIP;                                            \
    }                                                                   \
  } while (0)

// reduce-then-scan of n int2 values from `src` into `sink`; the grand total goes to *total
// (copied to the host when total is not NULL).
// buf_id: the scratch arena of the upper levels; hierarchies in use at the same time need their own
// Lays out the 64-ary min levels above level 1 (l1, written by parse_expand2) in the arena
// buf_id; fill_levels computes them.
int plan_levels(alpacc_ctx *ctx, int32_t *base, int32_t *l1, uint64_t n, Levels &lv, int buf_id) {
  lv.nlev = 1;
  lv.L[0] = base;
  lv.len[0] = n;
  if (n <= 64) return 0;
  uint64_t len = (n + 63) / 64;
  lv.L[1] = l1;
  lv.len[1] = len;
  lv.nlev = 2;
  uint64_t total = 0, l = len;
  while (l > 64) {
    l = (l + 63) / 64;
    total += ((l + 3) & ~3ull) + 16;
  }
  int32_t *area = (int32_t *)ctx->buf(buf_id, (total + 16) * sizeof(int32_t));
  if (!area) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  uint64_t off = 0;
  while (len > 64) {
    if (lv.nlev >= MAX_LEVELS) {
      ctx->err = "input too large for the min levels";
      return ALPACC_ERR_ARG;
    }
    const uint64_t nl = (len + 63) / 64;
    lv.L[lv.nlev] = area + off;
    lv.len[lv.nlev] = nl;
    lv.nlev++;
    off += ((nl + 3) & ~3ull) + 16;
    len = nl;
  }
  return 0;
}

// Fills the levels that plan_levels laid out, for both hierarchies: level 2 across the GPU, the
// small levels above it in one block.
int fill_levels(alpacc_ctx *ctx, const Levels &la, const Levels &lb, hipStream_t st) {
  const uint64_t groups = (la.nlev > 2 ? la.len[2] : 0) + (lb.nlev > 2 ? lb.len[2] : 0);
  if (groups == 0) return 0;
  min_level2<<<grid_for(groups * WAVE, 256), 256, 0, st>>>(la, lb);
  ctx->launches++;
  if (la.nlev > 3 || lb.nlev > 3) {
    min_levels_top<<<1, 1024, 0, st>>>(la, lb);
    ctx->launches++;
  }
  CK(hipGetLastError());
  return 0;
}

// The lexer's scratch for one unchunked call over n bytes: result words, the lane-exclusive
// endomorphisms, the tile and scan-block endomorphisms, the tile start states and the look-back
// words of the count scan.
struct LexScratch {
  uint32_t *small;       // 16 words: [0] count-tile counter, [4..6] accept, tokens, time-out
  uint16_t *lane_excl;   // [tiles * 64]
  uint16_t *tile_agg;    // [tiles]
  uint16_t *tile_excl;   // [tiles]
  uint16_t *block_agg;   // [scan blocks]
  uint8_t *tile_state;   // [tiles]
  uint64_t *st_count;    // [count tiles]
};

constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// Bytes of a LexScratch carved out of one buffer (every part 256-byte aligned), and the carving.
constexpr uint64_t lex_scratch_bytes(uint64_t n) {
  return align256(64) +
         align256(((n + LEX_TILE - 1) / LEX_TILE) * WAVE * 2) +
         2 * align256(((n + LEX_TILE - 1) / LEX_TILE) * 2) +
         align256((((n + LEX_TILE - 1) / LEX_TILE + LEX_SCAN_BLOCK - 1) / LEX_SCAN_BLOCK) * 2) +
         align256((n + LEX_TILE - 1) / LEX_TILE) +
         align256((((n + LEX_TILE - 1) / LEX_TILE + LEX_CT_WAVES - 1) / LEX_CT_WAVES) * 8);
}

void lex_scratch_carve(uint8_t *base, uint64_t n, LexScratch &s) {
  const uint64_t tiles = (n + LEX_TILE - 1) / LEX_TILE;
  const uint64_t ctiles = (tiles + LEX_CT_WAVES - 1) / LEX_CT_WAVES;
  const uint64_t sblocks = (tiles + LEX_SCAN_BLOCK - 1) / LEX_SCAN_BLOCK;
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    uint8_t *p = base + off;
    off += align256(bytes);
    return p;
  };
  s.small = (uint32_t *)take(64);
  s.lane_excl = (uint16_t *)take(tiles * WAVE * 2);
  s.tile_agg = (uint16_t *)take(tiles * 2);
  s.tile_excl = (uint16_t *)take(tiles * 2);
  s.block_agg = (uint16_t *)take(sblocks * 2);
  s.tile_state = take(tiles);
  s.st_count = (uint64_t *)take(ctiles * 8);
}

}  // namespace

// The allocation-free lexer entry's scratch: one array per GPU, held until the process ends, for
// documents of up to NOALLOC_MAX_BYTES. Every alpacc_lex_noalloc call on a GPU uses that GPU's copy
// (calls that overlap on one GPU race on it). Host code never passes the array's name to a kernel
// or a copy: it is a host-side stand-in there. The device address comes from hipGetSymbolAddress,
// taken once per device (noalloc_scratch_setup, called by alpacc_create with the device current).
constexpr uint64_t NOALLOC_MAX_BYTES = 64ull << 20;
constexpr uint64_t NOALLOC_SCRATCH_BYTES = lex_scratch_bytes(NOALLOC_MAX_BYTES);
static_assert(NOALLOC_MAX_BYTES == ALPACC_NOALLOC_MAX_BYTES, "limit differs from the header's");
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on `device` (the current device); nullptr on failure.
void *noalloc_scratch_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[device]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_noalloc_ptr[device] = ptr;
  }
  return g_noalloc_ptr[device];
}

// The lexer. chunk == 0: unchunked (kinds/starts/ends as u32). chunk > 0: the reference's chunk
// loop with chunks of `chunk` bytes (kinds, starts64/ends64).
// sync = false (unchunked only): return ALPACC_ACCEPT once the kernels are queued; the verdict,
// the token count and the time-out flag stay in the device words result[0..2] (B_SMALL + 4).
// fixed (unchunked only): scratch carved from a caller-provided buffer; nothing is allocated.
int run_lex(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint32_t *kinds, uint32_t *starts,
            uint32_t *ends, uint64_t *num_tokens, hipStream_t st, uint64_t chunk = 0,
            uint64_t *starts64 = nullptr, uint64_t *ends64 = nullptr, bool sync = true,
            const LexScratch *fixed = nullptr, uint8_t *kinds8 = nullptr) {
  if (!ctx->has_lexer) {
    ctx->err = "bundle has no lexer";
    return ALPACC_ERR_ARG;
  }
  if (n == 0) {  // the identity state never accepts (lexer.fut:170-172)
    *num_tokens = 0;
    return ALPACC_REJECT;
  }
  const uint64_t n_orig = n;
  uint64_t K = 1;
  if (chunk > 0) {
    if (chunk > n) chunk = n;  // one chunk: the same as any chunk size >= n
    K = (n + chunk - 1) / chunk;
    n = n + K - 1;  // the virtual string
  }
  if (n >= (1ull << 31) - 2) {
    ctx->err = "documents must be shorter than 2^31 - 2 bytes";
    return ALPACC_ERR_ARG;
  }
  if (fixed && chunk > 0) {
    ctx->err = "the chunked lexer has no fixed-scratch form";
    return ALPACC_ERR_ARG;
  }
  if (chunk > 0) {
    uint8_t *virt = (uint8_t *)ctx->buf(B_VIRT, n + 16);
    if (!virt) {
      ctx->err = "scratch allocation failed";
      return ALPACC_ERR_HIP;
    }
    lex_chunk_expand<<<grid_for(n, 256, 1u << 14), 256, 0, st>>>(d_in, (uint32_t)chunk, n, virt);
    ctx->launches++;
    d_in = virt;
  }
  uint64_t *chunk_info = nullptr;
  if (chunk > 0) {
    chunk_info = (uint64_t *)ctx->buf(B_CHUNK_INFO, K * 4 * sizeof(uint64_t));
    if (!chunk_info) {
      ctx->err = "scratch allocation failed";
      return ALPACC_ERR_HIP;
    }
  }
  const uint64_t tiles = (n + LEX_TILE - 1) / LEX_TILE;
  const uint64_t ctiles = (tiles + LEX_CT_WAVES - 1) / LEX_CT_WAVES;
  const uint64_t sblocks = (tiles + LEX_SCAN_BLOCK - 1) / LEX_SCAN_BLOCK;
  LexScratch sc;
  if (fixed) {
    sc = *fixed;
  } else {
    sc.small = (uint32_t *)ctx->buf(B_SMALL, 64);
    sc.lane_excl = (uint16_t *)ctx->buf(B_LANE_EXCL, tiles * WAVE * sizeof(uint16_t));
    sc.tile_agg = (uint16_t *)ctx->buf(B_TILE_AGG, tiles * sizeof(uint16_t));
    sc.tile_excl = (uint16_t *)ctx->buf(B_TILE_EXCL, tiles * sizeof(uint16_t));
    sc.block_agg = (uint16_t *)ctx->buf(B_BLOCK_AGG, sblocks * sizeof(uint16_t));
    sc.tile_state = (uint8_t *)ctx->buf(B_TILE_STATE, tiles);
    sc.st_count = (uint64_t *)ctx->buf(B_ST_COUNT, ctiles * sizeof(uint64_t));
  }
  uint32_t *small = sc.small;
  uint16_t *lane_excl = sc.lane_excl, *tile_agg = sc.tile_agg, *tile_excl = sc.tile_excl;
  uint16_t *block_agg = sc.block_agg;
  uint8_t *tile_state = sc.tile_state;
  uint64_t *st_count = sc.st_count;
  if (!small || !lane_excl || !tile_agg || !tile_excl || !block_agg || !tile_state || !st_count) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  clear_u32<<<1, 64, 0, st>>>(small, 16);
  if (ctx->force_timeout) set_u32<<<1, 1, 0, st>>>(small + 6, 1u);
  clear_u32<<<grid_for(2 * ctiles, 256), 256, 0, st>>>((uint32_t *)st_count, 2 * ctiles);
  LexArgs a;
  a.in = d_in;
  a.n = n;
  a.kinds = kinds;
  a.kinds8 = chunk > 0 ? nullptr : kinds8;
  a.starts = starts;
  a.ends = ends;
  a.byte_class = ctx->d_cls;
  a.endo_step = ctx->d_step;
  a.endo_comp = ctx->d_comp;
  a.endo_apply = ctx->d_apply;
  a.dfa_walk = ctx->d_walk;
  a.e = ctx->e;
  a.C = ctx->C;
  a.m1 = ctx->m1;
  a.init = ctx->init;
  a.ignore = ctx->ignore;
  a.num_tiles = (uint32_t)tiles;
  a.num_ctiles = (uint32_t)ctiles;
  a.aligned = ((uintptr_t)d_in & 15) == 0 ? 1u : 0u;
  a.lane_excl = lane_excl;
  a.tile_agg = tile_agg;
  a.tile_excl = tile_excl;
  a.block_agg = block_agg;
  a.tile_state = tile_state;
  a.tile_counter = small;
  a.st_count = st_count;
  a.result = small + 4;
  a.chunk = (uint32_t)chunk;
  a.nchunks = (uint32_t)K;
  a.n_orig = n_orig;
  a.starts64 = starts64;
  a.ends64 = ends64;
  a.chunk_info = chunk_info;
  a.docs = nullptr;
  a.tile_doc = nullptr;
  a.tile_rec = nullptr;
  a.chunks = nullptr;
  a.ndocs = 0;
  a.doc_res = nullptr;
  const uint64_t wpb = LEX_BLOCK / WAVE;
  uint64_t blocks = (tiles + wpb - 1) / wpb;
  blocks = std::min<uint64_t>(blocks, (uint64_t)std::max(1, ctx->cus) * ctx->lex_bpc);
  ctx->prof_begin(S_LEX_REDUCE, st);
  launch_lex_reduce<false>(ctx->lex_tab, (unsigned)blocks, ctx->lex_lds, a, st);
  ctx->prof_end(S_LEX_REDUCE, st);
  ctx->prof_begin(S_LEX_SCAN, st);
  lex_scan_tiles<<<(unsigned)sblocks, LEX_SCAN_BLOCK, 0, st>>>(a);
  lex_scan_blocks<<<1, LEX_SCAN_BLOCK, 0, st>>>(a, (uint32_t)sblocks);
  lex_tile_states<<<grid_for(tiles, 256), 256, 0, st>>>(a);
  ctx->prof_end(S_LEX_SCAN, st);
  ctx->prof_begin(S_LEX_EMIT, st);
  if (chunk > 0) {
    launch_lex_emit<true, false>(ctx->dfa_lds, (unsigned)ctiles, a, st);
    lex_chunk_fixup<<<grid_for(K, 256, 1024), 256, 0, st>>>(a);
    ctx->launches++;
  } else {
    launch_lex_emit<false, false>(ctx->dfa_lds, (unsigned)ctiles, a, st);
  }
  ctx->prof_end(S_LEX_EMIT, st);
  ctx->launches += 7;
  CK(hipGetLastError());
  if (!sync) return ALPACC_ACCEPT;
  CK(hipMemcpyAsync(ctx->h_small, small + 4, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  if (ctx->h_small[2]) {
    ctx->err = "lexer look-back did not complete";
    return ALPACC_ERR_TIMEOUT;
  }
  *num_tokens = ctx->h_small[1];
  return ctx->h_small[0] ? ALPACC_ACCEPT : ALPACC_REJECT;
}

// A batched lexer pass planned on the host: every buffer sized and allocated, the document
// descriptors staged in pinned host memory. enqueue_lex_batch then only queues one copy and the
// kernels, with no allocation and no host synchronisation, so that a batched call can be
// captured into a HIP graph and replayed.
struct LexBatchPlan {
  uint32_t ndocs;
  uint64_t tiles, ctiles, nchunks;
  unsigned lex_blocks;
  LexArgs a;
  LexDoc *d_docs;  // device copy of the descriptors
  LexDoc *h_docs;  // pinned staging of the descriptors
  uint32_t *small;
  uint32_t *doc_res;
  int force_timeout, lex_tab, dfa_lds;
  size_t lex_lds;
};

// Documents (offset, length) of d_in, each lexed on its own; document d's tokens go to slots
// [tok_base_d, tok_base_d + len_d + 1) of kinds/starts/ends (tok_base = exclusive sum of len + 1),
// its (accept, token count) to doc_res[2d], [2d+1] (device). h_docs: pinned, ndocs entries.
int plan_lex_batch(alpacc_ctx *ctx, const uint8_t *d_in, const uint64_t *offsets,
                   const uint64_t *lengths, uint32_t ndocs, uint32_t *kinds, uint32_t *starts,
                   uint32_t *ends, uint32_t *doc_res, LexDoc *h_docs, LexBatchPlan &pl) {
  memset(&pl, 0, sizeof(pl));
  if (!ctx->has_lexer) {
    ctx->err = "bundle has no lexer";
    return ALPACC_ERR_ARG;
  }
  uint64_t tiles = 0, toks = 0, chunks = 0;
  for (uint32_t d = 0; d < ndocs; ++d) {
    if (lengths[d] >= (1ull << 31) - 2) {
      ctx->err = "documents must be shorter than 2^31 - 2 bytes";
      return ALPACC_ERR_ARG;
    }
    const uint64_t nt = (lengths[d] + LEX_TILE - 1) / LEX_TILE;
    const uint64_t nch = (nt + LEX_SCAN_BLOCK - 1) / LEX_SCAN_BLOCK;
    h_docs[d] = LexDoc{offsets[d], lengths[d], toks, (uint32_t)tiles, (uint32_t)nt,
                       (uint32_t)chunks, (uint32_t)nch};
    tiles += nt;
    chunks += nch;
    toks += lengths[d] + 1;
    if (tiles >= (1ull << 31) / LEX_TILE * 2) {
      ctx->err = "batch too large";
      return ALPACC_ERR_ARG;
    }
  }
  pl.ndocs = ndocs;
  pl.h_docs = h_docs;
  pl.doc_res = doc_res;
  pl.d_docs = (LexDoc *)ctx->buf(B_DOCS, (ndocs + 1) * sizeof(LexDoc));
  pl.small = (uint32_t *)ctx->buf(B_SMALL, 64 * sizeof(uint32_t));
  if (!pl.d_docs || !pl.small) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  pl.force_timeout = ctx->force_timeout ? 1 : 0;
  pl.lex_tab = ctx->lex_tab;
  pl.dfa_lds = ctx->dfa_lds ? 1 : 0;
  pl.lex_lds = ctx->lex_lds;
  pl.tiles = tiles;
  if (tiles == 0) return 0;
  const uint64_t ctiles = (tiles + LEX_CT_WAVES - 1) / LEX_CT_WAVES;
  pl.ctiles = ctiles;
  LexArgs &a = pl.a;
  memset(&a, 0, sizeof(a));
  a.lane_excl = (uint16_t *)ctx->buf(B_LANE_EXCL, tiles * WAVE * sizeof(uint16_t));
  a.tile_agg = (uint16_t *)ctx->buf(B_TILE_AGG, tiles * sizeof(uint16_t));
  a.tile_state = (uint8_t *)ctx->buf(B_TILE_STATE, tiles);
  a.st_count = (uint64_t *)ctx->buf(B_ST_COUNT, ctiles * sizeof(uint64_t));
  a.tile_doc = (uint32_t *)ctx->buf(B_TILE_DOC, tiles * sizeof(uint32_t));
  a.tile_rec = (uint4 *)ctx->buf(B_TILE_REC, tiles * sizeof(uint4));
  a.chunks = (uint2 *)ctx->buf(B_CHUNKS, chunks * sizeof(uint2));
  a.tile_excl = (uint16_t *)ctx->buf(B_TILE_EXCL, tiles * sizeof(uint16_t));
  a.block_agg = (uint16_t *)ctx->buf(B_BLOCK_AGG, chunks * sizeof(uint16_t));
  pl.nchunks = chunks;
  if (!a.lane_excl || !a.tile_agg || !a.tile_state || !a.st_count || !a.tile_doc || !a.tile_rec ||
      !a.chunks || !a.tile_excl || !a.block_agg) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  a.in = d_in;
  a.n = 0;
  a.kinds = kinds;
  a.starts = starts;
  a.ends = ends;
  a.byte_class = ctx->d_cls;
  a.endo_step = ctx->d_step;
  a.endo_comp = ctx->d_comp;
  a.endo_apply = ctx->d_apply;
  a.dfa_walk = ctx->d_walk;
  a.e = ctx->e;
  a.C = ctx->C;
  a.m1 = ctx->m1;
  a.init = ctx->init;
  a.ignore = ctx->ignore;
  a.num_tiles = (uint32_t)tiles;
  a.num_ctiles = (uint32_t)ctiles;
  a.tile_counter = pl.small;
  a.result = pl.small + 4;
  a.docs = pl.d_docs;
  a.ndocs = ndocs;
  a.doc_res = doc_res;
  const uint64_t wpb = LEX_BLOCK / WAVE;
  const uint64_t blocks = (tiles + wpb - 1) / wpb;
  pl.lex_blocks = (unsigned)std::min<uint64_t>(blocks, (uint64_t)std::max(1, ctx->cus) * ctx->lex_bpc);
  return 0;
}

void enqueue_lex_batch(alpacc_ctx *ctx, const LexBatchPlan &pl, hipStream_t st) {
  (void)hipMemcpyAsync(pl.d_docs, pl.h_docs, pl.ndocs * sizeof(LexDoc), hipMemcpyHostToDevice, st);
  clear_u32<<<grid_for(2 * (uint64_t)pl.ndocs, 256), 256, 0, st>>>(pl.doc_res, 2 * (uint64_t)pl.ndocs);
  clear_u32<<<1, 64, 0, st>>>(pl.small, 16);
  if (pl.force_timeout) set_u32<<<1, 1, 0, st>>>(pl.small + 6, 1u);
  ctx->launches += 2;
  if (pl.tiles == 0) return;
  const LexArgs &a = pl.a;
  clear_u32<<<grid_for(2 * pl.ctiles, 256), 256, 0, st>>>((uint32_t *)a.st_count, 2 * pl.ctiles);
  lex_tile_docs<<<pl.ndocs, 256, 0, st>>>(a);
  ctx->prof_begin(S_LEX_REDUCE, st);
  launch_lex_reduce<true>(pl.lex_tab, pl.lex_blocks, pl.lex_lds, a, st);
  ctx->prof_end(S_LEX_REDUCE, st);
  ctx->prof_begin(S_LEX_SCAN, st);
  lex_scan_chunks<<<(unsigned)pl.nchunks, LEX_SCAN_BLOCK, 0, st>>>(a);
  lex_scan_doc_chunks<<<pl.ndocs, LEX_SCAN_BLOCK, 0, st>>>(a);
  lex_chunk_states<<<(unsigned)pl.nchunks, LEX_SCAN_BLOCK, 0, st>>>(a);
  ctx->prof_end(S_LEX_SCAN, st);
  ctx->prof_begin(S_LEX_EMIT, st);
  launch_lex_emit<false, true>(pl.dfa_lds != 0, (unsigned)pl.ctiles, a, st);
  ctx->prof_end(S_LEX_EMIT, st);
  ctx->launches += 7;
}

// The batched lexer queued on `st` (descriptors through the context's pinned staging);
// *d_docs_out receives the device descriptors.
int run_lex_batch(alpacc_ctx *ctx, const uint8_t *d_in, const uint64_t *offsets,
                  const uint64_t *lengths, uint32_t ndocs, uint32_t *kinds, uint32_t *starts,
                  uint32_t *ends, uint32_t *doc_res, LexDoc **d_docs_out, hipStream_t st) {
  LexDoc *h_docs = (LexDoc *)ctx->stage((ndocs + 1) * sizeof(LexDoc), st);
  if (!h_docs) {
    ctx->err = "pinned staging allocation failed";
    return ALPACC_ERR_HIP;
  }
  LexBatchPlan pl;
  int rc = plan_lex_batch(ctx, d_in, offsets, lengths, ndocs, kinds, starts, ends, doc_res, h_docs,
                          pl);
  if (rc) return rc;
  enqueue_lex_batch(ctx, pl, st);
  *d_docs_out = pl.d_docs;
  CK(hipGetLastError());
  return 0;
}

// Parser arguments common to the single and batched paths.
void parse_args_init(alpacc_ctx *ctx, PArgs &a) {
  memset(&a, 0, sizeof(a));
  a.q = ctx->q;
  a.k = ctx->k;
  a.nT = ctx->nT;
  a.start_t = ctx->start_t;
  a.end_t = ctx->end_t;
  a.empty_t = ctx->empty_t;
  a.index8 = ctx->d_index8;
  a.index32 = ctx->d_pair_index;
  a.index_len = (uint32_t)ctx->index_len;
  a.hkeys = ctx->d_hkeys;
  a.hvals = ctx->d_hvals;
  a.hbits = ctx->hbits;
  a.hprobe = ctx->hprobe;
  a.einfo = ctx->d_einfo;
  a.E = ctx->E;
  a.stackw = ctx->d_stackw;
  a.prodw = ctx->d_prodw;
  a.ecan = ctx->d_ecan;
  a.nstack = (uint32_t)ctx->nstackw;
  a.nprod = (uint32_t)ctx->nprodw;
}

// keys, tile totals and the tile scan (ntiles_max: tiles to cover, the device count in batch)
template <bool BATCH>
int parse_phase1(alpacc_ctx *ctx, PArgs &a, uint64_t ntiles_max, hipStream_t st,
                 SegTup *btot = nullptr) {
  ctx->prof_begin(S_PARSE_KEYS, st);
  const uint64_t nsuper = (ntiles_max + PKK_TILES - 1) / PKK_TILES;
  const unsigned kgrid =
      (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(nsuper, (uint64_t)std::max(1, ctx->cus) * ctx->keys_bpc));
  const bool small_w = ctx->q + ctx->k <= 4;
  if (ctx->d_index8 && small_w)
    parse_keys2<true, BATCH, 4><<<kgrid, PKK_THREADS, ctx->index_len, st>>>(a);
  else if (ctx->d_index8)
    parse_keys2<true, BATCH, PK_MAXW><<<kgrid, PKK_THREADS, ctx->index_len, st>>>(a);
  else if (small_w)
    parse_keys2<false, BATCH, 4><<<kgrid, PKK_THREADS, 0, st>>>(a);
  else
    parse_keys2<false, BATCH, PK_MAXW><<<kgrid, PKK_THREADS, 0, st>>>(a);
  const uint64_t nb = std::max<uint64_t>(1, (ntiles_max + PSB - 1) / PSB);
  if (!btot) btot = (SegTup *)ctx->buf(B_BTOT, (nb + 1) * sizeof(SegTup));
  if (!btot) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  pscan_reduce<BATCH><<<(unsigned)nb, PSB, 0, st>>>(a, btot);
  pscan_top<<<1, PSB, 0, st>>>(a, btot, nb);
  pscan_down<BATCH><<<(unsigned)nb, PSB, 0, st>>>(a, btot);
  ctx->prof_end(S_PARSE_KEYS, st);
  ctx->launches += 4;
  CK(hipGetLastError());
  return 0;
}

// element scratch for B brackets and N productions
struct ElemBufs {
  int32_t *depth, *dl1, *sx, *sl1;
  uint16_t *br;
};

int elem_bufs(alpacc_ctx *ctx, uint64_t B, uint64_t N, ElemBufs &e, hipStream_t st) {
  const uint64_t nbc = (B + WAVE - 1) / WAVE, npc = (N + WAVE - 1) / WAVE;
  e.depth = (int32_t *)ctx->buf(B_DEPTH, (B + 16) * sizeof(int32_t));
  e.br = (uint16_t *)ctx->buf(B_BR, (B + 16) * sizeof(uint16_t));
  e.dl1 = (int32_t *)ctx->buf(B_DL1, (nbc + 16) * sizeof(int32_t));
  e.sx = (int32_t *)ctx->buf(B_SX, (N + 16) * sizeof(int32_t));
  e.sl1 = (int32_t *)ctx->buf(B_SL1, (npc + 16) * sizeof(int32_t));
  if (!e.depth || !e.br || !e.dl1 || !e.sx || !e.sl1) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  elem_init<<<grid_for(std::max(nbc, npc) + 16, 256, 4096), 256, 0, st>>>(e.dl1, nbc + 16, e.sl1,
                                                                          npc + 16);
  ctx->launches++;
  return 0;
}

// bracket matching over B brackets, then parents of N nodes (N = 0: no parents); the two min
// hierarchies live in separate scratch arenas
template <bool BATCH>
int run_search(alpacc_ctx *ctx, const PArgs &a, const ElemBufs &e, uint64_t B, uint64_t N,
               uint32_t *parent, uint32_t *flags, hipStream_t st) {
  if (B == 0 && N == 0) return 0;
  Levels lvb, lvs;
  lvb.nlev = lvs.nlev = 0;
  int rc = 0;
  if (B > 0) rc = plan_levels(ctx, e.depth, e.dl1, B, lvb, B_LEVELS);
  if (!rc && N > 0) rc = plan_levels(ctx, e.sx, e.sl1, N, lvs, B_LEVELS_S);
  if (!rc) rc = fill_levels(ctx, lvb, lvs, st);
  if (rc) return rc;
  if (B > 0) {
    ctx->prof_begin(S_BRACKET_MATCH, st);
    launch_bracket_match<BATCH>(ctx->search_bs, ctx->search_cap, e.br, B, lvb, flags, a, st);
    ctx->prof_end(S_BRACKET_MATCH, st);
    ctx->launches++;
  }
  if (N > 0) {
    ctx->prof_begin(S_PARSE_PARENTS, st);
    launch_parse_parents<BATCH>(ctx->search_bs, ctx->search_cap, N, lvs, parent, a, st);
    ctx->prof_end(S_PARSE_PARENTS, st);
    ctx->launches++;
  }
  CK(hipGetLastError());
  return 0;
}

// The parser on a token stream of T tokens, with one host synchronisation after the tile scan,
// which sizes the element scratch (so that parse_expand2 runs one workgroup per actual tile) and
// decides the missing-pair and depth checks.
// lexed = true: the tokens come straight from run_lex(sync = false) in this call (kinds as bytes
// in kinds8); T is then an upper bound (the input length), the key and scan kernels read the
// token count from the lexer's device words, and the synchronisation also reads the lexer's
// verdict.
template <typename S>
int run_parse_tokens(alpacc_ctx *ctx, const uint32_t *kinds, const S *starts, const S *ends,
                     uint64_t T, uint32_t *parent, uint32_t *node, S *span, uint64_t cap,
                     uint64_t *num_nodes, hipStream_t st, bool productions_only = false,
                     bool lexed = false, const uint8_t *kinds8 = nullptr) {
  if (!ctx->has_parser) {
    ctx->err = "bundle has no parser";
    return ALPACC_ERR_NO_PARSER;
  }
  *num_nodes = 0;
  uint64_t P = T + 2;
  const uint64_t elem_max = (uint64_t)std::max(ctx->max_nb, ctx->max_np);
  if (!lexed && P * elem_max >= (1ull << 31)) {
    ctx->err = "token stream too long for 32-bit element counts";
    return ALPACC_ERR_ARG;
  }
  uint64_t ntiles = (P + PK_TILE - 1) / PK_TILE;
  uint32_t *small = (uint32_t *)ctx->buf(B_SMALL, 64 * sizeof(uint32_t));
  uint16_t *ent = (uint16_t *)ctx->buf(B_ENT, (ntiles * PK_TILE + 8) * sizeof(uint16_t));
  PTup *tiles = (PTup *)ctx->buf(B_PTILES, (ntiles + 1) * sizeof(PTup));
  PTup *pre = (PTup *)ctx->buf(B_PPRE, (ntiles + 1) * sizeof(PTup));
  if (!small || !ent || !tiles || !pre) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  uint32_t *flags = small + 8;
  PTup *total = reinterpret_cast<PTup *>(small + 16);
  PArgs a;
  parse_args_init(ctx, a);
  a.kinds = kinds;
  a.kinds8 = kinds8;
  a.T = T;
  a.P = P;
  a.T_dev = lexed ? small + 5 : nullptr;  // the lexer's result words: verdict, count, time-out
  a.ent = ent;
  a.tiles = tiles;
  a.pre = pre;
  a.total = total;
  a.flags = flags;
  a.ntiles = ntiles;
  ctx->prof_begin(S_PARSE, st);
  clear_u32<<<1, 64, 0, st>>>(flags, 4);
  ctx->launches++;
  int rc = parse_phase1<false>(ctx, a, ntiles, st);
  if (rc) return rc;
  // with lexed: the lexer's words (small + 4 .. + 7) come along
  const int sh = lexed ? 4 : 0;
  CK(hipMemcpyAsync(ctx->h_small, small + 8 - sh, (16 + sh) * sizeof(uint32_t),
                    hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  if (lexed) {
    if (ctx->h_small[2]) {
      ctx->err = "lexer look-back did not complete";
      return ALPACC_ERR_TIMEOUT;
    }
    if (!ctx->h_small[0]) return ALPACC_REJECT;  // the lexer rejected the input
    T = ctx->h_small[1];
    P = T + 2;
    if (P * elem_max >= (1ull << 31)) {
      ctx->err = "token stream too long for 32-bit element counts";
      return ALPACC_ERR_ARG;
    }
    ntiles = (P + PK_TILE - 1) / PK_TILE;
    a.T = T;
    a.P = P;
    a.ntiles = ntiles;
    a.T_dev = nullptr;  // from here on the counts are known on the host
  }
  const uint32_t *hs = ctx->h_small + sh;
  if (hs[0]) return ALPACC_REJECT;  // a missing pair (to_keys, parser.fut:235-240)
  PTup tot;
  memcpy(&tot, hs + 8, sizeof(PTup));
  // depths (parser.fut:153-166): a prefix below zero, or a non-zero end
  // (no bracket at all: mind is INT32_MAX and dd 0; cancelled brackets still count here)
  if (tot.mind < 0 || tot.dd != 0) return ALPACC_REJECT;
  const uint64_t B = (uint32_t)tot.nb, N = (uint32_t)tot.np;
  a.subst = (!productions_only && (uint64_t)(uint32_t)tot.nt == T) ? 1 : 0;
  ctx->last_counts[0] = T;
  ctx->last_counts[1] = P;
  ctx->last_counts[2] = B;
  ctx->last_counts[3] = N;
  *num_nodes = N;
  const int write_nodes = N <= cap ? 1 : 0;
  ElemBufs e;
  rc = elem_bufs(ctx, B, N, e, st);
  if (rc) return rc;
  const size_t xlds = px_lds_bytes(ctx->nstackw, ctx->nprodw, ctx->E);
  ctx->prof_begin(S_PARSE_EXPAND, st);
  parse_expand2<S, false><<<(unsigned)ntiles, PK_THREADS, xlds, st>>>(
      a, starts, ends, e.depth, e.br, e.dl1, e.sx, e.sl1, node, span, write_nodes);
  ctx->prof_end(S_PARSE_EXPAND, st);
  ctx->launches++;
  CK(hipGetLastError());
  rc = run_search<false>(ctx, a, e, B, write_nodes ? N : 0, parent, flags, st);
  if (rc) return rc;
  ctx->prof_end(S_PARSE, st);
  if (!write_nodes || N == 0) {
    CK(hipMemcpyAsync(ctx->h_small, flags, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    CK(hipStreamSynchronize(st));
    if (ctx->h_small[0]) return ALPACC_REJECT;
    if (!write_nodes) {
      ctx->err = "output capacity too small";
      return ALPACC_ERR_CAPACITY;
    }
    return ALPACC_ACCEPT;
  }
  CK(hipMemcpyAsync(ctx->h_small, flags, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  if (ctx->h_small[0]) return ALPACC_REJECT;  // brackets_matches (parser.fut:174-184)
  return ALPACC_ACCEPT;
}

// Final per-document results of the batched parser.
__global__ void pb_finish(PArgs a, alpacc_doc_result *out) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= a.ndocs) return;
  const PDoc pd = a.pdocs[d];
  alpacc_doc_result r;
  r.status = (pd.status && (a.doc_flags[d] & 2u) == 0) ? ALPACC_ACCEPT : ALPACC_REJECT;
  r.reserved = 0;
  r.offset = pd.node_off;
  r.count = r.status == ALPACC_ACCEPT ? pd.N : 0;
  out[d] = r;
}

__global__ void lb_finish(const LexDoc *docs, const uint32_t *doc_res, uint32_t ndocs,
                          alpacc_doc_result *out) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndocs) return;
  alpacc_doc_result r;
  r.status = doc_res[2 * d] ? ALPACC_ACCEPT : ALPACC_REJECT;
  r.reserved = 0;
  r.offset = docs[d].tok_base;
  r.count = r.status == ALPACC_ACCEPT ? doc_res[2 * d + 1] : 0;
  out[d] = r;
}

// A batched lex + parse planned on the host (as LexBatchPlan): the bracket and production scratch
// is sized before the launch, for the productions from the output capacity and for the brackets
// from an estimate, so that the whole call is one queue with no host synchronisation before its
// end. A batch whose totals exceed those capacities writes no elements (batch_elems_fit); the
// host then plans it again with the sizes the first pass reported.
struct ParseBatchPlan {
  LexBatchPlan lx;
  PArgs a;
  uint64_t tiles_max;
  uint64_t tiles_est;      // parse_expand2's grid
  unsigned stride_blocks;  // the striding pass's grid (0: none needed)
  SegTup *btot;
  ElemBufs e;
  uint64_t nbc, npc;  // 64-element chunks of the bracket / production capacities
  Levels lvb, lvs;
  uint32_t *parent, *node, *span;
  const uint32_t *starts, *ends;
  alpacc_doc_result *d_results, *h_results;
  uint32_t *h_words;  // pinned copy of the result words small[0 .. 32)
  size_t xlds;
  unsigned search_bs, search_cap;
};

int plan_parse_batch(alpacc_ctx *ctx, const uint8_t *d_in, const uint64_t *offsets,
                     const uint64_t *lengths, uint32_t ndocs, uint32_t *parent, uint32_t *node,
                     uint32_t *span, uint64_t cap, uint64_t bcap, ParseBatchPlan &pl,
                     hipStream_t st) {
  memset(&pl, 0, sizeof(pl));
  uint64_t nbytes = 0, tiles_max = 0;
  for (uint32_t d = 0; d < ndocs; ++d) {
    nbytes += lengths[d];
    tiles_max += (lengths[d] + 3 + PK_TILE - 1) / PK_TILE;
  }
  const uint64_t tslots = nbytes + ndocs + 1;
  if (tslots >= (1ull << 32) - 1 || tiles_max >= (1ull << 32) / PK_TILE) {
    ctx->err = "batch too large: token slots and key positions must stay below 2^32";
    return ALPACC_ERR_ARG;
  }
  // element offsets are 32-bit. pscan_top flags a total of 2^31 or more on the device, which is
  // exact while every block of PSB tiles stays below 2^31 elements; with pair entries of 2048
  // elements or more that is not certain, and the host bound (T + 2 <= length + 2) decides.
  const uint64_t elem_max = (uint64_t)std::max(ctx->max_nb, ctx->max_np);
  if (elem_max >= 2048 && (tslots + 2ull * ndocs) * elem_max >= (1ull << 31)) {
    ctx->err = "batch too large for 32-bit element counts";
    return ALPACC_ERR_ARG;
  }
  const uint64_t ncap = std::min<uint64_t>(cap, 1ull << 31);
  const uint64_t nbc = (bcap + WAVE - 1) / WAVE, npc = (ncap + WAVE - 1) / WAVE;
  uint8_t *kinds = (uint8_t *)ctx->buf(B_TOK_KIND, tslots + 16);  // kinds as bytes
  uint32_t *starts = (uint32_t *)ctx->buf(B_TOK_START, tslots * 4);
  uint32_t *ends = (uint32_t *)ctx->buf(B_TOK_END, tslots * 4);
  uint32_t *doc_res = (uint32_t *)ctx->buf(B_DOC_RES, (2ull * ndocs + 2) * 4);
  PDoc *pdocs = (PDoc *)ctx->buf(B_PDOCS, (ndocs + 1) * sizeof(PDoc));
  uint32_t *doc_flags = (uint32_t *)ctx->buf(B_DOC_FLAGS, (ndocs + 1) * 4);
  uint32_t *ptile_doc = (uint32_t *)ctx->buf(B_PTILE_DOC, (tiles_max + 1) * 4);
  uint16_t *ent = (uint16_t *)ctx->buf(B_ENT, (tiles_max * PK_TILE + 8) * sizeof(uint16_t));
  PTup *tiles = (PTup *)ctx->buf(B_PTILES, (tiles_max + 1) * sizeof(PTup));
  PTup *pre = (PTup *)ctx->buf(B_PPRE, (tiles_max + 1) * sizeof(PTup));
  pl.d_results = (alpacc_doc_result *)ctx->buf(B_RESULTS, (ndocs + 1) * sizeof(alpacc_doc_result));
  const uint64_t nbs = std::max<uint64_t>(1, (tiles_max + PSB - 1) / PSB);
  pl.btot = (SegTup *)ctx->buf(B_BTOT, (nbs + 1) * sizeof(SegTup));
  pl.e.depth = (int32_t *)ctx->buf(B_DEPTH, (bcap + 16) * sizeof(int32_t));
  pl.e.br = (uint16_t *)ctx->buf(B_BR, (bcap + 16) * sizeof(uint16_t));
  pl.e.dl1 = (int32_t *)ctx->buf(B_DL1, (nbc + 16) * sizeof(int32_t));
  pl.e.sx = (int32_t *)ctx->buf(B_SX, (ncap + 16) * sizeof(int32_t));
  pl.e.sl1 = (int32_t *)ctx->buf(B_SL1, (npc + 16) * sizeof(int32_t));
  uint32_t *chunk_doc = (uint32_t *)ctx->buf(B_CHUNK_DOC, (npc + 2) * 4);
  uint32_t *chunk_base = (uint32_t *)ctx->buf(B_CHUNK_BASE, (npc + 2) * 4);
  const size_t docs_bytes = ((ndocs + 1) * sizeof(LexDoc) + 255) & ~(size_t)255;
  const size_t res_bytes = ((ndocs + 1) * sizeof(alpacc_doc_result) + 255) & ~(size_t)255;
  uint8_t *stg = ctx->stage(docs_bytes + res_bytes + 256, st);
  if (!kinds || !starts || !ends || !doc_res || !pdocs || !doc_flags || !ptile_doc || !ent ||
      !tiles || !pre || !pl.d_results || !pl.btot || !pl.e.depth || !pl.e.br || !pl.e.dl1 ||
      !pl.e.sx || !pl.e.sl1 || !chunk_doc || !chunk_base || !stg) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  pl.h_results = reinterpret_cast<alpacc_doc_result *>(stg + docs_bytes);
  pl.h_words = reinterpret_cast<uint32_t *>(stg + docs_bytes + res_bytes);
  int rc = plan_lex_batch(ctx, d_in, offsets, lengths, ndocs, nullptr, starts, ends, doc_res,
                          reinterpret_cast<LexDoc *>(stg), pl.lx);
  if (rc) return rc;
  pl.lx.a.kinds8 = kinds;
  uint32_t *small = pl.lx.small;
  PArgs &a = pl.a;
  parse_args_init(ctx, a);
  a.kinds = nullptr;
  a.kinds8 = kinds;
  a.ent = ent;
  a.tiles = tiles;
  a.pre = pre;
  a.total = reinterpret_cast<PTup *>(small + 16);
  a.flags = small + 8;
  a.pdocs = pdocs;
  a.ptile_doc = ptile_doc;
  a.ntiles_dev = reinterpret_cast<uint64_t *>(small + 24);
  a.ndocs = ndocs;
  a.doc_flags = doc_flags;
  a.chunk_doc = chunk_doc;
  a.chunk_base = chunk_base;
  a.bcap = bcap;
  a.ncap = ncap;
  pl.tiles_max = tiles_max;
  // parse_expand2's grid: the tiles of an estimate of the key positions (p_per_byte of the input,
  // grown to what calls needed; the bound from the input bytes is several times the count), and
  // a striding pass over any tiles past it
  pl.tiles_est = std::min<uint64_t>(
      tiles_max, (uint64_t)(ctx->p_per_byte * (double)nbytes) / PK_TILE + ndocs + 1);
  pl.stride_blocks =
      pl.tiles_est < tiles_max ? (unsigned)std::max(1, ctx->cus) * 4 : 0u;
  pl.nbc = nbc;
  pl.npc = npc;
  pl.parent = parent;
  pl.node = node;
  pl.span = span;
  pl.starts = starts;
  pl.ends = ends;
  pl.lvb.nlev = pl.lvs.nlev = 0;
  if (bcap > 0) rc = plan_levels(ctx, pl.e.depth, pl.e.dl1, bcap, pl.lvb, B_LEVELS);
  if (!rc && ncap > 0) rc = plan_levels(ctx, pl.e.sx, pl.e.sl1, ncap, pl.lvs, B_LEVELS_S);
  if (rc) return rc;
  pl.xlds = px_lds_bytes(ctx->nstackw, ctx->nprodw, ctx->E);
  pl.search_bs = ctx->search_bs;
  pl.search_cap = ctx->search_cap;
  return 0;
}

// Queues the planned call on `st`: copies and kernels only (no allocation, no synchronisation).
void enqueue_parse_batch(alpacc_ctx *ctx, const ParseBatchPlan &pl, hipStream_t st) {
  enqueue_lex_batch(ctx, pl.lx, st);
  PArgs a = pl.a;
  const uint32_t ndocs = pl.lx.ndocs;
  ctx->prof_begin(S_PARSE, st);
  pb_setup<<<1, PS_THREADS, 0, st>>>(a, pl.lx.d_docs, pl.lx.doc_res,
                                      const_cast<uint64_t *>(a.ntiles_dev));
  pb_tile_docs<<<ndocs, 256, 0, st>>>(a);
  ctx->launches += 2;
  (void)parse_phase1<true>(ctx, a, pl.tiles_max, st, pl.btot);
  pb_docs<<<(ndocs + 255) / 256, 256, 0, st>>>(a);
  elem_init<<<grid_for(std::max(pl.nbc, pl.npc) + 16, 256, 4096), 256, 0, st>>>(
      pl.e.dl1, pl.nbc + 16, pl.e.sl1, pl.npc + 16);
  ctx->launches += 2;
  if (pl.tiles_max > 0) {
    ctx->prof_begin(S_PARSE_EXPAND, st);
    parse_expand2<uint32_t, true><<<(unsigned)std::max<uint64_t>(1, pl.tiles_est), PK_THREADS,
                                     pl.xlds, st>>>(
        a, pl.starts, pl.ends, pl.e.depth, pl.e.br, pl.e.dl1, pl.e.sx, pl.e.sl1, pl.node, pl.span,
        1);
    ctx->launches++;
    if (pl.stride_blocks) {
      PArgs as = a;
      as.expand_tile0 = pl.tiles_est;
      parse_expand2<uint32_t, true, true><<<pl.stride_blocks, PK_THREADS, pl.xlds, st>>>(
          as, pl.starts, pl.ends, pl.e.depth, pl.e.br, pl.e.dl1, pl.e.sx, pl.e.sl1, pl.node,
          pl.span, 1);
      ctx->launches++;
    }
    ctx->prof_end(S_PARSE_EXPAND, st);
  }
  (void)fill_levels(ctx, pl.lvb, pl.lvs, st);
  if (a.bcap > 0) {
    ctx->prof_begin(S_BRACKET_MATCH, st);
    launch_bracket_match<true>(pl.search_bs, pl.search_cap, pl.e.br, a.bcap, pl.lvb, a.flags, a,
                               st);
    ctx->prof_end(S_BRACKET_MATCH, st);
    ctx->launches++;
  }
  if (a.ncap > 0) {
    ctx->prof_begin(S_PARSE_PARENTS, st);
    launch_parse_parents<true>(pl.search_bs, pl.search_cap, a.ncap, pl.lvs, pl.parent, a, st);
    ctx->prof_end(S_PARSE_PARENTS, st);
    ctx->launches++;
  }
  pb_finish<<<(ndocs + 255) / 256, 256, 0, st>>>(a, pl.d_results);
  ctx->launches++;
  ctx->prof_end(S_PARSE, st);
  (void)hipMemcpyAsync(pl.h_results, pl.d_results, ndocs * sizeof(alpacc_doc_result),
                       hipMemcpyDeviceToHost, st);
  (void)hipMemcpyAsync(pl.h_words, pl.lx.small, 32 * sizeof(uint32_t), hipMemcpyDeviceToHost, st);
}

// Runs a planned call on `st`: queued directly, or through a HIP graph. graph_mode 1 captures
// the queue when a call's plan (every size, pointer and flag) equals the previous call's, and
// replays that graph while the plans keep repeating; graph_mode 2 captures every new plan.
// Profiling (HIP events per stage) always queues directly.
int launch_parse_batch(alpacc_ctx *ctx, const ParseBatchPlan &pl, hipStream_t st) {
  if (ctx->graph_mode == 0 || ctx->prof) {
    enqueue_parse_batch(ctx, pl, st);
    CK(hipGetLastError());
    return 0;
  }
  const uint8_t *pb = reinterpret_cast<const uint8_t *>(&pl);
  std::vector<uint8_t> plan(pb, pb + sizeof(pl));
  if (ctx->gexec && plan == ctx->gplan) {
    CK(hipGraphLaunch(ctx->gexec, st));
    ctx->graph_launches++;
    return 0;
  }
  if (ctx->graph_mode == 1 && plan != ctx->last_plan) {
    ctx->last_plan.swap(plan);
    enqueue_parse_batch(ctx, pl, st);
    CK(hipGetLastError());
    return 0;
  }
  ctx->drop_graph();
  if (!ctx->cap_stream) CK(hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking));
  CK(hipStreamBeginCapture(ctx->cap_stream, hipStreamCaptureModeThreadLocal));
  enqueue_parse_batch(ctx, pl, ctx->cap_stream);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->cap_stream, &graph);
  const hipError_t el = hipGetLastError();
  if (ec != hipSuccess || el != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    ctx->err = std::string("graph capture failed: ") + hipGetErrorString(ec != hipSuccess ? ec : el);
    return ALPACC_ERR_HIP;
  }
  const hipError_t ei = hipGraphInstantiate(&ctx->gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    ctx->gexec = nullptr;
    ctx->err = std::string("graph instantiation failed: ") + hipGetErrorString(ei);
    return ALPACC_ERR_HIP;
  }
  ctx->gplan.assign(pb, pb + sizeof(pl));
  ctx->last_plan = ctx->gplan;
  ctx->graph_captures++;
  CK(hipGraphLaunch(ctx->gexec, st));
  ctx->graph_launches++;
  return 0;
}

// Batched lex + parse of many documents: one lexer pass and one parser pass over all of them,
// queued as a whole (or replayed as a graph) and followed by one host synchronisation. A batch
// with more brackets than the scratch was sized for runs again with the exact size.
int run_parse_batch(alpacc_ctx *ctx, const uint8_t *d_in, const uint64_t *offsets,
                    const uint64_t *lengths, uint32_t ndocs, uint32_t *parent, uint32_t *node,
                    uint32_t *span, uint64_t cap, uint64_t *total_nodes,
                    alpacc_doc_result *results, hipStream_t st) {
  if (!ctx->has_parser) {
    ctx->err = "bundle has no parser";
    return ALPACC_ERR_NO_PARSER;
  }
  *total_nodes = 0;
  uint64_t nbytes = 0;
  for (uint32_t d = 0; d < ndocs; ++d) nbytes += lengths[d];
  // brackets: at most max_nb per key position (T + 2 <= length + 2 per document)
  const uint64_t b_bound = (nbytes + 2ull * ndocs) * std::max<uint32_t>(ctx->max_nb, 1);
  uint64_t bcap = std::min<uint64_t>(
      b_bound, std::max<uint64_t>(4096, (uint64_t)(ctx->b_per_byte * (double)nbytes)));
  for (int attempt = 0; attempt < 2; ++attempt) {
    ParseBatchPlan pl;
    int rc = plan_parse_batch(ctx, d_in, offsets, lengths, ndocs, parent, node, span, cap, bcap,
                              pl, st);
    if (rc) return rc;
    rc = launch_parse_batch(ctx, pl, st);
    if (rc) return rc;
    CK(hipStreamSynchronize(st));
    const uint32_t *h = pl.h_words;
    // words 4..7: the lexer's results (time-out flag at 6); 8..11: parser flags (overflow at 9);
    // 16..21: grand total; 24..25: position tiles
    if (h[6]) {
      ctx->err = "lexer look-back did not complete";
      return ALPACC_ERR_TIMEOUT;
    }
    PTup tot;
    memcpy(&tot, h + 16, sizeof(PTup));
    uint64_t ntiles;
    memcpy(&ntiles, h + 24, sizeof(uint64_t));
    if (h[9] || tot.nb < 0 || tot.np < 0) {
      ctx->err = "batch too large: more than 2^31 brackets or productions";
      return ALPACC_ERR_ARG;
    }
    const uint64_t B = (uint32_t)tot.nb, N = (uint32_t)tot.np;
    *total_nodes = N;
    ctx->last_counts[0] = 0;
    ctx->last_counts[1] = ntiles * PK_TILE;
    ctx->last_counts[2] = B;
    ctx->last_counts[3] = N;
    if (N > cap) {
      ctx->err = "output capacity too small";
      return ALPACC_ERR_CAPACITY;
    }
    if (ntiles > pl.tiles_est && nbytes)  // the striding pass ran: widen the estimate
      ctx->p_per_byte = std::min(
          4.0, std::max(ctx->p_per_byte, 1.25 * (double)(ntiles * PK_TILE) / (double)nbytes));
    if (B > bcap) {  // the first pass wrote nothing: size the brackets for this batch
      bcap = B + B / 8 + 4096;
      if (nbytes) ctx->b_per_byte = std::max(ctx->b_per_byte, (double)bcap / (double)nbytes);
      continue;
    }
    memcpy(results, pl.h_results, ndocs * sizeof(alpacc_doc_result));
    return ALPACC_ACCEPT;
  }
  ctx->err = "bracket scratch sizing did not converge";
  return ALPACC_ERR_HIP;
}

// Where the lexer's tables sit: lex_reduce keeps the (endo x class) table in LDS when it fits in
// 160 KiB (and the composition table too when both fit in 48 KiB), else reads both from L2;
// lex_emit keeps the DFA table in LDS up to LEX_MAX_DFA entries, else reads it from L2. With the
// tuning choice lexer_tables_in_lds = 0 every table stays in L2.
int place_lexer_tables(alpacc_ctx *ctx, std::string &err) {
  size_t lds = (size_t)ctx->e * ctx->C * 2 + 256;
  lds = (lds + 15) & ~(size_t)15;
  const size_t comp_bytes = (size_t)ctx->e * ctx->e * 2;
  if (!ctx->lex_tables_lds) {
    ctx->lex_tab = 0;
    lds = 256;
  } else if (lds + comp_bytes <= 48 * 1024) {
    ctx->lex_tab = 2;
    lds += comp_bytes;
  } else if (lds <= 160 * 1024) {
    ctx->lex_tab = 1;
  } else {
    ctx->lex_tab = 0;
    lds = 256;
  }
  ctx->dfa_lds = ctx->lex_tables_lds && (size_t)ctx->m1 * ctx->C <= (size_t)LEX_MAX_DFA;
  if (lds > 64 * 1024) {
    const int want = (int)lds;
    if (hipFuncSetAttribute((const void *)lex_reduce<1, false>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, want) != hipSuccess ||
        hipFuncSetAttribute((const void *)lex_reduce<1, true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, want) != hipSuccess) {
      err = "lexer tables do not fit in LDS";
      return -1;
    }
  }
  ctx->lex_lds = lds;
  return 0;
}

// Per pair entry totals and per element relative values (see the parser kernels) from the
// pair table, stacks and productions of the generated bundle (Encode.hs:336-375 layout restated
// in alpacc_amd/gen/tables.py).
bool build_parser_tables(alpacc_ctx *ctx, const alpacc_tables *t, std::string &err) {
  const uint32_t E = t->num_pair_entries;
  if (E == 0 || E >= 65535 || t->num_productions >= 65535 || t->num_terminals >= 65535 ||
      t->q + t->k > (uint32_t)PK_MAXW || t->q + t->k == 0) {
    err = "parser tables out of range";
    return false;
  }
  std::vector<EntInfo> info(E);
  std::vector<uint32_t> stackw;
  std::vector<uint2> prodw;
  std::vector<uint32_t> ecan(E);
  for (uint32_t e = 0; e < E; ++e) {
    const uint32_t *e4 = t->pair_entry + 4ull * e;
    if (e4[1] < e4[0] || e4[3] < e4[2] || e4[1] > t->stacks_len || e4[3] > t->productions_len) {
      err = "pair entry out of range";
      return false;
    }
    // the entry's leading right and trailing left brackets (at most 255 of each counted): the
    // ones that can cancel against the neighbouring positions' brackets
    uint32_t lr = 0, tl = 0;
    while (lr < 255 && e4[0] + lr < e4[1] && !(t->stacks[e4[0] + lr] & 0x8000u)) ++lr;
    while (tl < 255 && e4[1] - tl > e4[0] && (t->stacks[e4[1] - 1 - tl] & 0x8000u)) ++tl;
    // bit 16: brackets between the counted leading rights and trailing lefts (they never cancel)
    ecan[e] = tl | (lr << 8) | ((e4[1] - e4[0] > lr + tl) ? 0x10000u : 0u);
    EntInfo &f = info[e];
    f.nb = (uint16_t)(e4[1] - e4[0]);
    f.np = (uint16_t)(e4[3] - e4[2]);
    f.sst = (uint16_t)stackw.size();
    f.pst = (uint16_t)prodw.size();
    int32_t r = 0, mind = INT16_MAX;
    for (uint32_t s = e4[0]; s < e4[1]; ++s) {
      const uint16_t b = t->stacks[s];
      const bool left = (b & 0x8000u) != 0;
      r += left ? 1 : -1;
      mind = std::min(mind, r);
      const int32_t rel = r - (left ? 1 : 0);  // depths: scan - [is_left]
      stackw.push_back((uint32_t)b | ((uint32_t)(uint16_t)(int16_t)rel << 16));
    }
    f.dd = (int16_t)r;
    f.mind = (int16_t)mind;
    int32_t ds = 0, nt = 0;
    for (uint32_t s = e4[2]; s < e4[3]; ++s) {
      const uint16_t pr = t->productions[s];
      if (pr >= t->num_productions) {
        err = "production id out of range";
        return false;
      }
      const int32_t term = t->prod_to_terminal[pr];
      uint2 w;
      w.x = (uint32_t)pr | ((uint32_t)(term + 1) << 16);
      w.y = (uint32_t)(uint16_t)(int16_t)ds | ((uint32_t)nt << 16);
      prodw.push_back(w);
      ds += t->arities[pr] - 1;
      nt += term >= 0 ? 1 : 0;
    }
    f.ds = (int16_t)ds;
    f.nt = (uint16_t)nt;
    ctx->max_nb = std::max<uint32_t>(ctx->max_nb, f.nb);
    ctx->max_np = std::max<uint32_t>(ctx->max_np, f.np);
    if (stackw.size() >= 65535 || prodw.size() >= 65535 || std::abs(r) > 30000 ||
        std::abs(ds) > 30000) {
      err = "pair entries too long";
      return false;
    }
  }
  // number of key windows, saturating at 2^64 - 1
  uint64_t want_len = 1;
  for (uint32_t j = 0; j < t->q + t->k; ++j)
    want_len = want_len > ~0ull / std::max<uint64_t>(t->num_terminals, 1)
                   ? ~0ull : want_len * t->num_terminals;
  const bool hashed = t->pair_hash_len > 0;
  if (hashed) {
    const uint64_t H = t->pair_hash_len;
    if ((H & (H - 1)) != 0 || H < 2 || H > (1ull << 32) || !t->pair_hash_keys ||
        !t->pair_hash_vals || want_len == ~0ull) {
      err = "hashed pair table malformed";
      return false;
    }
    uint32_t bits = 0;
    while ((1ull << bits) < H) ++bits;
    // every stored key must be reachable from its home slot; the longest probe bounds the search
    uint64_t probe_max = 1, used = 0;
    for (uint64_t i = 0; i < H; ++i) {
      const uint64_t key = t->pair_hash_keys[i];
      if (key == ~0ull) continue;
      ++used;
      if (key >= want_len || t->pair_hash_vals[i] >= E) {
        err = "hashed pair table entry out of range";
        return false;
      }
      uint64_t h = (key * 0x9E3779B97F4A7C15ull) >> (64 - bits), d = 1;
      while (h != i) {
        if (t->pair_hash_keys[h] == ~0ull) {
          err = "hashed pair table: a key is not reachable from its home slot";
          return false;
        }
        h = (h + 1) & (H - 1);
        ++d;
      }
      probe_max = std::max(probe_max, d);
    }
    if (used == H) {
      err = "hashed pair table has no free slot";
      return false;
    }
    ctx->hbits = bits;
    ctx->hprobe = (uint32_t)std::min<uint64_t>(probe_max + 1, H);
  } else {
    for (uint64_t i = 0; i < t->pair_index_len; ++i) {
      if (t->pair_index[i] != 0xFFFFFFFFu && t->pair_index[i] >= E) {
        err = "pair index entry out of range";
        return false;
      }
    }
    if (want_len != t->pair_index_len || !t->pair_index) {
      err = "pair index length is not num_terminals^(q+k)";
      return false;
    }
  }
  if (stackw.empty()) stackw.push_back(0);
  if (prodw.empty()) prodw.push_back(make_uint2(0, 0));
  if (px_lds_bytes(stackw.size(), prodw.size(), E) > PX_TABLE_LDS_MAX) {
    err = "pair entry tables too large for LDS";
    return false;
  }
  ctx->index_len = t->pair_index_len;
  if (hashed) {
    ctx->d_hkeys = upload(t->pair_hash_keys, t->pair_hash_len, err);
    ctx->d_hvals = upload(t->pair_hash_vals, t->pair_hash_len, err);
    ctx->table_bytes += t->pair_hash_len * 12;
  } else {
    ctx->d_pair_index = upload(t->pair_index, t->pair_index_len, err);
  }
  if (!hashed && E < 255 && t->pair_index_len <= (uint64_t)PK_IDX8_MAX) {
    std::vector<uint8_t> idx8(t->pair_index_len);
    for (uint64_t i = 0; i < t->pair_index_len; ++i)
      idx8[i] = t->pair_index[i] == 0xFFFFFFFFu ? (uint8_t)0xFF : (uint8_t)t->pair_index[i];
    ctx->d_index8 = upload(idx8.data(), idx8.size(), err);
  }
  ctx->nstackw = stackw.size();
  ctx->nprodw = prodw.size();
  ctx->d_einfo = upload(info.data(), info.size(), err);
  ctx->d_stackw = upload(stackw.data(), stackw.size(), err);
  ctx->d_prodw = upload(prodw.data(), prodw.size(), err);
  ctx->d_ecan = upload(ecan.data(), ecan.size(), err);
  ctx->table_bytes += ctx->index_len * 4 + (ctx->d_index8 ? ctx->index_len : 0) +
                      info.size() * sizeof(EntInfo) + stackw.size() * 4 + prodw.size() * 8 +
                      ecan.size() * 4;
  return err.empty();
}

}  // namespace

extern "C" {

const char *alpacc_create_error(void) { return g_create_error.c_str(); }

alpacc_ctx *alpacc_create(const alpacc_tables *t, int device) {
  g_create_error.clear();
  if (!t) {
    g_create_error = "null tables";
    return nullptr;
  }
  if (hipSetDevice(device) != hipSuccess) {
    g_create_error = "hipSetDevice failed";
    return nullptr;
  }
  alpacc_ctx *ctx = new alpacc_ctx();
  ctx->device = device;
  // the allocation-free entry's scratch address on this device (nullptr: that entry reports it)
  ctx->noalloc_scratch = noalloc_scratch_setup(device);
  (void)hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device);
  std::string &err = g_create_error;
  if (hipHostMalloc((void **)&ctx->h_small, 256, hipHostMallocDefault) != hipSuccess) {
    err = "hipHostMalloc failed";
    alpacc_destroy(ctx);
    return nullptr;
  }
  if (t->num_endos > 0) {
    if (t->num_endos > 65535 || t->num_states > 256 || t->num_classes > 256 ||
        t->init_state >= t->num_states) {
      err = "lexer tables out of range";
      alpacc_destroy(ctx);
      return nullptr;
    }
    ctx->has_lexer = true;
    ctx->e = t->num_endos;
    ctx->C = t->num_classes;
    ctx->m1 = t->num_states;
    ctx->init = t->init_state;
    ctx->ignore = t->ignore_terminal;
    ctx->dead = t->dead_terminal;
    ctx->d_cls = upload(t->byte_class, 256, err);
    ctx->d_step = upload(t->endo_step, (size_t)ctx->e * ctx->C, err);
    ctx->d_comp = upload(t->endo_comp, (size_t)ctx->e * ctx->e, err);
    ctx->d_apply = upload(t->endo_apply, (size_t)ctx->e * ctx->m1, err);
    {
      // lex_emit's walk table, one word per (state, class) of dfa_step: the next state's row
      // as a byte offset into this table (bits 0-17: at most 255 x 256 x 4), whether the label of
      // the byte is kept (not the ignored terminal; bit 18), whether a token boundary falls before
      // the byte (the produce bit; bit 19), whether the next state accepts (bit 20) and the label
      // (bits 24-30).
      const size_t nw = (size_t)ctx->m1 * ctx->C;
      std::vector<uint32_t> walk(nw);
      for (size_t i = 0; i < nw && err.empty(); ++i) {
        const uint32_t e = t->dfa_step[i];
        const uint32_t ns = e & 0xFFu, prod = (e >> 8) & 1u, lab = e >> 9;
        if (ns >= ctx->m1) {
          err = "lexer tables out of range";
          break;
        }
        const bool kept = ctx->ignore < 0 || lab != (uint32_t)ctx->ignore;
        walk[i] = ns * ctx->C * 4u | (kept ? 1u << 18 : 0u) | prod << 19 |
                  (t->state_accept[ns] ? 1u << 20 : 0u) | lab << 24;
      }
      if (err.empty()) ctx->d_walk = upload(walk.data(), nw, err);
    }
    if (!err.empty()) {
      alpacc_destroy(ctx);
      return nullptr;
    }
    if (place_lexer_tables(ctx, err) != 0) {
      alpacc_destroy(ctx);
      return nullptr;
    }
  }
  if (ctx->has_lexer)
    ctx->table_bytes += 256 + (uint64_t)ctx->e * ctx->C * 2 + (uint64_t)ctx->e * ctx->e * 2 +
                        (uint64_t)ctx->e * ctx->m1 + (uint64_t)ctx->m1 * ctx->C * 4;
  if (t->num_productions > 0 && (t->pair_index || t->pair_hash_len)) {
    ctx->has_parser = true;
    ctx->q = t->q;
    ctx->k = t->k;
    ctx->nT = t->num_terminals;
    ctx->start_t = t->start_terminal;
    ctx->end_t = t->end_terminal;
    ctx->empty_t = t->empty_terminal;
    ctx->E = t->num_pair_entries;
    if (!build_parser_tables(ctx, t, err)) {
      alpacc_destroy(ctx);
      return nullptr;
    }
  }
  return ctx;
}

void alpacc_destroy(alpacc_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  for (auto &s : ctx->scratch)
    if (s.first) (void)hipFree(s.first);
  void *tabs[] = {ctx->d_cls,    ctx->d_step,     ctx->d_comp,   ctx->d_apply,
                  ctx->d_walk,   ctx->d_pair_index, ctx->d_index8,
                  ctx->d_hkeys,  ctx->d_hvals,
                  ctx->d_einfo,  ctx->d_stackw,   ctx->d_prodw,  ctx->d_ecan};
  for (void *p : tabs)
    if (p) (void)hipFree(p);
  if (ctx->h_small) (void)hipHostFree(ctx->h_small);
  ctx->drop_graph();
  if (ctx->cap_stream) (void)hipStreamDestroy(ctx->cap_stream);
  if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
  for (auto &pair : ctx->ev)
    for (auto &e : pair)
      if (e) (void)hipEventDestroy(e);
  delete ctx;
}

const char *alpacc_last_error(alpacc_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int alpacc_last_launch_count(alpacc_ctx *ctx) { return ctx ? ctx->launches : 0; }

int alpacc_test_force_timeout(alpacc_ctx *ctx, int on) {
  if (!ctx) return ALPACC_ERR_ARG;
  ctx->force_timeout = on != 0;
  return 0;
}

int alpacc_set_tuning(alpacc_ctx *ctx, const alpacc_tuning *t) {
  if (!ctx || !t) return ALPACC_ERR_ARG;
  if (t->graph_mode < 0 || t->graph_mode > 2 || t->lex_blocks_per_cu < 1 ||
      t->lex_blocks_per_cu > 8 || t->keys_blocks_per_cu < 1 || t->keys_blocks_per_cu > 8 ||
      t->search_max_blocks < 64 || t->search_max_blocks > (1u << 20) ||
      t->brackets_per_kib < 1 || t->brackets_per_kib > (1u << 20) ||
      (t->search_block_threads != 256 && t->search_block_threads != 512 &&
       t->search_block_threads != 1024) ||
      (t->lexer_tables_in_lds != 0 && t->lexer_tables_in_lds != 1)) {
    ctx->err = "tuning value out of range";
    return ALPACC_ERR_ARG;
  }
  if (t->graph_mode != ctx->graph_mode) {
    ctx->drop_graph();
    ctx->last_plan.clear();
  }
  ctx->graph_mode = t->graph_mode;
  ctx->lex_bpc = t->lex_blocks_per_cu;
  ctx->keys_bpc = t->keys_blocks_per_cu;
  ctx->search_cap = t->search_max_blocks;
  ctx->b_per_byte = t->brackets_per_kib / 1024.0;
  ctx->search_bs = t->search_block_threads;
  if ((t->lexer_tables_in_lds != 0) != (ctx->lex_tables_lds != 0) && ctx->has_lexer) {
    ctx->lex_tables_lds = t->lexer_tables_in_lds != 0;
    ctx->drop_graph();
    ctx->last_plan.clear();
    std::string err;
    if (place_lexer_tables(ctx, err) != 0) {
      ctx->err = err;
      return ALPACC_ERR_HIP;
    }
  }
  return 0;
}

int alpacc_get_tuning(alpacc_ctx *ctx, alpacc_tuning *t) {
  if (!ctx || !t) return ALPACC_ERR_ARG;
  t->graph_mode = ctx->graph_mode;
  t->lex_blocks_per_cu = ctx->lex_bpc;
  t->keys_blocks_per_cu = ctx->keys_bpc;
  t->search_max_blocks = ctx->search_cap;
  t->brackets_per_kib = (uint32_t)std::min(ctx->b_per_byte * 1024.0 + 0.5, (double)(1u << 20));
  t->search_block_threads = ctx->search_bs;
  t->lexer_tables_in_lds = ctx->lex_tables_lds;
  t->graph_launches = ctx->graph_launches;
  t->graph_captures = ctx->graph_captures;
  return 0;
}

uint64_t alpacc_device_bytes(alpacc_ctx *ctx) {
  if (!ctx) return 0;
  uint64_t b = ctx->table_bytes;
  for (const auto &s : ctx->scratch) b += s.second;
  return b;
}

int alpacc_last_counts(alpacc_ctx *ctx, uint64_t *counts) {
  if (!ctx || !counts) return ALPACC_ERR_ARG;
  for (int i = 0; i < 4; ++i) counts[i] = ctx->last_counts[i];
  return 0;
}

int alpacc_profile_enable(alpacc_ctx *ctx, int on) {
  if (!ctx) return ALPACC_ERR_ARG;
  if (on && !ctx->ev[0][0]) {
    for (auto &pair : ctx->ev)
      for (auto &e : pair)
        if (hipEventCreate(&e) != hipSuccess) return ALPACC_ERR_HIP;
  }
  ctx->prof = on != 0;
  for (int i = 0; i < NSTAGE; ++i) {
    ctx->prof_ms[i] = 0.0;
    ctx->prof_calls[i] = 0;
    ctx->rec[i] = false;
  }
  return 0;
}

int alpacc_profile_read(alpacc_ctx *ctx, const char **names, double *ms, uint64_t *calls,
                        int max_stages) {
  if (!ctx) return ALPACC_ERR_ARG;
  const int m = std::min(max_stages, (int)NSTAGE);
  for (int i = 0; i < m; ++i) {
    if (names) names[i] = kStageNames[i];
    if (ms) ms[i] = ctx->prof_ms[i];
    if (calls) calls[i] = ctx->prof_calls[i];
  }
  for (int i = 0; i < NSTAGE; ++i) {
    ctx->prof_ms[i] = 0.0;
    ctx->prof_calls[i] = 0;
  }
  return m;
}

int alpacc_lex(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint32_t *d_kinds,
               uint32_t *d_starts, uint32_t *d_ends, uint64_t cap, uint64_t *num_tokens,
               void *stream) {
  if (!ctx || !num_tokens || (n > 0 && (!d_in || !d_kinds || !d_starts || !d_ends))) return ALPACC_ERR_ARG;
  if (cap < n) {
    ctx->err = "token capacity must be at least the input length";
    return ALPACC_ERR_CAPACITY;
  }
  ctx->launches = 0;
  ctx->err.clear();
  (void)hipSetDevice(ctx->device);
  const int rc = run_lex(ctx, d_in, n, d_kinds, d_starts, d_ends, num_tokens, (hipStream_t)stream);
  if (ctx->prof) ctx->prof_collect();
  return rc;
}

int alpacc_lex_noalloc(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint32_t *d_kinds,
                       uint32_t *d_starts, uint32_t *d_ends, uint64_t cap, uint64_t *num_tokens,
                       void *stream) {
  if (!ctx || !num_tokens || (n > 0 && (!d_in || !d_kinds || !d_starts || !d_ends))) return ALPACC_ERR_ARG;
  ctx->launches = 0;
  ctx->err.clear();
  if (n > NOALLOC_MAX_BYTES) {
    ctx->err = "alpacc_lex_noalloc takes documents of at most ALPACC_NOALLOC_MAX_BYTES (64 MiB)";
    return ALPACC_ERR_ARG;
  }
  if (cap < n) {
    ctx->err = "token capacity must be at least the input length";
    return ALPACC_ERR_CAPACITY;
  }
  if (!ctx->noalloc_scratch) {
    ctx->err = "no address for the allocation-free scratch on this device";
    return ALPACC_ERR_HIP;
  }
  (void)hipSetDevice(ctx->device);
  LexScratch sc;
  lex_scratch_carve((uint8_t *)ctx->noalloc_scratch, n, sc);
  const int rc = run_lex(ctx, d_in, n, d_kinds, d_starts, d_ends, num_tokens, (hipStream_t)stream,
                         0, nullptr, nullptr, true, &sc);
  if (ctx->prof) ctx->prof_collect();
  return rc;
}

int alpacc_parse_tokens(alpacc_ctx *ctx, const uint32_t *d_kinds, const uint32_t *d_starts,
                        const uint32_t *d_ends, uint64_t num_tokens, uint32_t *d_parent,
                        uint32_t *d_node, uint32_t *d_span, uint64_t cap, uint64_t *num_nodes,
                        void *stream) {
  if (!ctx || !num_nodes) return ALPACC_ERR_ARG;
  ctx->launches = 0;
  ctx->err.clear();
  (void)hipSetDevice(ctx->device);
  const int rc = run_parse_tokens<uint32_t>(ctx, d_kinds, d_starts, d_ends, num_tokens, d_parent,
                                            d_node, d_span, cap, num_nodes, (hipStream_t)stream);
  if (rc != ALPACC_ACCEPT) ctx->prof_drop(S_PARSE);
  if (ctx->prof) ctx->prof_collect();
  return rc;
}

int alpacc_parse(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint32_t *d_parent,
                 uint32_t *d_node, uint32_t *d_span, uint64_t cap, uint64_t *num_nodes,
                 void *stream) {
  if (!ctx || !num_nodes) return ALPACC_ERR_ARG;
  ctx->launches = 0;
  ctx->err.clear();
  *num_nodes = 0;
  if (!ctx->has_parser) {
    ctx->err = "bundle has no parser";
    return ALPACC_ERR_NO_PARSER;
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const uint64_t tcap = n > 0 ? n : 1;
  // token kinds as bytes (labels have 7 bits): only the key lookups read them
  uint8_t *kinds8 = (uint8_t *)ctx->buf(B_TOK_KIND, tcap + 16);
  uint32_t *starts = (uint32_t *)ctx->buf(B_TOK_START, tcap * 4);
  uint32_t *ends = (uint32_t *)ctx->buf(B_TOK_END, tcap * 4);
  if (!kinds8 || !starts || !ends) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  // the parser starts on the lexer's device results; the lexer's verdict and token count come
  // back with the parser's first host synchronisation
  uint64_t T = 0;
  int rc = run_lex(ctx, d_in, n, nullptr, starts, ends, &T, st, 0, nullptr, nullptr, false,
                   nullptr, kinds8);
  if (rc != ALPACC_ACCEPT) {
    if (ctx->prof) ctx->prof_collect();
    return rc;
  }
  const int lex_launches = ctx->launches;
  rc = run_parse_tokens<uint32_t>(ctx, nullptr, starts, ends, n, d_parent, d_node, d_span, cap,
                                  num_nodes, st, false, true, kinds8);
  ctx->launches += lex_launches;
  if (rc != ALPACC_ACCEPT) ctx->prof_drop(S_PARSE);
  if (ctx->prof) ctx->prof_collect();
  return rc;
}

int alpacc_lex_chunked(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint64_t chunk_bytes,
                       uint32_t *d_kinds, uint64_t *d_starts, uint64_t *d_ends, uint64_t cap,
                       uint64_t *num_tokens, void *stream) {
  if (!ctx || !num_tokens || chunk_bytes == 0 ||
      (n > 0 && (!d_in || !d_kinds || !d_starts || !d_ends)))
    return ALPACC_ERR_ARG;
  const uint64_t c = std::min(chunk_bytes, std::max<uint64_t>(n, 1));
  if (cap < n + (n + c - 1) / c) {
    ctx->err = "token capacity must be at least n + ceil(n / chunk_bytes)";
    return ALPACC_ERR_CAPACITY;
  }
  ctx->launches = 0;
  ctx->err.clear();
  (void)hipSetDevice(ctx->device);
  const int rc = run_lex(ctx, d_in, n, d_kinds, nullptr, nullptr, num_tokens, (hipStream_t)stream,
                         chunk_bytes, d_starts, d_ends);
  if (ctx->prof) ctx->prof_collect();
  return rc;
}

int alpacc_parse_chunked(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint64_t chunk_bytes,
                         uint32_t *d_parent, uint32_t *d_node, uint64_t *d_span, uint64_t cap,
                         uint64_t *num_nodes, void *stream) {
  if (!ctx || !num_nodes || chunk_bytes == 0) return ALPACC_ERR_ARG;
  ctx->launches = 0;
  ctx->err.clear();
  *num_nodes = 0;
  if (!ctx->has_parser) {
    ctx->err = "bundle has no parser";
    return ALPACC_ERR_NO_PARSER;
  }
  (void)hipSetDevice(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  const uint64_t c = std::min(chunk_bytes, std::max<uint64_t>(n, 1));
  const uint64_t tcap = n + (n + c - 1) / c + 1;
  uint32_t *kinds = (uint32_t *)ctx->buf(B_TOK_KIND, tcap * 4);
  uint64_t *starts = (uint64_t *)ctx->buf(B_TOK_START64, tcap * 8);
  uint64_t *ends = (uint64_t *)ctx->buf(B_TOK_END64, tcap * 8);
  if (!kinds || !starts || !ends) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  uint64_t T = 0;
  int rc = run_lex(ctx, d_in, n, kinds, nullptr, nullptr, &T, st, chunk_bytes, starts, ends);
  if (rc != ALPACC_ACCEPT) {
    if (ctx->prof) ctx->prof_collect();
    return rc;
  }
  const int lex_launches = ctx->launches;
  rc = run_parse_tokens<uint64_t>(ctx, kinds, starts, ends, T, d_parent, d_node, d_span, cap,
                                  num_nodes, st);
  ctx->launches += lex_launches;
  if (rc != ALPACC_ACCEPT) ctx->prof_drop(S_PARSE);
  if (ctx->prof) ctx->prof_collect();
  return rc;
}

int alpacc_lex_batch(alpacc_ctx *ctx, const uint8_t *d_in, const uint64_t *offsets,
                     const uint64_t *lengths, uint32_t num_docs, uint32_t *d_kinds,
                     uint32_t *d_starts, uint32_t *d_ends, uint64_t cap,
                     alpacc_doc_result *results, void *stream) {
  if (!ctx || !offsets || !lengths || !results || !d_kinds || !d_starts || !d_ends || num_docs == 0)
    return ALPACC_ERR_ARG;
  uint64_t need = 0;
  for (uint32_t d = 0; d < num_docs; ++d) need += lengths[d] + 1;
  if (cap < need) {
    ctx->err = "token capacity must be at least the sum of (length + 1)";
    return ALPACC_ERR_CAPACITY;
  }
  ctx->launches = 0;
  ctx->err.clear();
  (void)hipSetDevice(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  uint32_t *doc_res = (uint32_t *)ctx->buf(B_DOC_RES, (2ull * num_docs + 2) * 4);
  alpacc_doc_result *d_results =
      (alpacc_doc_result *)ctx->buf(B_RESULTS, (num_docs + 1) * sizeof(alpacc_doc_result));
  if (!doc_res || !d_results) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  LexDoc *ldocs = nullptr;
  int rc = run_lex_batch(ctx, d_in, offsets, lengths, num_docs, d_kinds, d_starts, d_ends, doc_res,
                         &ldocs, st);
  if (rc == 0) {
    lb_finish<<<(num_docs + 255) / 256, 256, 0, st>>>(ldocs, doc_res, num_docs, d_results);
    ctx->launches++;
    uint32_t *small = (uint32_t *)ctx->buf(B_SMALL, 64 * sizeof(uint32_t));
    if (hipMemcpyAsync(results, d_results, num_docs * sizeof(alpacc_doc_result),
                       hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(ctx->h_small, small + 4, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                       st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      ctx->err = "HIP error in the batched lexer";
      rc = ALPACC_ERR_HIP;
    } else if (ctx->h_small[2]) {
      // a bounded look-back wait gave up: the offsets behind it are wrong
      ctx->err = "lexer look-back did not complete";
      rc = ALPACC_ERR_TIMEOUT;
    } else {
      rc = ALPACC_ACCEPT;
    }
  }
  if (ctx->prof) ctx->prof_collect();
  return rc;
}

int alpacc_parse_batch(alpacc_ctx *ctx, const uint8_t *d_in, const uint64_t *offsets,
                       const uint64_t *lengths, uint32_t num_docs, uint32_t *d_parent,
                       uint32_t *d_node, uint32_t *d_span, uint64_t cap, uint64_t *total_nodes,
                       alpacc_doc_result *results, void *stream) {
  if (!ctx || !offsets || !lengths || !results || !total_nodes || num_docs == 0)
    return ALPACC_ERR_ARG;
  ctx->launches = 0;
  ctx->err.clear();
  (void)hipSetDevice(ctx->device);
  const int rc = run_parse_batch(ctx, d_in, offsets, lengths, num_docs, d_parent, d_node, d_span,
                                 cap, total_nodes, results, (hipStream_t)stream);
  if (ctx->prof) ctx->prof_collect();
  return rc;
}

}  // extern "C"

namespace {

// test.fut wire format: big-endian u64 (encode_u64, test.fut:9-18), one-byte booleans.
void put_u64(std::vector<uint8_t> &o, uint64_t v) {
  for (int k = 7; k >= 0; --k) o.push_back((uint8_t)(v >> (8 * k)));
}

bool get_u64(const uint8_t *b, uint64_t size, uint64_t &pos, uint64_t &v) {
  if (pos + 8 > size) return false;
  v = 0;
  for (int k = 0; k < 8; ++k) v = (v << 8) | b[pos + k];
  pos += 8;
  return true;
}

struct DevArr {
  void *p = nullptr;
  size_t n = 0;
  void *get(size_t bytes) {
    if (n < bytes) {
      if (p) (void)hipFree(p);
      p = nullptr;
      n = 0;
      if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) return nullptr;
      n = std::max<size_t>(bytes, 256);
    }
    return p;
  }
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
};

// Lexer+parser or lexer-only artifact: every input through parse_int / lex_int of the generated
// entries, i.e. with the reference's 16 MiB chunk (Generator.hs:24-76). Inputs up to one chunk go
// through the batched entry (one chunk: both modes agree), longer ones through the chunked call.
int test_bytes(alpacc_ctx *ctx, const std::vector<std::pair<uint64_t, uint64_t>> &ins,
               const uint8_t *blob, std::vector<uint8_t> &out, hipStream_t st) {
  const bool parse = ctx->has_parser;
  const uint64_t C = ALPACC_REFERENCE_CHUNK;
  const uint32_t D = (uint32_t)ins.size();
  std::vector<int32_t> status(D, 0);
  std::vector<std::vector<uint8_t>> rec(D);
  // documents up to one chunk: one batched call
  std::vector<uint32_t> small;
  std::vector<uint64_t> offs, lens;
  uint64_t total = 0;
  for (uint32_t i = 0; i < D; ++i)
    if (ins[i].second <= C) {
      small.push_back(i);
      offs.push_back(total);
      lens.push_back(ins[i].second);
      total += ins[i].second;
    }
  DevArr din, d0, d1, d2;
  uint8_t *dev = (uint8_t *)din.get(total + 16);
  if (!dev) return ALPACC_ERR_HIP;
  for (size_t j = 0; j < small.size(); ++j)
    if (lens[j] && hipMemcpyAsync(dev + offs[j], blob + ins[small[j]].first, lens[j],
                                  hipMemcpyHostToDevice, st) != hipSuccess)
      return ALPACC_ERR_HIP;
  if (!small.empty()) {
    std::vector<alpacc_doc_result> res(small.size());
    if (parse) {
      uint64_t cap = total + 16, nodes = 0;
      int rc;
      for (;;) {
        void *p0 = d0.get(cap * 4), *p1 = d1.get(cap * 4), *p2 = d2.get(cap * 8);
        if (!p0 || !p1 || !p2) return ALPACC_ERR_HIP;
        rc = run_parse_batch(ctx, dev, offs.data(), lens.data(), (uint32_t)small.size(),
                             (uint32_t *)p0, (uint32_t *)p1, (uint32_t *)p2, cap, &nodes,
                             res.data(), st);
        if (rc == ALPACC_ERR_CAPACITY) {
          cap = nodes + 16;
          continue;
        }
        break;
      }
      if (rc < 0) return rc;
      std::vector<uint32_t> par(nodes), nd(nodes), sp(2 * nodes);
      if (nodes &&
          (hipMemcpyAsync(par.data(), d0.p, nodes * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
           hipMemcpyAsync(nd.data(), d1.p, nodes * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
           hipMemcpyAsync(sp.data(), d2.p, nodes * 8, hipMemcpyDeviceToHost, st) != hipSuccess))
        return ALPACC_ERR_HIP;
      if (hipStreamSynchronize(st) != hipSuccess) return ALPACC_ERR_HIP;
      for (size_t j = 0; j < small.size(); ++j) {
        std::vector<uint8_t> &o = rec[small[j]];
        if (res[j].status != ALPACC_ACCEPT) continue;
        status[small[j]] = 1;
        put_u64(o, res[j].count);
        for (uint64_t k = res[j].offset; k < res[j].offset + res[j].count; ++k) {
          const bool term = (nd[k] & ALPACC_NODE_TERMINAL) != 0;
          o.push_back(term ? 1 : 0);
          put_u64(o, par[k]);
          put_u64(o, nd[k] & ~ALPACC_NODE_TERMINAL);
          put_u64(o, term ? sp[2 * k] : 0);
          put_u64(o, term ? sp[2 * k + 1] : 0);
        }
      }
    } else {
      const uint64_t cap = total + small.size() + 1;
      void *p0 = d0.get(cap * 4), *p1 = d1.get(cap * 4), *p2 = d2.get(cap * 4);
      if (!p0 || !p1 || !p2) return ALPACC_ERR_HIP;
      std::vector<uint32_t> doc_res_unused;
      uint32_t *doc_res = (uint32_t *)ctx->buf(B_DOC_RES, (2ull * small.size() + 2) * 4);
      alpacc_doc_result *d_results =
          (alpacc_doc_result *)ctx->buf(B_RESULTS, (small.size() + 1) * sizeof(alpacc_doc_result));
      if (!doc_res || !d_results) return ALPACC_ERR_HIP;
      LexDoc *ldocs = nullptr;
      int rc = run_lex_batch(ctx, dev, offs.data(), lens.data(), (uint32_t)small.size(),
                             (uint32_t *)p0, (uint32_t *)p1, (uint32_t *)p2, doc_res, &ldocs, st);
      if (rc < 0) return rc;
      lb_finish<<<((uint32_t)small.size() + 255) / 256, 256, 0, st>>>(ldocs, doc_res,
                                                                       (uint32_t)small.size(),
                                                                       d_results);
      std::vector<uint32_t> k(cap), sa(cap), ea(cap);
      uint32_t *dsmall = (uint32_t *)ctx->buf(B_SMALL, 64 * sizeof(uint32_t));
      if (hipMemcpyAsync(res.data(), d_results, small.size() * sizeof(alpacc_doc_result),
                         hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipMemcpyAsync(k.data(), p0, cap * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipMemcpyAsync(sa.data(), p1, cap * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipMemcpyAsync(ea.data(), p2, cap * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipMemcpyAsync(ctx->h_small, dsmall + 4, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                         st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return ALPACC_ERR_HIP;
      if (ctx->h_small[2]) {
        ctx->err = "lexer look-back did not complete";
        return ALPACC_ERR_TIMEOUT;
      }
      for (size_t j = 0; j < small.size(); ++j) {
        std::vector<uint8_t> &o = rec[small[j]];
        if (res[j].status != ALPACC_ACCEPT) continue;
        status[small[j]] = 1;
        put_u64(o, res[j].count);
        for (uint64_t t = res[j].offset; t < res[j].offset + res[j].count; ++t) {
          put_u64(o, k[t]);
          put_u64(o, sa[t]);
          put_u64(o, ea[t]);
        }
      }
    }
  }
  // documents longer than one chunk: the chunked calls
  for (uint32_t i = 0; i < D; ++i) {
    const uint64_t n = ins[i].second;
    if (n <= C) continue;
    uint8_t *d = (uint8_t *)din.get(n + 16);
    if (!d || hipMemcpyAsync(d, blob + ins[i].first, n, hipMemcpyHostToDevice, st) != hipSuccess)
      return ALPACC_ERR_HIP;
    std::vector<uint8_t> &o = rec[i];
    if (parse) {
      const uint64_t tcap = n + (n + C - 1) / C + 1;
      uint32_t *kinds = (uint32_t *)ctx->buf(B_TOK_KIND, tcap * 4);
      uint64_t *s64 = (uint64_t *)ctx->buf(B_TOK_START64, tcap * 8);
      uint64_t *e64 = (uint64_t *)ctx->buf(B_TOK_END64, tcap * 8);
      if (!kinds || !s64 || !e64) return ALPACC_ERR_HIP;
      uint64_t T = 0, N = 0;
      int rc = run_lex(ctx, d, n, kinds, nullptr, nullptr, &T, st, C, s64, e64);
      if (rc < 0) return rc;
      if (rc != ALPACC_ACCEPT) continue;
      uint64_t cap = 2 * n + 16;
      for (;;) {
        void *p0 = d0.get(cap * 4), *p1 = d1.get(cap * 4), *p2 = d2.get(cap * 16);
        if (!p0 || !p1 || !p2) return ALPACC_ERR_HIP;
        rc = run_parse_tokens<uint64_t>(ctx, kinds, s64, e64, T, (uint32_t *)p0, (uint32_t *)p1,
                                        (uint64_t *)p2, cap, &N, st);
        if (rc == ALPACC_ERR_CAPACITY) {
          cap = N + 16;
          continue;
        }
        break;
      }
      if (rc < 0) return rc;
      if (rc != ALPACC_ACCEPT) continue;
      std::vector<uint32_t> par(N), nd(N);
      std::vector<uint64_t> sp(2 * N);
      if (hipMemcpyAsync(par.data(), d0.p, N * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipMemcpyAsync(nd.data(), d1.p, N * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipMemcpyAsync(sp.data(), d2.p, N * 16, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return ALPACC_ERR_HIP;
      status[i] = 1;
      put_u64(o, N);
      for (uint64_t k = 0; k < N; ++k) {
        const bool term = (nd[k] & ALPACC_NODE_TERMINAL) != 0;
        o.push_back(term ? 1 : 0);
        put_u64(o, par[k]);
        put_u64(o, nd[k] & ~ALPACC_NODE_TERMINAL);
        put_u64(o, term ? sp[2 * k] : 0);
        put_u64(o, term ? sp[2 * k + 1] : 0);
      }
    } else {
      const uint64_t cap = n + (n + C - 1) / C + 1;
      void *p0 = d0.get(cap * 4), *p1 = d1.get(cap * 8), *p2 = d2.get(cap * 8);
      if (!p0 || !p1 || !p2) return ALPACC_ERR_HIP;
      uint64_t T = 0;
      int rc = run_lex(ctx, d, n, (uint32_t *)p0, nullptr, nullptr, &T, st, C, (uint64_t *)p1,
                       (uint64_t *)p2);
      if (rc < 0) return rc;
      if (rc != ALPACC_ACCEPT) continue;
      std::vector<uint32_t> k(T);
      std::vector<uint64_t> sa(T), ea(T);
      if (T && (hipMemcpyAsync(k.data(), p0, T * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(sa.data(), p1, T * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(ea.data(), p2, T * 8, hipMemcpyDeviceToHost, st) != hipSuccess))
        return ALPACC_ERR_HIP;
      if (hipStreamSynchronize(st) != hipSuccess) return ALPACC_ERR_HIP;
      status[i] = 1;
      put_u64(o, T);
      for (uint64_t t = 0; t < T; ++t) {
        put_u64(o, k[t]);
        put_u64(o, sa[t]);
        put_u64(o, ea[t]);
      }
    }
  }
  for (uint32_t i = 0; i < D; ++i) {
    out.push_back(status[i] ? 1 : 0);
    if (status[i]) out.insert(out.end(), rec[i].begin(), rec[i].end());
  }
  return 0;
}

// Parser-only artifact (parser_test, test.fut:60-98): every input is an array of u64 terminals,
// the output its production sequence (pre_productions_int).
int test_terminals(alpacc_ctx *ctx, const uint8_t *blob, uint64_t size, uint64_t pos,
                   uint64_t count, std::vector<uint8_t> &out, hipStream_t st) {
  DevArr dk, ds, d0, d1, d2;
  for (uint64_t i = 0; i < count; ++i) {
    uint64_t m = 0;
    if (!get_u64(blob, size, pos, m) || m > (size - pos) / 8) return ALPACC_ERR_ARG;
    std::vector<uint32_t> kinds(m);
    bool fits = true;
    for (uint64_t t = 0; t < m; ++t) {
      uint64_t v = 0;
      get_u64(blob, size, pos, v);
      fits &= v < ctx->nT;
      kinds[t] = (uint32_t)std::min<uint64_t>(v, 0xFFFFFFFFull);
    }
    if (!fits) {  // a terminal the table does not know: no key matches
      out.push_back(0);
      continue;
    }
    uint32_t *k = (uint32_t *)dk.get((m + 1) * 4);
    uint32_t *zs = (uint32_t *)ds.get((m + 1) * 4);
    if (!k || !zs) return ALPACC_ERR_HIP;
    if (m && hipMemcpyAsync(k, kinds.data(), m * 4, hipMemcpyHostToDevice, st) != hipSuccess)
      return ALPACC_ERR_HIP;
    uint64_t cap = 4 * m + 16, N = 0;
    int rc;
    for (;;) {
      void *p0 = d0.get(cap * 4), *p1 = d1.get(cap * 4), *p2 = d2.get(cap * 8);
      if (!p0 || !p1 || !p2) return ALPACC_ERR_HIP;
      rc = run_parse_tokens<uint32_t>(ctx, k, zs, zs, m, (uint32_t *)p0, (uint32_t *)p1,
                                      (uint32_t *)p2, cap, &N, st, true);
      if (rc == ALPACC_ERR_CAPACITY) {
        cap = N + 16;
        continue;
      }
      break;
    }
    if (rc < 0) return rc;
    if (rc != ALPACC_ACCEPT) {
      out.push_back(0);
      continue;
    }
    std::vector<uint32_t> nd(N);
    if (N && hipMemcpyAsync(nd.data(), d1.p, N * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
      return ALPACC_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return ALPACC_ERR_HIP;
    out.push_back(1);
    put_u64(out, N);
    for (uint64_t t = 0; t < N; ++t) put_u64(out, nd[t]);
  }
  return 0;
}

}  // namespace

extern "C" {

int alpacc_test(alpacc_ctx *ctx, const uint8_t *blob, uint64_t size, uint8_t **out,
                uint64_t *out_size, void *stream) {
  if (!ctx || !blob || !out || !out_size) return ALPACC_ERR_ARG;
  *out = nullptr;
  *out_size = 0;
  ctx->launches = 0;
  ctx->err.clear();
  (void)hipSetDevice(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  uint64_t pos = 0, count = 0;
  if (!get_u64(blob, size, pos, count)) {
    ctx->err = "test input shorter than its count";
    return ALPACC_ERR_ARG;
  }
  std::vector<uint8_t> o;
  put_u64(o, count);
  int rc;
  if (ctx->has_lexer) {
    std::vector<std::pair<uint64_t, uint64_t>> ins;
    for (uint64_t i = 0; i < count; ++i) {
      uint64_t n = 0;
      if (!get_u64(blob, size, pos, n) || n > size - pos) {
        ctx->err = "test input truncated";
        return ALPACC_ERR_ARG;
      }
      ins.emplace_back(pos, n);
      pos += n;
    }
    rc = ins.empty() ? 0 : test_bytes(ctx, ins, blob, o, st);
  } else if (ctx->has_parser) {
    rc = test_terminals(ctx, blob, size, pos, count, o, st);
  } else {
    rc = ALPACC_ERR_ARG;
  }
  if (rc < 0) {
    if (ctx->err.empty()) ctx->err = "test entry failed";
    return rc;
  }
  uint8_t *buf = (uint8_t *)malloc(o.size() ? o.size() : 1);
  if (!buf) return ALPACC_ERR_HIP;
  memcpy(buf, o.data(), o.size());
  *out = buf;
  *out_size = o.size();
  return ALPACC_ACCEPT;
}

void alpacc_free(void *ptr) { free(ptr); }

int alpacc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int alpacc_set_device(int device) { return hipSetDevice(device) == hipSuccess ? 0 : -1; }

void *alpacc_device_alloc(uint64_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) return nullptr;
  return p;
}

void alpacc_device_free(void *ptr) {
  if (ptr) (void)hipFree(ptr);
}

void *alpacc_host_alloc(uint64_t bytes) {
  void *p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

void alpacc_host_free(void *ptr) {
  if (ptr) (void)hipHostFree(ptr);
}

int alpacc_copy_to_device(void *dst, const void *src, uint64_t bytes, void *stream) {
  if (bytes == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) != hipSuccess) return -1;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : -1;
}

int alpacc_copy_to_host(void *dst, const void *src, uint64_t bytes, void *stream) {
  if (bytes == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : -1;
}

int alpacc_copy_device(void *dst, const void *src, uint64_t bytes, void *stream) {
  if (bytes == 0) return 0;
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess
             ? 0
             : -1;
}

int alpacc_device_memset(void *dst, int value, uint64_t bytes, void *stream) {
  if (bytes == 0) return 0;
  return hipMemsetAsync(dst, value, bytes, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

void *alpacc_stream_create(void) {
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
  return (void *)s;
}

void alpacc_stream_destroy(void *stream) {
  if (stream) (void)hipStreamDestroy((hipStream_t)stream);
}

int alpacc_stream_sync(void *stream) {
  return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : -1;
}

void *alpacc_event_create(void) {
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return (void *)e;
}

void alpacc_event_destroy(void *event) {
  if (event) (void)hipEventDestroy((hipEvent_t)event);
}

int alpacc_event_record(void *event, void *stream) {
  return hipEventRecord((hipEvent_t)event, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

int alpacc_event_sync(void *event) {
  return hipEventSynchronize((hipEvent_t)event) == hipSuccess ? 0 : -1;
}

float alpacc_event_elapsed_ms(void *start, void *end) {
  if (hipEventSynchronize((hipEvent_t)end) != hipSuccess) return -1.0f;
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, (hipEvent_t)start, (hipEvent_t)end) != hipSuccess) return -1.0f;
  return ms;
}

}  // extern "C"
