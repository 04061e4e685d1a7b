// alpacc_hip.hip - MI355X (gfx950, CDNA4, wave64) lexer and LLP parser for alpacc grammars.
//
// Device path for the runtime of the reference (futhark/lexer.fut, futhark/parser.fut):
//
//   lexer : one persistent kernel. Every wavefront owns a 2 KiB tile (64 lanes x 32 bytes) and
//           runs the whole pipeline for it: byte classes, a serial endomorphism walk per lane
//           with the (endo x class) table in LDS, a wave64 Kogge-Stone scan of the lane
//           endomorphisms with the (endo x endo) table, a decoupled look-back across tiles,
//           then two walks of the plain lexer DFA from the now known lane start states: one to
//           count token ends and token starts, one (after a second look-back over the counts)
//           to write (kind, start, end) of every token. Input bytes are read from HBM once.
//   parser: pair-table lookup per token window, segmented copy of bracket / production spans,
//           bracket depth scan, bracket matching and parents through a 64-ary min-tree
//           "previous smaller-or-equal" search, terminal-node substitution.
//
// All arithmetic is integer. Tables are produced by alpacc_amd/gen (the restated generator).

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string.h>

#include <algorithm>
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
constexpr int LEX_STAGE = 256;                   // tokens staged per wavefront and round

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

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
  return v;
}
__device__ __forceinline__ int32_t wave_max_i32(int32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, WAVE));
  return v;
}
__device__ __forceinline__ int32_t wave_min_i32(int32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = min(v, __shfl_xor(v, d, WAVE));
  return v;
}

struct LexArgs {
  const uint8_t *in;
  uint64_t n;
  uint32_t *kinds;
  uint32_t *starts;
  uint32_t *ends;
  const uint8_t *byte_class;
  const uint16_t *endo_step;
  const uint16_t *endo_comp;
  const uint8_t *endo_apply;
  const uint16_t *dfa_step;
  const uint8_t *state_accept;
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
};

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
__device__ __forceinline__ void load_classes(const LexArgs &p, const uint8_t *s_cls, uint64_t base,
                                             int valid, uint32_t (&cls)[LEX_IPT / 4]) {
  if (valid == LEX_IPT && p.aligned) {
    const uint4 a = *reinterpret_cast<const uint4 *>(p.in + base);
    const uint4 b = *reinterpret_cast<const uint4 *>(p.in + base + 16);
    const uint32_t wds[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const uint32_t x = wds[q];
      cls[q] = (uint32_t)s_cls[x & 0xFF] | ((uint32_t)s_cls[(x >> 8) & 0xFF] << 8) |
               ((uint32_t)s_cls[(x >> 16) & 0xFF] << 16) | ((uint32_t)s_cls[x >> 24] << 24);
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      uint32_t v = 0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = q * 4 + r;
        if (j < valid) v |= (uint32_t)s_cls[p.in[base + j]] << (8 * r);
      }
      cls[q] = v;
    }
  }
}

// K1: per wave tile, the endomorphism of every lane's bytes (serial walk through the
// (endo x class) table in LDS) and their wave-exclusive composition (Blelloch up/down sweep with
// the (endo x endo) table, 126 lookups per wave). Persistent, grid-stride over wave tiles.
template <bool COMP_IN_LDS>
__global__ void __launch_bounds__(LEX_BLOCK) lex_reduce(LexArgs p) {
  extern __shared__ __align__(16) uint8_t smem[];
  const uint32_t e = p.e, C = p.C;
  uint16_t *s_step = reinterpret_cast<uint16_t *>(smem);
  uint8_t *s_cls = reinterpret_cast<uint8_t *>(s_step + e * C);
  uint16_t *s_comp = reinterpret_cast<uint16_t *>(s_cls + 256);
  for (uint32_t i = threadIdx.x; i < e * C; i += blockDim.x) s_step[i] = p.endo_step[i];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_cls[i] = p.byte_class[i];
  if (COMP_IN_LDS)
    for (uint32_t i = threadIdx.x; i < e * e; i += blockDim.x) s_comp[i] = p.endo_comp[i];
  __syncthreads();
  const uint16_t *comp_tab = COMP_IN_LDS ? s_comp : p.endo_comp;
  const int lane = threadIdx.x & (WAVE - 1);
  const uint32_t wpb = blockDim.x / WAVE;
  const uint64_t n = p.n;
  for (uint32_t tile = blockIdx.x * wpb + threadIdx.x / WAVE; tile < p.num_tiles;
       tile += gridDim.x * wpb) {
    const uint64_t base = (uint64_t)tile * LEX_TILE + (uint64_t)lane * LEX_IPT;
    const int valid = base >= n ? 0 : (int)min((uint64_t)LEX_IPT, n - base);
    uint32_t cls[LEX_IPT / 4];
    load_classes(p, s_cls, base, valid, cls);
    uint32_t v = 0;
#pragma unroll
    for (int j = 0; j < LEX_IPT; ++j)
      if (j < valid) v = s_step[v * C + LEX_CLS(j)];
    // up-sweep: lane l with (l+1) % 2d == 0 holds the composition of lanes l-2d+1 .. l
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      const uint32_t left = __shfl(v, (lane - d) & (WAVE - 1), WAVE);
      if (((lane + 1) & (2 * d - 1)) == 0) v = comp_lookup(comp_tab, e, left, v);
    }
    const uint32_t total = __shfl(v, WAVE - 1, WAVE);
    if (lane == WAVE - 1) v = 0;
    // down-sweep: right node r gets prefix(r) then sum(left subtree); left node gets prefix(r)
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) {
      const uint32_t left = __shfl(v, (lane - d) & (WAVE - 1), WAVE);
      const uint32_t right = __shfl(v, (lane + d) & (WAVE - 1), WAVE);
      if (((lane + 1) & (2 * d - 1)) == 0)
        v = comp_lookup(comp_tab, e, v, left);
      else if (((lane + 1 + d) & (2 * d - 1)) == 0)
        v = right;
    }
    p.lane_excl[(uint64_t)tile * WAVE + lane] = (uint16_t)v;
    if (lane == 0) p.tile_agg[tile] = (uint16_t)total;
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

// K2b: one block, exclusive composition of the block totals in place (chunks with a carry).
__global__ void __launch_bounds__(LEX_SCAN_BLOCK) lex_scan_blocks(LexArgs p, uint32_t nblocks) {
  __shared__ uint16_t s[LEX_SCAN_BLOCK];
  __shared__ uint16_t carry_s;
  const uint32_t i = threadIdx.x;
  if (i == 0) carry_s = 0;
  for (uint32_t off = 0; off < nblocks; off += LEX_SCAN_BLOCK) {
    s[i] = off + i < nblocks ? p.block_agg[off + i] : 0;
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
    if (off + i < nblocks) p.block_agg[off + i] = (uint16_t)comp_lookup(p.endo_comp, p.e, carry, s[i]);
    __syncthreads();
    if (i == LEX_SCAN_BLOCK - 1) carry_s = (uint16_t)comp_lookup(p.endo_comp, p.e, carry, chunk_total);
    __syncthreads();
  }
}

// K2c: DFA state at the start of every wave tile.
__global__ void lex_tile_states(LexArgs p) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.num_tiles) return;
  const uint32_t pre = comp_lookup(p.endo_comp, p.e, p.block_agg[t / LEX_SCAN_BLOCK], p.tile_excl[t]);
  p.tile_state[t] = p.endo_apply[pre * p.m1 + p.init];
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
      if (++spins > SPIN_LIMIT || ld_relaxed(p.result + 2) != 0) {
        if (lane == 0) st_relaxed(p.result + 2, 1u);
        cnt = 0;
        last = -1;
        return;
      }
      __builtin_amdgcn_s_sleep(1);
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
template <bool CHUNKED>
__global__ void __launch_bounds__(LEX_BLOCK) lex_emit(LexArgs p) {
  __shared__ uint16_t s_dfa[LEX_MAX_DFA];
  __shared__ uint8_t s_cls[256];
  __shared__ uint32_t s_wcnt[LEX_CT_WAVES];
  __shared__ int32_t s_wlast[LEX_CT_WAVES];
  __shared__ uint32_t s_ctile;
  __shared__ uint32_t s_stage[LEX_CT_WAVES][3][LEX_STAGE];
  const uint32_t C = p.C;
  for (uint32_t i = threadIdx.x; i < p.m1 * C; i += blockDim.x) s_dfa[i] = p.dfa_step[i];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_cls[i] = p.byte_class[i];
  if (threadIdx.x == 0) s_ctile = atomicAdd(p.tile_counter, 1u);
  __syncthreads();
  const uint32_t ct = s_ctile;
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  const uint32_t tile = ct * LEX_CT_WAVES + w;
  const uint64_t n = p.n;
  const uint32_t ign = p.ignore < 0 ? 0xFFFFFFFFu : (uint32_t)p.ignore;
  const uint64_t tbase = (uint64_t)tile * LEX_TILE;
  const uint64_t base = tbase + (uint64_t)lane * LEX_IPT;
  const int valid = (tile >= p.num_tiles || base >= n) ? 0 : (int)min((uint64_t)LEX_IPT, n - base);
  uint32_t cls[LEX_IPT / 4];
  load_classes(p, s_cls, base, valid, cls);
  uint32_t s_lane = 0;
  if (tile < p.num_tiles) {
    const uint32_t s_tile = p.tile_state[tile];
    const uint32_t ex = p.lane_excl[(uint64_t)tile * WAVE + lane];
    s_lane = p.endo_apply[ex * p.m1 + s_tile];
  }
  uint32_t c_next = __shfl_down(LEX_CLS(0), 1, WAVE);
  if (lane == WAVE - 1) {
    const uint64_t nb = tbase + LEX_TILE;
    c_next = nb < n ? s_cls[p.in[nb]] : 0u;
  }

  // ---- one walk: token-end mask, token-start mask and the label after every byte
  uint32_t endm = 0, startm = 0;
  uint32_t labs[LEX_IPT / 4] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t s_end = s_lane;
  uint32_t lab_end = 0;
  if (valid > 0) {
    uint32_t ent = s_dfa[s_lane * C + LEX_CLS(0)];
    uint32_t s = ent & 0xFFu;
    uint32_t lab = ent >> 9;
    if (((ent >> 8) & 1u) && lab != ign) startm |= 1u;
#pragma unroll
    for (int j = 0; j < LEX_IPT; ++j) {
      if (j < valid) {
        labs[j >> 2] |= (lab & 0xFFu) << (8 * (j & 3));
        const bool inner = j + 1 < valid;
        const bool has_next = inner || (base + j + 1 < n);
        const uint32_t cn = inner ? LEX_CLS((j + 1) & (LEX_IPT - 1)) : c_next;
        const uint32_t ent2 = has_next ? s_dfa[s * C + cn] : 0u;
        const bool prod2 = (ent2 >> 8) & 1u;
        const uint32_t lab2 = ent2 >> 9;
        if (has_next && prod2 && lab != ign) endm |= 1u << j;
        if (inner) {
          if (prod2 && lab2 != ign) startm |= 1u << (j + 1);
          s = ent2 & 0xFFu;
          lab = lab2;
        }
      }
    }
    s_end = s;
    lab_end = lab;
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
  uint32_t cincl = cnt;
  int32_t lincl = last;
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const uint32_t oc = __shfl_up(cincl, d, WAVE);
    const int32_t ol = __shfl_up(lincl, d, WAVE);
    if (lane >= d) {
      cincl += oc;
      lincl = max(lincl, ol);
    }
  }
  if (lane == WAVE - 1) {
    s_wcnt[w] = cincl;
    s_wlast[w] = lincl;
  }
  __syncthreads();
  if (w == 0) {
    uint32_t wc = lane < LEX_CT_WAVES ? s_wcnt[lane] : 0u;
    int32_t wl = lane < LEX_CT_WAVES ? s_wlast[lane] : -1;
    uint32_t wci = wc;
    int32_t wli = wl;
#pragma unroll
    for (int d = 1; d < LEX_CT_WAVES; d <<= 1) {
      const uint32_t oc = __shfl_up(wci, d, WAVE);
      const int32_t ol = __shfl_up(wli, d, WAVE);
      if (lane >= d) {
        wci += oc;
        wli = max(wli, ol);
      }
    }
    const uint32_t ctot = __shfl(wci, LEX_CT_WAVES - 1, WAVE);
    const int32_t ltot = __shfl(wli, LEX_CT_WAVES - 1, WAVE);
    if (lane == 0) st_relaxed64(p.st_count + ct, pack_count(ctot, ltot, ct == 0 ? CT_INC : CT_AGG));
    uint32_t cpre = 0;
    int32_t lpre = -1;
    if (ct > 0) {
      lookback_count(p, ct, lane, cpre, lpre);
      if (lane == 0) st_relaxed64(p.st_count + ct, pack_count(cpre + ctot, max(lpre, ltot), CT_INC));
    }
    int32_t wlex = __shfl_up(wli, 1, WAVE);
    const uint32_t wcex = wci - wc;
    if (lane == 0) wlex = -1;
    if (lane < LEX_CT_WAVES) {
      s_wcnt[lane] = cpre + wcex;
      s_wlast[lane] = max(lpre, wlex);
    }
  }
  __syncthreads();

  // ---- write tokens: stage in LDS by wave-local index, copy out coalesced, in rounds
  const uint32_t wout = s_wcnt[w];                  // global index of the wave's first token
  const uint32_t wtotal = __shfl(cincl, WAVE - 1, WAVE);
  const uint32_t lofs = cincl - cnt;                // wave-local index of this lane's first token
  int32_t lexcl = __shfl_up(lincl, 1, WAVE);
  if (lane == 0) lexcl = -1;
  const int32_t start_in = max(max(s_wlast[w], lexcl), 0);
  uint32_t *stk = s_stage[w][0];
  uint32_t *sts = s_stage[w][1];
  uint32_t *ste = s_stage[w][2];
  for (uint32_t r0 = 0; r0 < wtotal; r0 += LEX_STAGE) {
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
      const int32_t cur = sm ? (int32_t)(base + 31 - __clz(sm)) : start_in;
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
      stk[o] = (lw >> (8 * (j & 3))) & 0xFFu;
      sts[o] = (uint32_t)cur;
      ste[o] = (uint32_t)(base + j + (CHUNKED ? 0 : 1));
      ++k;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t m = min((uint32_t)LEX_STAGE, wtotal - r0);
    for (uint32_t i = lane; i < m; i += WAVE) {
      const uint64_t g = (uint64_t)wout + r0 + i;
      p.kinds[g] = stk[i];
      if (CHUNKED) {
        uint64_t cs, ce;
        chunk_span(p, sts[i], ste[i], cs, ce);
        p.starts64[g] = cs;
        p.ends64[g] = ce;
      } else {
        p.starts[g] = sts[i];
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
      p.kinds[out] = lab_end;
      p.starts[out] = (uint32_t)cur;
      p.ends[out] = (uint32_t)n;
      total = out + 1;
    }
    p.result[1] = total;
    p.result[0] = p.state_accept[s_end] ? 1u : 0u;
  }
}

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

__global__ void clear_u32(uint32_t *p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// ----------------------------------------------------------------------------------------------
// device-wide exclusive/inclusive scans of int2 (two independent int32 sums), reduce-then-scan.
// Loads are coalesced: a wavefront owns 16 rows of 64 consecutive elements. Values come from a
// source functor and results go to a sink functor, so producers and consumers fuse into the scan.

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ROWS = 8;
constexpr int SCAN_WAVE_ITEMS = WAVE * SCAN_ROWS;
constexpr int SCAN_TILE = (SCAN_BLOCK / WAVE) * SCAN_WAVE_ITEMS;

struct I2 {
  int32_t x, y;
};

__device__ __forceinline__ I2 add2(I2 a, I2 b) { return I2{a.x + b.x, a.y + b.y}; }

__device__ __forceinline__ I2 wave_sum2(I2 v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    v.x += __shfl_xor(v.x, d, WAVE);
    v.y += __shfl_xor(v.y, d, WAVE);
  }
  return v;
}

__device__ __forceinline__ I2 wave_incl2(I2 v, int lane) {
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const int32_t ox = __shfl_up(v.x, d, WAVE);
    const int32_t oy = __shfl_up(v.y, d, WAVE);
    if (lane >= d) {
      v.x += ox;
      v.y += oy;
    }
  }
  return v;
}

struct SrcArray {
  const I2 *a;
  __device__ I2 operator()(uint64_t i) const { return a[i]; }
};

// +1 for a left (pushed) bracket, -1 for a right (popped) one (depths, parser.fut:153-166)
struct SrcBracket {
  const uint16_t *br;
  __device__ I2 operator()(uint64_t i) const { return I2{(br[i] & 0x8000u) ? 1 : -1, 0}; }
};

// (arity - 1, is terminal production) (parents, parser.fut:286-296; terminal_offsets :305-314)
struct SrcProd {
  const uint16_t *prod;
  const int32_t *arity;
  const int32_t *p2t;
  __device__ I2 operator()(uint64_t i) const {
    const uint16_t p = prod[i];
    return I2{arity[p] - 1, p2t[p] >= 0 ? 1 : 0};
  }
};

struct SinkArray {  // exclusive prefix to an array
  I2 *out;
  __device__ void operator()(uint64_t i, bool valid, I2 v, I2 ex, I2 in, int lane) const {
    if (valid) out[i] = ex;
  }
};

// bracket depth = inclusive scan - [left]; reject on a negative prefix or a non-zero total;
// 64-element minima of the depths (level 1 of the min hierarchy).
struct SinkDepth {
  int32_t *depth;
  int32_t *l1;
  uint32_t *flags;
  uint64_t n;
  __device__ void operator()(uint64_t i, bool valid, I2 v, I2 ex, I2 in, int lane) const {
    int32_t d = INT32_MAX;
    if (valid) {
      if (in.x < 0 || (i == n - 1 && in.x != 0)) atomicOr(flags, 1u);
      d = in.x - (v.x > 0 ? 1 : 0);
      depth[i] = d;
    }
    const int32_t m = wave_min_i32(d);
    if (lane == 0 && i < n) l1[i / WAVE] = m;
  }
};

// s = exscan(arity - 1), terminal rank = exscan(is terminal); minima of s.
struct SinkProd {
  int32_t *s;
  int32_t *rank;
  int32_t *l1;
  uint64_t n;
  __device__ void operator()(uint64_t i, bool valid, I2 v, I2 ex, I2 in, int lane) const {
    int32_t x = INT32_MAX;
    if (valid) {
      x = ex.x;
      s[i] = x;
      rank[i] = ex.y;
    }
    const int32_t m = wave_min_i32(x);
    if (lane == 0 && i < n) l1[i / WAVE] = m;
  }
};

template <class Src>
__global__ void __launch_bounds__(SCAN_BLOCK) scan_reduce(Src src, uint64_t n, I2 *bsum) {
  __shared__ I2 wsum[SCAN_BLOCK / WAVE];
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  I2 s{0, 0};
#pragma unroll 4
  for (int r = 0; r < SCAN_TILE / SCAN_BLOCK; ++r) {
    const uint64_t i = base + (uint64_t)r * SCAN_BLOCK + threadIdx.x;
    if (i < n) s = add2(s, src(i));
  }
  s = wave_sum2(s);
  if ((threadIdx.x & (WAVE - 1)) == 0) wsum[threadIdx.x / WAVE] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    I2 t{0, 0};
    for (int w = 0; w < SCAN_BLOCK / WAVE; ++w) t = add2(t, wsum[w]);
    bsum[blockIdx.x] = t;
  }
}

// one block of 1024 threads: exclusive scan of the block sums in place (each thread owns a
// contiguous run); the grand total goes to bsum[nb]
constexpr int SCANB_THREADS = 1024;
__global__ void __launch_bounds__(SCANB_THREADS) scan_blocks(I2 *bsum, uint64_t nb) {
  __shared__ I2 wsum[SCANB_THREADS / WAVE];
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  const uint64_t per = (nb + SCANB_THREADS - 1) / SCANB_THREADS;
  const uint64_t lo = min(nb, (uint64_t)threadIdx.x * per), hi = min(nb, lo + per);
  I2 t{0, 0};
  for (uint64_t i = lo; i < hi; ++i) t = add2(t, bsum[i]);
  const I2 incl = wave_incl2(t, lane);
  if (lane == WAVE - 1) wsum[w] = incl;
  __syncthreads();
  I2 run{0, 0}, tot{0, 0};
  for (int k = 0; k < SCANB_THREADS / WAVE; ++k) {
    if (k < w) run = add2(run, wsum[k]);
    tot = add2(tot, wsum[k]);
  }
  run = add2(run, incl);
  run.x -= t.x;
  run.y -= t.y;
  for (uint64_t i = lo; i < hi; ++i) {
    const I2 v = bsum[i];
    bsum[i] = run;
    run = add2(run, v);
  }
  if (threadIdx.x == 0) bsum[nb] = tot;
}

template <class Src, class Sink>
__global__ void __launch_bounds__(SCAN_BLOCK) scan_down(Src src, uint64_t n, const I2 *bsum,
                                                        Sink sink) {
  __shared__ I2 wsum[SCAN_BLOCK / WAVE];
  const int lane = threadIdx.x & (WAVE - 1);
  const int w = threadIdx.x / WAVE;
  const uint64_t wbase = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)w * SCAN_WAVE_ITEMS;
  I2 vals[SCAN_ROWS];
  I2 tot{0, 0};
#pragma unroll
  for (int r = 0; r < SCAN_ROWS; ++r) {
    const uint64_t i = wbase + (uint64_t)r * WAVE + lane;
    vals[r] = i < n ? src(i) : I2{0, 0};
    tot = add2(tot, vals[r]);
  }
  tot = wave_sum2(tot);
  if (lane == 0) wsum[w] = tot;
  __syncthreads();
  I2 carry = bsum[blockIdx.x];
  for (int k = 0; k < w; ++k) carry = add2(carry, wsum[k]);
#pragma unroll
  for (int r = 0; r < SCAN_ROWS; ++r) {
    const uint64_t i = wbase + (uint64_t)r * WAVE + lane;
    const I2 incl_row = wave_incl2(vals[r], lane);
    const I2 in = add2(carry, incl_row);
    I2 ex = in;
    ex.x -= vals[r].x;
    ex.y -= vals[r].y;
    sink(i, i < n, vals[r], ex, in, lane);
    carry.x += __shfl(incl_row.x, WAVE - 1, WAVE);
    carry.y += __shfl(incl_row.y, WAVE - 1, WAVE);
  }
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

// minimum of every 64-element chunk: one wavefront per chunk, coalesced loads
__global__ void min_reduce64(const int32_t *in, uint64_t n, int32_t *out, uint64_t nout) {
  const int lane = threadIdx.x & (WAVE - 1);
  const uint64_t wstride = (uint64_t)gridDim.x * (blockDim.x / WAVE);
  for (uint64_t c = (uint64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; c < nout;
       c += wstride) {
    const uint64_t i = c * WAVE + lane;
    const int32_t m = wave_min_i32(i < n ? in[i] : INT32_MAX);
    if (lane == 0) out[c] = m;
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
__device__ __forceinline__ int lift_left(int32_t x, int pos, int32_t v) {
  int32_t M[6];
  M[0] = x;
#pragma unroll
  for (int k = 1; k < 6; ++k) {
    const int32_t o = __shfl_down(M[k - 1], 1 << (k - 1), WAVE);
    const int lane = threadIdx.x & (WAVE - 1);
    M[k] = (lane + (1 << (k - 1)) < WAVE) ? min(M[k - 1], o) : M[k - 1];
  }
#pragma unroll
  for (int k = 5; k >= 0; --k) {
    const int step = 1 << k;
    const int src = pos - step;
    // M[k] at lane src covers x_src .. x_{src+step-1} = x_{pos-step} .. x_{pos-1}
    const int32_t m = __shfl(M[k], src < 0 ? 0 : src, WAVE);
    if (src >= 0 && m > v) pos = src;
  }
  // the steps add up to 63: from pos = 64 the loop stops at 1 without having looked at x_0
  const int32_t x0 = __builtin_amdgcn_readlane(x, 0);
  if (pos == 1 && x0 > v) pos = 0;
  return pos;
}

// Minimum of the 64 chunks before `chunk` (lane l: chunk - 64 + l), the window the far search uses.
__device__ __forceinline__ int32_t window_min(const Levels &lv, uint64_t chunk, int lane) {
  if (lv.nlev < 2) return INT32_MAX;
  const int64_t wc = (int64_t)chunk - WAVE + lane;
  return wc >= 0 ? lv.L[1][wc] : INT32_MAX;
}

// Wave-cooperative previous smaller-or-equal for the 64 consecutive elements of chunk `chunk`
// (lane l holds element chunk*64+l with value v; lanes past the end hold INT32_MAX; wv is
// window_min of the chunk): binary lifting inside the chunk; then binary lifting over the window
// picks the chunk holding the answer, and each distinct such chunk is loaded once (coalesced)
// and searched by lifting again; a per-lane climb of the min hierarchy handles the rest.
// All 64 lanes must call this together.
__device__ int64_t prev_le_wave(const Levels &lv, uint64_t chunk, int lane, int32_t v, int32_t wv,
                                bool need) {
  const int64_t cbase = (int64_t)(chunk * WAVE);
  const int p0 = lift_left(v, lane, v);
  int64_t res = -1;
  bool done = !need;
  if (!done && p0 > 0) {
    res = cbase + p0 - 1;
    done = true;
  }
  if (__ballot(!done) == 0ull) return res;
  if (lv.nlev >= 2) {
    const int p1 = lift_left(wv, WAVE, v);
    // target chunk of every unresolved lane (-1: not in the window)
    int64_t cb = (!done && p1 > 0) ? (int64_t)chunk - WAVE + p1 - 1 : -1;
    uint64_t pend = __ballot(cb >= 0);
    while (pend) {
      const int t = __ffsll((long long)pend) - 1;
      const int64_t c = (int64_t)(uint32_t)__builtin_amdgcn_readlane((int)(cb & 0xFFFFFFFF), t) |
                        ((int64_t)__builtin_amdgcn_readlane((int)(cb >> 32), t) << 32);
      const int32_t x = lv.L[0][c * WAVE + lane];
      const bool mine = cb == c;
      const int p2 = lift_left(x, WAVE, mine ? v : INT32_MIN);
      if (mine) {
        res = c * WAVE + p2 - 1;
        done = true;
        cb = -1;
      }
      pend = __ballot(cb >= 0);
    }
  }
  if (!done) res = prev_le_far(lv, chunk > (uint64_t)WAVE ? chunk - WAVE : 0, v);
  return res;
}

// ----------------------------------------------------------------------------------------------
// parser kernels

struct ParseArgs {
  const uint32_t *kinds;
  uint64_t T;  // tokens
  uint32_t q, k, nT, start_t, end_t, empty_t;
  const uint32_t *pair_index;
  const uint32_t *pair_entry;
  const uint16_t *stacks;
  const uint16_t *prods;
  const int32_t *arity;
  const int32_t *p2t;
  uint32_t *flags;  // [0] reject
};

// keys over [start] ++ kinds ++ [end] (get_key / lookup, parser.fut:117-151)
__global__ void parse_keys(ParseArgs a, uint32_t *ent, I2 *lens) {
  const uint64_t P = a.T + 2;
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < P;
       p += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t di = 0;
    for (uint32_t j = 0; j < a.q + a.k; ++j) {
      const int64_t pos = (int64_t)p + (int64_t)j - (int64_t)a.q;
      uint32_t t;
      if (pos < 0 || pos >= (int64_t)P) t = a.empty_t;
      else if (pos == 0) t = a.start_t;
      else if (pos == (int64_t)P - 1) t = a.end_t;
      else t = a.kinds[pos - 1];
      di = di * a.nT + (t < a.nT ? t : a.empty_t);
    }
    const uint32_t id = a.pair_index[di];
    if (id == 0xFFFFFFFFu) {
      atomicOr(a.flags, 1u);
      ent[p] = 0;
      lens[p] = I2{0, 0};
    } else {
      ent[p] = id;
      const uint32_t *e4 = a.pair_entry + 4ull * id;
      lens[p] = I2{(int32_t)(e4[1] - e4[0]), (int32_t)(e4[3] - e4[2])};
    }
  }
}

// segmented copy of stack and production spans (segmented_copy, parser.fut:216-233)
__global__ void parse_expand(ParseArgs a, const uint32_t *ent, const I2 *offs, uint16_t *br,
                             uint16_t *prod) {
  const uint64_t P = a.T + 2;
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < P;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t *e4 = a.pair_entry + 4ull * ent[p];
    const I2 o = offs[p];
    for (uint32_t s = e4[0]; s < e4[1]; ++s) br[(uint64_t)o.x + (s - e4[0])] = a.stacks[s];
    for (uint32_t s = e4[2]; s < e4[3]; ++s) prod[(uint64_t)o.y + (s - e4[2])] = a.prods[s];
  }
}

// brackets_matches (parser.fut:174-184): every right bracket equals (without the left bit) the
// bracket at its previous smaller-or-equal depth. One wavefront per 64 brackets.
__global__ void bracket_match(const uint16_t *br, uint64_t B, Levels lv, uint32_t *flags) {
  const int lane = threadIdx.x & (WAVE - 1);
  const uint64_t nch = (B + WAVE - 1) / WAVE;
  const uint64_t wstride = (uint64_t)gridDim.x * (blockDim.x / WAVE);
  for (uint64_t c = (uint64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; c < nch;
       c += wstride) {
    const uint64_t i = c * WAVE + lane;
    const bool valid = i < B;
    const uint16_t b = valid ? br[i] : (uint16_t)0x8000u;
    const int32_t v = valid ? lv.L[0][i] : INT32_MAX;
    const int32_t wv = window_min(lv, c, lane);
    const bool right = valid && !(b & 0x8000u);
    const int64_t j = prev_le_wave(lv, c, lane, v, wv, right);
    if (right && (j < 0 || (br[j] & 0x7FFFu) != (b & 0x7FFFu))) atomicOr(flags, 1u);
  }
}

// parents (parser.fut:286-296) and terminal nodes (parser.fut:305-332). One wavefront per chunk.
// S: span word (u32, or u64 for the chunked lexer's spans).
// subst = 0 when the terminal productions do not pair up with the tokens: safe_zip then yields no
// pairs (parser.fut:300-303) and every node stays a production node.
template <typename S>
__global__ void parse_nodes(ParseArgs a, const uint16_t *prod, const int32_t *rank, uint64_t N,
                            Levels lv, const S *starts, const S *ends, uint32_t *parent,
                            uint32_t *node, S *span, int subst) {
  const int lane = threadIdx.x & (WAVE - 1);
  const uint64_t nch = (N + WAVE - 1) / WAVE;
  const uint64_t wstride = (uint64_t)gridDim.x * (blockDim.x / WAVE);
  for (uint64_t c = (uint64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; c < nch;
       c += wstride) {
    const uint64_t i = c * WAVE + lane;
    const bool valid = i < N;
    const int32_t v = valid ? lv.L[0][i] : INT32_MAX;
    const int32_t wv = window_min(lv, c, lane);
    // loads that do not depend on the search go out first
    const uint16_t pr = valid ? prod[i] : (uint16_t)0;
    const int32_t t = valid && subst ? a.p2t[pr] : -1;
    S s0 = 0, s1 = 0;
    if (t >= 0) {
      const uint32_t r = (uint32_t)rank[i];
      s0 = starts[r];
      s1 = ends[r];
    }
    const int64_t j = prev_le_wave(lv, c, lane, v, wv, valid && i > 0);
    if (!valid) continue;
    parent[i] = i == 0 ? 0u : (uint32_t)j;
    node[i] = t >= 0 ? (ALPACC_NODE_TERMINAL | (uint32_t)t) : (uint32_t)pr;
    if (sizeof(S) == 4) {
      reinterpret_cast<uint2 *>(span)[i] = make_uint2((uint32_t)s0, (uint32_t)s1);
    } else {
      span[2 * i] = s0;
      span[2 * i + 1] = s1;
    }
  }
}

inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 1u << 16) {
  uint64_t g = (n + block - 1) / block;
  if (g == 0) g = 1;
  return (unsigned)std::min<uint64_t>(g, cap);
}

}  // namespace

// ----------------------------------------------------------------------------------------------
// host side

static std::string g_create_error;

enum Stage { S_LEX_REDUCE = 0, S_LEX_SCAN, S_LEX_EMIT, S_PARSE, NSTAGE };
static const char *kStageNames[NSTAGE] = {"lex_reduce", "lex_scan", "lex_emit", "parse"};

struct alpacc_ctx {
  int device = 0;
  int cus = 0;
  std::string err;
  int launches = 0;
  // lexer tables
  bool has_lexer = false;
  uint32_t e = 0, C = 0, m1 = 0, init = 0, dead = 0;
  int32_t ignore = -1;
  uint8_t *d_cls = nullptr;
  uint16_t *d_step = nullptr, *d_comp = nullptr, *d_dfa = nullptr;
  uint8_t *d_apply = nullptr, *d_accept = nullptr;
  bool comp_in_lds = false;
  size_t lex_lds = 0;
  // parser tables
  bool has_parser = false;
  uint32_t q = 0, k = 0, nT = 0, start_t = 0, end_t = 0, empty_t = 0, P = 0;
  uint32_t *d_pair_index = nullptr, *d_pair_entry = nullptr;
  uint16_t *d_stacks = nullptr, *d_prods = nullptr;
  int32_t *d_arity = nullptr, *d_p2t = nullptr;
  // scratch (grow-only)
  std::vector<std::pair<void *, size_t>> scratch;
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
  void prof_drop(int s) { rec[s] = false; }
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
  B_LENS,
  B_OFFS,
  B_BSUM,
  B_BR,
  B_PROD,
  B_DEPTH,
  B_LEVELS,
  B_RANK,
  B_SX,
  B_DL1,
  B_SL1,
  B_VIRT,
  B_CHUNK_INFO,
  B_TOK_START64,
  B_TOK_END64,
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
template <class Src, class Sink>
int scan2(alpacc_ctx *ctx, Src src, Sink sink, uint64_t n, I2 *total, hipStream_t st) {
  const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  I2 *bsum = (I2 *)ctx->buf(B_BSUM, (nb + 1) * sizeof(I2));
  if (!bsum) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  if (n > 0) {
    scan_reduce<Src><<<(unsigned)nb, SCAN_BLOCK, 0, st>>>(src, n, bsum);
    scan_blocks<<<1, SCANB_THREADS, 0, st>>>(bsum, nb);
    scan_down<Src, Sink><<<(unsigned)nb, SCAN_BLOCK, 0, st>>>(src, n, bsum, sink);
    ctx->launches += 3;
    CK(hipGetLastError());
  }
  if (total) {
    if (n == 0) {
      *total = I2{0, 0};
    } else {
      CK(hipMemcpyAsync(total, bsum + nb, sizeof(I2), hipMemcpyDeviceToHost, st));
      CK(hipStreamSynchronize(st));
    }
  }
  return 0;
}

// Level arrays of the min hierarchy over `base` (n values): level 1 must already be in
// `l1`; levels 2.. are built here. Every level is padded so that 16-byte quads can be read.
int build_levels(alpacc_ctx *ctx, int32_t *base, int32_t *l1, uint64_t n, Levels &lv,
                 hipStream_t st) {
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
  int32_t *area = (int32_t *)ctx->buf(B_LEVELS, (total + 16) * sizeof(int32_t));
  if (!area) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  int32_t *cur = l1;
  uint64_t off = 0;
  while (len > 64) {
    if (lv.nlev >= MAX_LEVELS) {
      ctx->err = "input too large for the min levels";
      return ALPACC_ERR_ARG;
    }
    const uint64_t nl = (len + 63) / 64;
    int32_t *next = area + off;
    min_reduce64<<<grid_for(nl * WAVE, 256), 256, 0, st>>>(cur, len, next, nl);
    ctx->launches++;
    lv.L[lv.nlev] = next;
    lv.len[lv.nlev] = nl;
    lv.nlev++;
    off += ((nl + 3) & ~3ull) + 16;
    cur = next;
    len = nl;
  }
  CK(hipGetLastError());
  return 0;
}

// The lexer. chunk == 0: unchunked (kinds/starts/ends as u32). chunk > 0: the reference's chunk
// loop with chunks of `chunk` bytes (kinds, starts64/ends64).
int run_lex(alpacc_ctx *ctx, const uint8_t *d_in, uint64_t n, uint32_t *kinds, uint32_t *starts,
            uint32_t *ends, uint64_t *num_tokens, hipStream_t st, uint64_t chunk = 0,
            uint64_t *starts64 = nullptr, uint64_t *ends64 = nullptr) {
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
  uint32_t *small = (uint32_t *)ctx->buf(B_SMALL, 64);
  uint16_t *lane_excl = (uint16_t *)ctx->buf(B_LANE_EXCL, tiles * WAVE * sizeof(uint16_t));
  uint16_t *tile_agg = (uint16_t *)ctx->buf(B_TILE_AGG, tiles * sizeof(uint16_t));
  uint16_t *tile_excl = (uint16_t *)ctx->buf(B_TILE_EXCL, tiles * sizeof(uint16_t));
  uint16_t *block_agg = (uint16_t *)ctx->buf(B_BLOCK_AGG, sblocks * sizeof(uint16_t));
  uint8_t *tile_state = (uint8_t *)ctx->buf(B_TILE_STATE, tiles);
  uint64_t *st_count = (uint64_t *)ctx->buf(B_ST_COUNT, ctiles * sizeof(uint64_t));
  if (!small || !lane_excl || !tile_agg || !tile_excl || !block_agg || !tile_state || !st_count) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  clear_u32<<<1, 64, 0, st>>>(small, 16);
  clear_u32<<<grid_for(2 * ctiles, 256), 256, 0, st>>>((uint32_t *)st_count, 2 * ctiles);
  LexArgs a;
  a.in = d_in;
  a.n = n;
  a.kinds = kinds;
  a.starts = starts;
  a.ends = ends;
  a.byte_class = ctx->d_cls;
  a.endo_step = ctx->d_step;
  a.endo_comp = ctx->d_comp;
  a.endo_apply = ctx->d_apply;
  a.dfa_step = ctx->d_dfa;
  a.state_accept = ctx->d_accept;
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
  const uint64_t wpb = LEX_BLOCK / WAVE;
  uint64_t blocks = (tiles + wpb - 1) / wpb;
  blocks = std::min<uint64_t>(blocks, (uint64_t)std::max(1, ctx->cus) * 2);
  ctx->prof_begin(S_LEX_REDUCE, st);
  if (ctx->comp_in_lds)
    lex_reduce<true><<<(unsigned)blocks, LEX_BLOCK, ctx->lex_lds, st>>>(a);
  else
    lex_reduce<false><<<(unsigned)blocks, LEX_BLOCK, ctx->lex_lds, st>>>(a);
  ctx->prof_end(S_LEX_REDUCE, st);
  ctx->prof_begin(S_LEX_SCAN, st);
  lex_scan_tiles<<<(unsigned)sblocks, LEX_SCAN_BLOCK, 0, st>>>(a);
  lex_scan_blocks<<<1, LEX_SCAN_BLOCK, 0, st>>>(a, (uint32_t)sblocks);
  lex_tile_states<<<grid_for(tiles, 256), 256, 0, st>>>(a);
  ctx->prof_end(S_LEX_SCAN, st);
  ctx->prof_begin(S_LEX_EMIT, st);
  if (chunk > 0) {
    lex_emit<true><<<(unsigned)ctiles, LEX_BLOCK, 0, st>>>(a);
    lex_chunk_fixup<<<grid_for(K, 256, 1024), 256, 0, st>>>(a);
    ctx->launches++;
  } else {
    lex_emit<false><<<(unsigned)ctiles, LEX_BLOCK, 0, st>>>(a);
  }
  ctx->prof_end(S_LEX_EMIT, st);
  ctx->launches += 7;
  CK(hipGetLastError());
  CK(hipMemcpyAsync(ctx->h_small, small + 4, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  if (ctx->h_small[2]) {
    ctx->err = "lexer look-back did not complete";
    return ALPACC_ERR_TIMEOUT;
  }
  *num_tokens = ctx->h_small[1];
  return ctx->h_small[0] ? ALPACC_ACCEPT : ALPACC_REJECT;
}

template <typename S>
int run_parse_tokens(alpacc_ctx *ctx, const uint32_t *kinds, const S *starts, const S *ends,
                     uint64_t T, uint32_t *parent, uint32_t *node, S *span, uint64_t cap,
                     uint64_t *num_nodes, hipStream_t st) {
  if (!ctx->has_parser) {
    ctx->err = "bundle has no parser";
    return ALPACC_ERR_NO_PARSER;
  }
  *num_nodes = 0;
  const uint64_t P = T + 2;
  uint32_t *flags = (uint32_t *)ctx->buf(B_SMALL, 64) + 8;
  uint32_t *ent = (uint32_t *)ctx->buf(B_ENT, P * sizeof(uint32_t));
  I2 *lens = (I2 *)ctx->buf(B_LENS, P * sizeof(I2));
  I2 *offs = (I2 *)ctx->buf(B_OFFS, P * sizeof(I2));
  if (!flags || !ent || !lens || !offs) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  ParseArgs a;
  a.kinds = kinds;
  a.T = T;
  a.q = ctx->q;
  a.k = ctx->k;
  a.nT = ctx->nT;
  a.start_t = ctx->start_t;
  a.end_t = ctx->end_t;
  a.empty_t = ctx->empty_t;
  a.pair_index = ctx->d_pair_index;
  a.pair_entry = ctx->d_pair_entry;
  a.stacks = ctx->d_stacks;
  a.prods = ctx->d_prods;
  a.arity = ctx->d_arity;
  a.p2t = ctx->d_p2t;
  a.flags = flags;
  ctx->prof_begin(S_PARSE, st);
  clear_u32<<<1, 64, 0, st>>>(flags, 4);
  parse_keys<<<grid_for(P, 256), 256, 0, st>>>(a, ent, lens);
  ctx->launches += 2;
  CK(hipGetLastError());
  I2 tot;
  int rc = scan2(ctx, SrcArray{lens}, SinkArray{offs}, P, &tot, st);
  if (rc) return rc;
  CK(hipMemcpyAsync(ctx->h_small, flags, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  if (ctx->h_small[0]) return ALPACC_REJECT;  // a missing pair (to_keys, parser.fut:235-240)
  const uint64_t B = (uint32_t)tot.x, N = (uint32_t)tot.y;
  uint16_t *br = (uint16_t *)ctx->buf(B_BR, B * sizeof(uint16_t) + 2);
  uint16_t *prod = (uint16_t *)ctx->buf(B_PROD, N * sizeof(uint16_t) + 2);
  if (!br || !prod) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  parse_expand<<<grid_for(P, 256), 256, 0, st>>>(a, ent, offs, br, prod);
  ctx->launches++;
  // bracket depths and matching
  if (B > 0) {
    int32_t *depth = (int32_t *)ctx->buf(B_DEPTH, (B + 16) * sizeof(int32_t));
    int32_t *dl1 = (int32_t *)ctx->buf(B_DL1, ((B + 63) / 64 + 16) * sizeof(int32_t));
    if (!depth || !dl1) {
      ctx->err = "scratch allocation failed";
      return ALPACC_ERR_HIP;
    }
    rc = scan2(ctx, SrcBracket{br}, SinkDepth{depth, dl1, flags, B}, B, nullptr, st);
    if (rc) return rc;
    Levels lv;
    rc = build_levels(ctx, depth, dl1, B, lv, st);
    if (rc) return rc;
    bracket_match<<<grid_for(B, 256), 256, 0, st>>>(br, B, lv, flags);
    ctx->launches++;
    CK(hipGetLastError());
  }
  CK(hipMemcpyAsync(ctx->h_small, flags, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  if (ctx->h_small[0]) return ALPACC_REJECT;
  *num_nodes = N;
  if (N > cap) {
    ctx->err = "output capacity too small";
    return ALPACC_ERR_CAPACITY;
  }
  if (N == 0) return ALPACC_ACCEPT;
  int32_t *sx = (int32_t *)ctx->buf(B_SX, (N + 16) * sizeof(int32_t));
  int32_t *rank = (int32_t *)ctx->buf(B_RANK, (N + 16) * sizeof(int32_t));
  int32_t *sl1 = (int32_t *)ctx->buf(B_SL1, ((N + 63) / 64 + 16) * sizeof(int32_t));
  if (!sx || !rank || !sl1) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  I2 ptot;
  rc = scan2(ctx, SrcProd{prod, ctx->d_arity, ctx->d_p2t}, SinkProd{sx, rank, sl1, N}, N, &ptot, st);
  if (rc) return rc;
  const int subst = (uint64_t)(uint32_t)ptot.y == T ? 1 : 0;
  Levels lv;
  rc = build_levels(ctx, sx, sl1, N, lv, st);
  if (rc) return rc;
  parse_nodes<S><<<grid_for(N, 256), 256, 0, st>>>(a, prod, rank, N, lv, starts, ends, parent,
                                                   node, span, subst);
  ctx->launches++;
  CK(hipGetLastError());
  ctx->prof_end(S_PARSE, st);
  CK(hipStreamSynchronize(st));
  return ALPACC_ACCEPT;
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
  hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device);
  std::string &err = g_create_error;
  if (hipHostMalloc((void **)&ctx->h_small, 64, hipHostMallocDefault) != hipSuccess) {
    err = "hipHostMalloc failed";
    delete ctx;
    return nullptr;
  }
  if (t->num_endos > 0) {
    if (t->num_endos > 65535 || t->num_states > 256 || t->num_classes > 256 ||
        t->init_state >= t->num_states) {
      err = "lexer tables out of range";
      delete ctx;
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
    ctx->d_dfa = upload(t->dfa_step, (size_t)ctx->m1 * ctx->C, err);
    ctx->d_accept = upload(t->state_accept, ctx->m1, err);
    if (!err.empty()) {
      delete ctx;
      return nullptr;
    }
    if ((size_t)ctx->m1 * ctx->C > (size_t)LEX_MAX_DFA) {
      err = "lexer DFA table too large";
      delete ctx;
      return nullptr;
    }
    size_t lds = (size_t)ctx->e * ctx->C * 2 + 256;
    lds = (lds + 15) & ~(size_t)15;
    const size_t comp_bytes = (size_t)ctx->e * ctx->e * 2;
    ctx->comp_in_lds = lds + comp_bytes <= 48 * 1024;
    if (ctx->comp_in_lds) lds += comp_bytes;
    if (lds > 160 * 1024) {
      err = "lexer tables do not fit in LDS";
      delete ctx;
      return nullptr;
    }
    if (lds > 64 * 1024) {
      const int want = (int)lds;
      if (hipFuncSetAttribute((const void *)lex_reduce<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, want) != hipSuccess) {
        err = "lexer tables do not fit in LDS";
        delete ctx;
        return nullptr;
      }
    }
    ctx->lex_lds = lds;
  }
  if (t->num_productions > 0 && t->pair_index) {
    ctx->has_parser = true;
    ctx->q = t->q;
    ctx->k = t->k;
    ctx->nT = t->num_terminals;
    ctx->start_t = t->start_terminal;
    ctx->end_t = t->end_terminal;
    ctx->empty_t = t->empty_terminal;
    ctx->P = t->num_productions;
    ctx->d_pair_index = upload(t->pair_index, t->pair_index_len, err);
    ctx->d_pair_entry = upload(t->pair_entry, (size_t)t->num_pair_entries * 4, err);
    ctx->d_stacks = upload(t->stacks, t->stacks_len, err);
    ctx->d_prods = upload(t->productions, t->productions_len, err);
    ctx->d_arity = upload(t->arities, t->num_productions, err);
    ctx->d_p2t = upload(t->prod_to_terminal, t->num_productions, err);
    if (!err.empty()) {
      delete ctx;
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
  void *tabs[] = {ctx->d_cls,          ctx->d_step,         ctx->d_comp,   ctx->d_apply,
                  ctx->d_dfa,          ctx->d_accept,       ctx->d_pair_index, ctx->d_pair_entry,
                  ctx->d_stacks,       ctx->d_prods,        ctx->d_arity,  ctx->d_p2t};
  for (void *p : tabs)
    if (p) (void)hipFree(p);
  if (ctx->h_small) (void)hipHostFree(ctx->h_small);
  for (auto &pair : ctx->ev)
    for (auto &e : pair)
      if (e) (void)hipEventDestroy(e);
  delete ctx;
}

const char *alpacc_last_error(alpacc_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int alpacc_last_launch_count(alpacc_ctx *ctx) { return ctx ? ctx->launches : 0; }

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
  uint32_t *kinds = (uint32_t *)ctx->buf(B_TOK_KIND, tcap * 4);
  uint32_t *starts = (uint32_t *)ctx->buf(B_TOK_START, tcap * 4);
  uint32_t *ends = (uint32_t *)ctx->buf(B_TOK_END, tcap * 4);
  if (!kinds || !starts || !ends) {
    ctx->err = "scratch allocation failed";
    return ALPACC_ERR_HIP;
  }
  uint64_t T = 0;
  int rc = run_lex(ctx, d_in, n, kinds, starts, ends, &T, st);
  if (rc != ALPACC_ACCEPT) {
    if (ctx->prof) ctx->prof_collect();
    return rc;
  }
  const int lex_launches = ctx->launches;
  rc = run_parse_tokens<uint32_t>(ctx, kinds, starts, ends, T, d_parent, d_node, d_span, cap,
                                  num_nodes, st);
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

float alpacc_event_elapsed_ms(void *start, void *end) {
  if (hipEventSynchronize((hipEvent_t)end) != hipSuccess) return -1.0f;
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, (hipEvent_t)start, (hipEvent_t)end) != hipSuccess) return -1.0f;
  return ms;
}

}  // extern "C"
