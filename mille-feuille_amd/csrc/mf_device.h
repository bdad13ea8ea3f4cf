// mf_device.h -- device-side building blocks shared by the SpMV and solver
// kernels: the streamed wave64 tiled SpMV, deterministic block / grid
// reductions and the grid barrier.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mf_internal.h"
#include "mf_mlp.h"

#define MF_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return mf::fail(MF_ERR_HIP, std::string(#call) + ": " +               \
                                      hipGetErrorString(e_));                \
  } while (0)

#define MF_TRY(x)               \
  do {                          \
    mf_status s_ = (x);         \
    if (s_ != MF_OK) return s_; \
  } while (0)

namespace mf {

constexpr int kWave = 64;

// Pointers of the device layout passed by value into kernels.
struct TileView {
  const int *tile_ptr;  // tilem+1
  const int *tile_col;  // tilenum
  const int *tile_off;  // tilenum+1
  const int *tile_row;  // tilenum
  const uint8_t *tile_nz;  // tilenum: nnz of the tile - 1
  const uint8_t *rc;    // payload slots: (row << 4) | col inside the tile
  const uint8_t *rp16;  // 16 per tile: row starts inside the tile
  const double *v64;
  const float *v32;
  const uint16_t *v16;
  const int8_t *v8;
  const float *s16;
  const float *s8;
  const uint8_t *prec;
  int tilem;
  int tilenum;
  int npad;
};

inline TileView view_of(const mf_tiled *t) {
  const DeviceLayout &d = t->dev;
  return TileView{d.tile_ptr, d.tile_col, d.tile_off, d.tile_row, d.tile_nz, d.rc, d.rp16,
                  d.v64,      d.v32,      d.v16,      d.v8,       d.s16,  d.s8,
                  d.prec,     t->tilem,   t->tilenum, d.npad};
}

__device__ __forceinline__ float half_bits_to_float(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}

// SpMV modes: every tile FP64; per-tile levels (A.prec); per-segment levels
// (seglev[tile column], the rule policy's flag indexed by the 16-entry segment
// of x that the tile multiplies, src/main-cg-mixed.cu:236, 1134).
enum SpmvMode { kAll64 = 0, kPerTile = 1, kPerSeg = 2 };

// Per-lane counters of tile visits per precision level.
struct PrecCount {
  unsigned c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0;
  __device__ __forceinline__ void add(int p) {
    c0 += p == 0;
    c1 += p == 1;
    c2 += p == 2;
    c3 += p == 3;
    c4 += p == 4;
  }
};

// ---------------------------------------------------------------------------
// Streamed tiled SpMV.
//
// A wave owns a contiguous range of tile-rows, i.e. a contiguous range of
// tiles and of payload slots (tiles are stored tile-row by tile-row, each
// tile's entries padded to a multiple of 4).  It walks that range in windows
// of at most kWinTiles tiles, kWinRows tile-rows and kWinChunks 4-entry
// chunks.  Per window the tile metadata (payload offset, column, tile-row,
// entry count, precision, scale) and the tile of every chunk are staged in
// wave-private LDS, then the window's chunks are streamed: lane L takes chunk
// base/4 + L, so one wave instruction covers 256 payload slots and the
// values / (row,col) bytes arrive as 32-, 16-, 8- or 4-byte vector loads.  A
// chunk never straddles two tiles, so one lookup gives its tile, level,
// column and row base; the lane gathers x, and adds its products into per-row
// FP64 accumulators in LDS, merging consecutive entries of one row in
// registers first.  When a window ends, every tile-row it completed is handed
// to the epilogue row by row (one lane per row); an unfinished last tile-row
// keeps its accumulators for the next window.
//
// Mixed precision: FP64 tiles multiply x; FP32 / FP16 / INT8 tiles multiply
// the FP32 copy x32 with FP32 products (FP16/INT8 scaled by the tile's power
// of two) that are accumulated in FP64; SKIP tiles contribute nothing.
// force64 bit 0 treats every tile as FP64 (the solver's fallback); bit 1
// (per-tile levels) makes FP64 tiles multiply the FP32 copy x32 as well, in
// FP64 arithmetic (the mixed solvers keep their vectors in FP32 only).
// ---------------------------------------------------------------------------

struct WaveLds {
  int tcol[kWinTiles];      // 16 * tile column
  float tscale[kWinTiles];  // FP16 / INT8 scale of the tile's current level
  int tend[kWinTiles];      // end of the tile's real entries, relative to the window
  uint8_t trow[kWinTiles];  // tile-row relative to the window's first tile-row
  uint8_t tprec[kWinTiles];
  uint8_t ctile[kWinChunks];  // window-relative tile of each 4-entry chunk
  static constexpr int kAccStride = 16;  // accumulators per tile-row
  double acc[kWinRows * 16];
};

template <int MODE>
__device__ __forceinline__ int tile_level(const TileView &A, const uint8_t *seglev, int t,
                                          int force64) {
  if (MODE == kAll64 || (force64 & 1)) return MF_PREC_FP64;
  if (MODE == kPerTile) return A.prec[t];
  return seglev[A.tile_col[t]];
}

// Product of one entry at its tile's level (FP64 result).
template <int MODE>
__device__ __forceinline__ double entry_product(const TileView &A, int lvl, int k,
                                                int xi, const double *__restrict__ x,
                                                const float *__restrict__ x32, float scale) {
  if (lvl == MF_PREC_FP64) return A.v64[k] * x[xi];
  if (lvl == MF_PREC_FP32) return (double)(A.v32[k] * x32[xi]);
  if (lvl == MF_PREC_FP16) return (double)(half_bits_to_float(A.v16[k]) * x32[xi] * scale);
  if (lvl == MF_PREC_INT8) return (double)((float)A.v8[k] * x32[xi] * scale);
  return 0.0;
}

// One lane's chunk of a slice: the four payload slots k0..k0+3 of the window,
// all in tile j (tiles start at multiples of 4); slots past the tile's last
// entry are padding.  The payload is loaded at the tile's level (FP64: 8
// words, FP32: 4, FP16: 2, INT8: 1, packed); x is gathered once the (row,col)
// bytes are in.
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

struct Chunk {
  uint32_t rc4;
  int vmask;      // bit u: slot u holds an entry of the window; 0: none
  int j;          // window-relative tile
  int lvl;        // level of tile j
  int wsel;       // fixed-width loads: FP16 / INT8 chunk's word in a (see chunk_load_pay_fixed)
  u32x4_t a, b;   // raw payload words 0-3 and 4-7
  double x0, x1, x2, x3;  // x (FP64 tiles) or x32 (other levels) at each entry's column
};

// Word i of the payload (i is a compile-time constant after unrolling).
__device__ __forceinline__ uint32_t chunk_word(const Chunk &c, int i) {
  switch (i) {
    case 0: return c.a.x;
    case 1: return c.a.y;
    case 2: return c.a.z;
    case 3: return c.a.w;
    case 4: return c.b.x;
    case 5: return c.b.y;
    case 6: return c.b.z;
    default: return c.b.w;
  }
}

// x of entry u (u is a compile-time constant after unrolling; named members
// rather than an array keep the values in registers).
__device__ __forceinline__ double chunk_x(const Chunk &c, int u) {
  switch (u) {
    case 0: return c.x0;
    case 1: return c.x1;
    case 2: return c.x2;
    default: return c.x3;
  }
}

__device__ __forceinline__ void chunk_set_x(Chunk &c, int u, double v) {
  switch (u) {
    case 0: c.x0 = v; break;
    case 1: c.x1 = v; break;
    case 2: c.x2 = v; break;
    default: c.x3 = v; break;
  }
}

// Stage 1a: the (row,col) bytes, and the payload when every tile is FP64;
// neither needs the window metadata.  c = chunk index within the window.
__device__ __forceinline__ void chunk_load_rc(const TileView &A, int Z0, int c, int nch,
                                              bool all64, Chunk &ch) {
  if (c >= nch) return;
  const int k0 = Z0 + 4 * c;
  ch.rc4 = *(const uint32_t *)(A.rc + k0);
  if (all64) {
    const u32x4_t *q = (const u32x4_t *)(A.v64 + k0);
    ch.a = q[0];
    ch.b = q[1];
  }
}

// Stage 1a without a branch: a chunk past the window reloads the window's
// last chunk (in bounds), so the number of loads in flight stays fixed.
__device__ __forceinline__ void chunk_load64_fixed(const TileView &A, int Z0, int c, int nch,
                                                   Chunk &ch) {
  const int k0 = Z0 + 4 * (c < nch ? c : nch - 1);
  ch.rc4 = *(const uint32_t *)(A.rc + k0);
  const u32x4_t *q = (const u32x4_t *)(A.v64 + k0);
  ch.a = q[0];
  ch.b = q[1];
}

// Stage 2 without a branch (FP64 payload): x of all four slots; a slot past
// the tile's entries gathers an in-bounds x of the tile's column block.
__device__ __forceinline__ void chunk_gather64_fixed(const WaveLds &L, const double *__restrict__ x,
                                                     Chunk &ch) {
  const int xo = L.tcol[ch.j];
#pragma unroll
  for (int u = 0; u < 4; ++u) chunk_set_x(ch, u, x[xo + ((ch.rc4 >> (8 * u)) & 15)]);
}

// Stage 1c without a branch (after chunk_locate): two 16-byte loads at the
// tile's level from 16-byte aligned addresses; FP64: the chunk's 32 bytes;
// FP32: its 16 bytes (loaded twice); FP16: the 16 bytes holding its 8 bytes
// (words 2*wsel, 2*wsel+1); INT8 / SKIP: the 16 bytes holding its 4 bytes
// (word wsel).  A chunk past the window loads the window's last chunk.
__device__ __forceinline__ void chunk_load_pay_fixed(const TileView &A, int Z0, int c, int nch,
                                                     Chunk &ch) {
  const int k0 = Z0 + 4 * (c < nch ? c : nch - 1);
  const int lv = ch.lvl;
  const char *p64 = (const char *)(A.v64 + k0);
  const char *p32 = (const char *)(A.v32 + k0);
  const char *p16 = (const char *)(A.v16 + (k0 & ~7));
  const char *p8 = (const char *)(A.v8 + (k0 & ~15));
  const char *p = lv == MF_PREC_FP64 ? p64 : lv == MF_PREC_FP32 ? p32 : lv == MF_PREC_FP16 ? p16 : p8;
  ch.wsel = lv == MF_PREC_FP16 ? (k0 >> 2) & 1 : (k0 >> 2) & 3;
  ch.a = *(const u32x4_t *)p;
  ch.b = *(const u32x4_t *)(p + (lv == MF_PREC_FP64 ? 16 : 0));
}

__device__ __forceinline__ void chunk_load_lv_fixed(const TileView &A, int Z0, int c, int nch,
                                                    Chunk &ch) {
  ch.rc4 = *(const uint32_t *)(A.rc + Z0 + 4 * (c < nch ? c : nch - 1));
  chunk_load_pay_fixed(A, Z0, c, nch, ch);
}

// Stage 2 without a branch, per-tile levels: for every slot an 8-byte load,
// x[xi] for FP64 tiles, else the aligned pair of x32 holding x32[xi] (the
// half is picked by the column's low bit when consumed).
__device__ __forceinline__ void chunk_gather_lv_fixed(const WaveLds &L,
                                                      const double *__restrict__ x,
                                                      const float *__restrict__ x32, Chunk &ch,
                                                      bool x32only) {
  const int xo = L.tcol[ch.j];
  const bool wide = ch.lvl == MF_PREC_FP64 && !x32only;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int xi = xo + ((ch.rc4 >> (8 * u)) & 15);
    const uint2 *q = wide ? (const uint2 *)(x + xi) : (const uint2 *)(x32 + (xi & ~1));
    const uint2 v = *q;
    chunk_set_x(ch, u, __builtin_bit_cast(double, ((uint64_t)v.y << 32) | v.x));
  }
}

// Adds a chunk's four slot products into the window's row accumulators.
// The slots of one tile are in CSR order, so slots of one row are adjacent:
// each run of equal rows is summed in registers and added once, by the slot
// that ends it.  Padding slots repeat the last entry's (row,col) byte and
// carry a zero product, so they extend that entry's run.  No branches on the
// slot data: the four adds are exec-masked.
__device__ __forceinline__ void chunk_accumulate(WaveLds &L, int k0, int k1, int k2, int k3,
                                                 double p0, double p1, double p2, double p3) {
  const double s1 = p1 + (k1 == k0 ? p0 : 0.0);
  const double s2 = p2 + (k2 == k1 ? s1 : 0.0);
  const double s3 = p3 + (k3 == k2 ? s2 : 0.0);
  if (k1 != k0) atomicAdd(&L.acc[k0], p0);
  if (k2 != k1) atomicAdd(&L.acc[k1], s1);
  if (k3 != k2) atomicAdd(&L.acc[k2], s2);
  atomicAdd(&L.acc[k3], s3);
}

// Row accumulator index of slot u: the tile's row base plus the row nibble.
__device__ __forceinline__ int chunk_row(const Chunk &ch, int ro, int u) {
  return ro + ((ch.rc4 >> (8 * u + 4)) & 15);
}

// Stage 3 for the fixed-width loads (per-tile levels), without branches on
// the slot data.  Every level's value is decoded and the tile's level picks
// one: FP64 a*x in FP64 (x32only: a*(double)x32); FP32 (a*x32); FP16 and
// INT8 (a*x32)*scale, rounded to FP32 then widened.  SKIP and invalid slots
// contribute 0.
__device__ __forceinline__ void chunk_consume_lv_flat(WaveLds &L, const Chunk &ch, bool x32only) {
  const int ro = L.trow[ch.j] * 16;
  const int lv = ch.lvl;
  const float sc = lv == MF_PREC_FP16 || lv == MF_PREC_INT8 ? L.tscale[ch.j] : 1.0f;
  const uint32_t w0 = ch.wsel == 0 ? ch.a.x : ch.wsel == 1 ? ch.a.y : ch.wsel == 2 ? ch.a.z : ch.a.w;
  const uint32_t h0 = ch.wsel == 0 ? ch.a.x : ch.a.z;  // FP16: words 2*wsel, 2*wsel+1
  const uint32_t h1 = ch.wsel == 0 ? ch.a.y : ch.a.w;
  const int live = lv == MF_PREC_SKIP ? 0 : ch.vmask;
  double p[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int col = (ch.rc4 >> (8 * u)) & 15;
    const uint64_t raw = __builtin_bit_cast(uint64_t, chunk_x(ch, u));
    const float xf = __builtin_bit_cast(float, (uint32_t)(col & 1 ? raw >> 32 : raw));
    const double a64 = __builtin_bit_cast(double, ((uint64_t)chunk_word(ch, 2 * u + 1) << 32) |
                                                      chunk_word(ch, 2 * u));
    const double x64 = x32only ? (double)xf : __builtin_bit_cast(double, raw);
    const float v32 = __builtin_bit_cast(float, chunk_word(ch, u));
    const float v16 = half_bits_to_float((uint16_t)((u < 2 ? h0 : h1) >> (16 * (u & 1))));
    const float v8 = (float)(int8_t)(w0 >> (8 * u));
    const float vn = lv == MF_PREC_FP32 ? v32 : lv == MF_PREC_FP16 ? v16 : v8;
    const double prod = lv == MF_PREC_FP64 ? a64 * x64 : (double)(vn * xf * sc);
    p[u] = (live >> u & 1) ? prod : 0.0;
  }
  chunk_accumulate(L, chunk_row(ch, ro, 0), chunk_row(ch, ro, 1), chunk_row(ch, ro, 2),
                   chunk_row(ch, ro, 3), p[0], p[1], p[2], p[3]);
}

// Stage 3 for the fixed-width loads (per-tile levels), branching per slot:
// fewer live registers than chunk_consume_lv_flat, for kernels whose
// register budget is shared with a solver loop.
__device__ __forceinline__ void chunk_consume_lv_fixed(WaveLds &L, const Chunk &ch, bool x32only) {
  if (!ch.vmask || ch.lvl == MF_PREC_SKIP) return;
  const int ro = L.trow[ch.j] * 16;
  const int lv = ch.lvl;
  const float sc = lv == MF_PREC_FP16 || lv == MF_PREC_INT8 ? L.tscale[ch.j] : 1.0f;
  const uint32_t w0 = ch.wsel == 0 ? ch.a.x : ch.wsel == 1 ? ch.a.y : ch.wsel == 2 ? ch.a.z : ch.a.w;
  const uint32_t h0 = ch.wsel == 0 ? ch.a.x : ch.a.z;  // FP16: words 2*wsel, 2*wsel+1
  const uint32_t h1 = ch.wsel == 0 ? ch.a.y : ch.a.w;
  int key = -1;
  double run = 0.0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (!(ch.vmask >> u & 1)) continue;
    const int col = (ch.rc4 >> (8 * u)) & 15;
    const int kk = ro + ((ch.rc4 >> (8 * u + 4)) & 15);
    const uint64_t raw = __builtin_bit_cast(uint64_t, chunk_x(ch, u));
    const float xf = __builtin_bit_cast(float, (uint32_t)(col & 1 ? raw >> 32 : raw));
    double prod;
    if (lv == MF_PREC_FP64) {
      prod = __builtin_bit_cast(double, ((uint64_t)chunk_word(ch, 2 * u + 1) << 32) |
                                            chunk_word(ch, 2 * u)) *
             (x32only ? (double)xf : __builtin_bit_cast(double, raw));
    } else {
      if (lv == MF_PREC_FP32) {
        prod = (double)(__builtin_bit_cast(float, chunk_word(ch, u)) * xf);
      } else if (lv == MF_PREC_FP16) {
        const uint16_t h = (uint16_t)((u < 2 ? h0 : h1) >> (16 * (u & 1)));
        prod = (double)(half_bits_to_float(h) * xf * sc);
      } else {
        const int8_t q = (int8_t)(w0 >> (8 * u));
        prod = (double)((float)q * xf * sc);
      }
    }
    if (kk != key) {
      if (key >= 0) atomicAdd(&L.acc[key], run);
      key = kk;
      run = prod;
    } else {
      run += prod;
    }
  }
  atomicAdd(&L.acc[key], run);
}

// Stage 1b: tile, level and valid slots of the chunk (LDS only).
__device__ __forceinline__ void chunk_locate(const WaveLds &L, int c, int nch, bool all64,
                                             Chunk &ch) {
  ch.vmask = 0;
  ch.j = 0;
  ch.lvl = MF_PREC_FP64;
  if (c >= nch) return;
  const int j = L.ctile[c];
  const int live = L.tend[j] - 4 * c;  // real entries from slot 0 on (>= 1)
  ch.j = j;
  ch.vmask = live >= 4 ? 15 : (1 << live) - 1;
  if (!all64) ch.lvl = L.tprec[j];
}

// Stage 1c (per-tile levels, after chunk_locate): the payload at the tile's level.
__device__ __forceinline__ void chunk_load_payload(const TileView &A, int Z0, int c, Chunk &ch) {
  if (!ch.vmask) return;
  const int k0 = Z0 + 4 * c;
  const int lv = ch.lvl;
  if (lv == MF_PREC_FP64) {
    const u32x4_t *q = (const u32x4_t *)(A.v64 + k0);
    ch.a = q[0];
    ch.b = q[1];
  } else if (lv == MF_PREC_FP32) {
    ch.a = *(const u32x4_t *)(A.v32 + k0);
  } else if (lv == MF_PREC_FP16) {
    const uint2 h = *(const uint2 *)(A.v16 + k0);
    ch.a.x = h.x;
    ch.a.y = h.y;
  } else if (lv == MF_PREC_INT8) {
    ch.a.x = *(const uint32_t *)(A.v8 + k0);
  }
}

// Stage 2 (needs the (row,col) bytes): gather x at every entry's column.
__device__ __forceinline__ void chunk_gather(const WaveLds &L, const double *__restrict__ x,
                                             const float *__restrict__ x32, bool all64,
                                             Chunk &ch) {
  if (!ch.vmask || ch.lvl == MF_PREC_SKIP) return;
  const int xo = L.tcol[ch.j];
  const bool wide = all64 || ch.lvl == MF_PREC_FP64;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (!(ch.vmask >> u & 1)) continue;
    const int xi = xo + ((ch.rc4 >> (8 * u)) & 15);
    chunk_set_x(ch, u, wide ? x[xi] : (double)x32[xi]);
  }
}

// Stage 3 (needs payload and x): products, merged per row in registers and
// added into the window's LDS row accumulators.
__device__ __forceinline__ void chunk_consume(WaveLds &L, const Chunk &ch, bool all64) {
  if (!ch.vmask || ch.lvl == MF_PREC_SKIP) return;
  const int ro = L.trow[ch.j] * 16;
  const int lv = all64 ? MF_PREC_FP64 : ch.lvl;
  const float sc = lv == MF_PREC_FP16 || lv == MF_PREC_INT8 ? L.tscale[ch.j] : 1.0f;
  int key = -1;
  double run = 0.0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (!(ch.vmask >> u & 1)) continue;
    const int kk = ro + ((ch.rc4 >> (8 * u + 4)) & 15);
    double prod;
    if (lv == MF_PREC_FP64) {
      prod = __builtin_bit_cast(double, ((uint64_t)chunk_word(ch, 2 * u + 1) << 32) |
                                            chunk_word(ch, 2 * u)) * chunk_x(ch, u);
    } else if (lv == MF_PREC_FP32) {
      prod = (double)(__builtin_bit_cast(float, chunk_word(ch, u)) * (float)chunk_x(ch, u));
    } else if (lv == MF_PREC_FP16) {
      const uint16_t h = (uint16_t)(chunk_word(ch, u >> 1) >> (16 * (u & 1)));
      prod = (double)(half_bits_to_float(h) * (float)chunk_x(ch, u) * sc);
    } else {
      const int8_t q = (int8_t)(ch.a.x >> (8 * u));
      prod = (double)((float)q * (float)chunk_x(ch, u) * sc);
    }
    if (kk != key) {
      if (key >= 0) atomicAdd(&L.acc[key], run);
      key = kk;
      run = prod;
    } else {
      run += prod;
    }
  }
  atomicAdd(&L.acc[key], run);
}

// ---------------------------------------------------------------------------
// Level-grouped windows (per-tile levels).
//
// A window's chunks are streamed grouped by their tile's level.  At staging,
// every tile's chunks get consecutive positions in its level's group (FP64,
// then FP32, FP16, INT8, and last SKIP) from fixed-order wave prefix sums of
// the per-level chunk counts, so the grouping, and with it the order of the
// row accumulations, is the same on every run.  L.ctile[pos] then holds the
// tile of the chunk at position pos, and L.tend[j] packs the end of tile j's
// entries (bits 0-12) with the offset from the tile's first position to its
// first chunk (bits 13-23, biased by 1024): chunk = pos + offset.  Each group
// is streamed by a loop compiled for its level: payload loads of the level's
// width (INT8 4 B, FP16 8 B, FP32 16 B, FP64 32 B per chunk), 4-byte x32
// gathers below FP64, and a consume without level selection.  The SKIP group
// is streamed only by an all-FP64 pass over every position.
//
// The functions take any LDS layout L with members tcol, tscale, tend, trow,
// ctile and acc (WaveLds for streamed windows, ResLds for the solvers'
// resident windows).
// ---------------------------------------------------------------------------
struct LvGroups {
  // positions: FP64 [0,b1), FP32 [b1,b2), FP16 [b2,b3), INT8 [b3,e), SKIP [e,eall)
  int b1, b2, b3, e, eall;
};

template <int LV>
struct LvChunk {
  uint32_t rc4;
  int valid, tcol, ro;    // valid: the lane's position lies in the group
  float sc;
  u32x4_t a, b;           // payload words (b: FP64 only)
  double d0, d1, d2, d3;  // x of FP64 tiles multiplying FP64 x
  float f0, f1, f2, f3;   // x32 of the other cases
};

// Inclusive prefix sums of three values over the wave's lanes.
__device__ __forceinline__ void wave_scan3(uint32_t &v0, uint32_t &v1, uint32_t &v2, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t u0 = __shfl_up(v0, d), u1 = __shfl_up(v1, d), u2 = __shfl_up(v2, d);
    if (lane >= d) {
      v0 += u0;
      v1 += u1;
      v2 += u2;
    }
  }
}

// Per-level chunk counts packed in three words: FP64 | FP32 << 16,
// FP16 | INT8 << 16, SKIP (a window holds at most 1024 chunks).
struct LvCount {
  uint32_t lo, hi, sk;
  __device__ __forceinline__ LvCount operator+(const LvCount &o) const {
    return LvCount{lo + o.lo, hi + o.hi, sk + o.sk};
  }
  __device__ __forceinline__ LvCount operator-(const LvCount &o) const {
    return LvCount{lo - o.lo, hi - o.hi, sk - o.sk};
  }
  __device__ __forceinline__ int field(int lvl) const {
    if (lvl >= MF_PREC_SKIP) return (int)sk;
    const uint32_t w = lvl < 2 ? lo : hi;
    return (int)((lvl & 1) ? w >> 16 : w & 0xFFFFu);
  }
};

__device__ __forceinline__ LvCount lv_pack(int lvl, int n) {
  return LvCount{lvl == MF_PREC_FP64 ? (uint32_t)n : lvl == MF_PREC_FP32 ? (uint32_t)n << 16 : 0u,
                 lvl == MF_PREC_FP16 ? (uint32_t)n : lvl == MF_PREC_INT8 ? (uint32_t)n << 16 : 0u,
                 lvl >= MF_PREC_SKIP ? (uint32_t)n : 0u};
}

// Inclusive wave scan of a count; tot receives lane 63's value.
__device__ __forceinline__ LvCount lv_scan(LvCount v, int lane, LvCount &tot) {
  wave_scan3(v.lo, v.hi, v.sk, lane);
  tot = LvCount{(uint32_t)__builtin_amdgcn_readlane(v.lo, 63), (uint32_t)__builtin_amdgcn_readlane(v.hi, 63),
                (uint32_t)__builtin_amdgcn_readlane(v.sk, 63)};
  return v;
}

__device__ __forceinline__ LvGroups lv_groups(const LvCount &T) {
  LvGroups g;
  g.b1 = (int)(T.lo & 0xFFFFu);
  g.b2 = g.b1 + (int)(T.lo >> 16);
  g.b3 = g.b2 + (int)(T.hi & 0xFFFFu);
  g.e = g.b3 + (int)(T.hi >> 16);
  g.eall = g.e + (int)T.sk;
  return g;
}

__device__ __forceinline__ int lv_base(const LvGroups &g, int l) {
  return l == MF_PREC_FP64 ? 0 : l == MF_PREC_FP32 ? g.b1 : l == MF_PREC_FP16 ? g.b2
       : l == MF_PREC_INT8 ? g.b3 : g.e;
}

// Tile j of the window into LDS for the level-grouped stream: o, e: its
// payload begin and end relative to the window (multiples of 4); pos: its
// first position in its level's group.
template <bool COUNT, class LDS>
__device__ __forceinline__ void stage_tile_lv(LDS &L, int j, int o, int e, int col, int trow, int nz,
                                              int lvl, float sc, int pos, PrecCount *cnt) {
  if (COUNT) cnt->add(lvl);
  L.tcol[j] = col << 4;
  L.trow[j] = (uint8_t)trow;
  L.tscale[j] = sc;
  L.tend[j] = (o + nz + 1) | (((o >> 2) - pos + 1024) << 13);
  for (int q = 0; q < ((e - o) >> 2); ++q) L.ctile[pos + q] = (uint8_t)j;
}

// Stages the window's tiles lane (a) and lane + 64 (b) of nt, grouped by
// level; returns the group bounds (wave-uniform).  o*, e*: payload begin and
// end relative to the window; r*: tile-row relative to the window.
template <bool COUNT>
__device__ __forceinline__ LvGroups stage_levels(WaveLds &L, int lane, int nt, int oa, int ea,
                                                 int ca, int ra, int na, int la, float sa, int ob,
                                                 int eb, int cb, int rb, int nb, int lb, float sb,
                                                 PrecCount *cnt) {
  const bool va = lane < nt, vb = lane + 64 < nt;
  const LvCount ca_ = lv_pack(la, va ? (ea - oa) >> 2 : 0);
  const LvCount cb_ = lv_pack(lb, vb ? (eb - ob) >> 2 : 0);
  LvCount ta, tb;
  const LvCount ia = lv_scan(ca_, lane, ta);
  const LvCount ib = lv_scan(cb_, lane, tb);
  const LvGroups g = lv_groups(ta + tb);
  if (va) stage_tile_lv<COUNT>(L, lane, oa, ea, ca, ra, na, la, sa, lv_base(g, la) + (ia - ca_).field(la), cnt);
  if (vb)
    stage_tile_lv<COUNT>(L, lane + 64, ob, eb, cb, rb, nb, lb, sb,
                         lv_base(g, lb) + (ib - cb_ + ta).field(lb), cnt);
  return g;
}

// Element at a 32-bit byte offset from a wave-uniform base: the load takes
// the base in scalar registers and one 32-bit offset register per lane (no
// 64-bit address arithmetic per lane).  Vector offsets stay below 2^32 bytes
// because handles are limited to 2^28 rows (mf_tile_create).
template <class T>
__device__ __forceinline__ T ld_at(const void *base, uint32_t byte_off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}

// Payload streams of a window: the level arrays from the window's first slot.
struct WinPay {
  const uint8_t *rc;
  const void *v64, *v32, *v16, *v8;
};

__device__ __forceinline__ WinPay win_pay(const TileView &A, int Z0) {
  return WinPay{A.rc + Z0, A.v64 + Z0, A.v32 + Z0, A.v16 + Z0, A.v8 + Z0};
}

// Chunk at group position pos (< end; a lane past the end repeats the last
// chunk and is marked invalid, so every slice issues the same loads): LDS
// lookups, then the (row,col) bytes and the payload at the level's width.
// A chunk needs no per-slot mask: the padding slots of a tile hold value 0
// and repeat its last (row,col) byte, so they add exact zeros to the row the
// chunk's last real slot adds to.
template <int LV, class LDS>
__device__ __forceinline__ void lv_load(const WinPay &W, const LDS &L, int pos, int end,
                                        LvChunk<LV> &s) {
  const bool valid = pos < end;
  const int p = valid ? pos : end - 1;
  const int j = L.ctile[p];
  const int te = L.tend[j];
  const int c = p + (te >> 13) - 1024;
  s.valid = valid;
  s.tcol = L.tcol[j];
  s.ro = L.trow[j] * LDS::kAccStride;
  s.sc = (LV == MF_PREC_FP16 || LV == MF_PREC_INT8) ? L.tscale[j] : 1.0f;
  const uint32_t c4 = 4u * (uint32_t)c;  // first slot of the chunk in the window
  s.rc4 = ld_at<uint32_t>(W.rc, c4);
  if (LV == MF_PREC_FP64) {
    s.a = ld_at<u32x4_t>(W.v64, 8u * c4);
    s.b = ld_at<u32x4_t>(W.v64, 8u * c4 + 16u);
  } else if (LV == MF_PREC_FP32) {
    s.a = ld_at<u32x4_t>(W.v32, 4u * c4);
  } else if (LV == MF_PREC_FP16) {
    const uint2 h = ld_at<uint2>(W.v16, 2u * c4);
    s.a.x = h.x;
    s.a.y = h.y;
  } else {
    s.a.x = ld_at<uint32_t>(W.v8, c4);
  }
}

// x at the chunk's four columns: FP64 x for FP64 tiles unless X32, else x32.
template <int LV, bool X32>
__device__ __forceinline__ void lv_gather(const double *__restrict__ x, const float *__restrict__ x32,
                                          LvChunk<LV> &s) {
  const uint32_t t = (uint32_t)s.tcol;
  const uint32_t x0 = t + (s.rc4 & 15), x1 = t + ((s.rc4 >> 8) & 15);
  const uint32_t x2 = t + ((s.rc4 >> 16) & 15), x3 = t + ((s.rc4 >> 24) & 15);
  if (LV == MF_PREC_FP64 && !X32) {
    s.d0 = ld_at<double>(x, 8u * x0);
    s.d1 = ld_at<double>(x, 8u * x1);
    s.d2 = ld_at<double>(x, 8u * x2);
    s.d3 = ld_at<double>(x, 8u * x3);
  } else {
    s.f0 = ld_at<float>(x32, 4u * x0);
    s.f1 = ld_at<float>(x32, 4u * x1);
    s.f2 = ld_at<float>(x32, 4u * x2);
    s.f3 = ld_at<float>(x32, 4u * x3);
  }
}

// Slot u's product: FP64 for FP64 tiles, FP32 for the others (the FP16 and
// INT8 tile scale applied once per slot).
template <int LV>
using LvAcc = typename std::conditional<LV == MF_PREC_FP64, double, float>::type;

template <int LV, bool X32>
__device__ __forceinline__ LvAcc<LV> lv_product(const LvChunk<LV> &s, int u) {
  const float xf = u == 0 ? s.f0 : u == 1 ? s.f1 : u == 2 ? s.f2 : s.f3;
  if constexpr (LV == MF_PREC_FP64) {
    const uint32_t lo = u == 0 ? s.a.x : u == 1 ? s.a.z : u == 2 ? s.b.x : s.b.z;
    const uint32_t hi = u == 0 ? s.a.y : u == 1 ? s.a.w : u == 2 ? s.b.y : s.b.w;
    const double a = __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
    const double xd = u == 0 ? s.d0 : u == 1 ? s.d1 : u == 2 ? s.d2 : s.d3;
    return a * (X32 ? (double)xf : xd);
  } else if constexpr (LV == MF_PREC_FP32) {
    const uint32_t w = u == 0 ? s.a.x : u == 1 ? s.a.y : u == 2 ? s.a.z : s.a.w;
    return __builtin_bit_cast(float, w) * xf;
  } else if constexpr (LV == MF_PREC_FP16) {
    const uint32_t w = u < 2 ? s.a.x : s.a.y;
    return half_bits_to_float((uint16_t)(w >> (16 * (u & 1)))) * xf * s.sc;
  } else {
    return (float)(int8_t)(s.a.x >> (8 * u)) * xf * s.sc;
  }
}

// Adds four slot products into row accumulators k0..k3 (see
// chunk_accumulate): runs of equal rows are summed in T (FP32 below FP64)
// and each run is added to its row once, in FP64.
template <class LDS, typename T>
__device__ __forceinline__ void lv_accumulate(LDS &L, int k0, int k1, int k2, int k3, T p0, T p1,
                                              T p2, T p3) {
  const T s1 = p1 + (k1 == k0 ? p0 : T(0));
  const T s2 = p2 + (k2 == k1 ? s1 : T(0));
  const T s3 = p3 + (k3 == k2 ? s2 : T(0));
  if (k1 != k0) atomicAdd(&L.acc[k0], (double)p0);
  if (k2 != k1) atomicAdd(&L.acc[k1], (double)s1);
  if (k3 != k2) atomicAdd(&L.acc[k2], (double)s2);
  atomicAdd(&L.acc[k3], (double)s3);
}

// Products of the chunk's slots into the window's row accumulators (nothing
// for a lane past the group's end).
template <int LV, bool X32, class LDS>
__device__ __forceinline__ void lv_consume(LDS &L, const LvChunk<LV> &s) {
  if (!s.valid) return;
  LvAcc<LV> p[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) p[u] = lv_product<LV, X32>(s, u);
  const int r0 = s.ro + ((s.rc4 >> 4) & 15), r1 = s.ro + ((s.rc4 >> 12) & 15);
  const int r2 = s.ro + ((s.rc4 >> 20) & 15), r3 = s.ro + ((s.rc4 >> 28) & 15);
  lv_accumulate(L, r0, r1, r2, r3, p[0], p[1], p[2], p[3]);
}

// Group positions [b, e) in slices of 64 chunks, D - 1 slices ahead: while
// slice k is consumed, slice k+1's gathers and the loads of slices k+2 ..
// k+D-1 are in flight.  Each step issues slice k+D-1's loads before slice
// k+1's gathers, so waiting for slice k+1's (row,col) bytes (loaded D - 2
// steps earlier) never waits for slice k's gathers (loads retire in issue
// order), and a step waits on no load younger than one step.  Every step
// issues the same loads (slices past the group repeat its last chunk,
// marked invalid), so the number of loads in flight at every wait is fixed;
// the steps are those of whole trips of the D-step body, then a tail of the
// remaining nsl mod D steps.  FP64 chunks with FP64 x hold the most
// registers and run three slices deep, the others four.
template <int LV, bool X32>
struct LvDepth {
  static constexpr int D = (LV == MF_PREC_FP64 && !X32) ? 3 : 4;
};

// Step i of a trip starting at slice k (c[] rotates with i, so i is static).
template <int LV, bool X32, int D, int I, class LDS>
__device__ __forceinline__ void level_step(const WinPay &W, LDS &L, int q, int e,
                                           const double *__restrict__ x,
                                           const float *__restrict__ x32, LvChunk<LV> (&c)[D]) {
  lv_load<LV>(W, L, q + 64 * I, e, c[(I + D - 1) % D]);
  __builtin_amdgcn_sched_barrier(0);
  lv_gather<LV, X32>(x, x32, c[(I + 1) % D]);
  __builtin_amdgcn_sched_barrier(0);
  lv_consume<LV, X32>(L, c[I]);
}

template <int LV, bool X32, class LDS>
__device__ __forceinline__ void level_group(const TileView &A, LDS &L, int Z0, int b, int e,
                                            const double *__restrict__ x,
                                            const float *__restrict__ x32, int lane) {
  constexpr int D = LvDepth<LV, X32>::D;
  static_assert(D == 3 || D == 4, "level_group depth");
  const int nsl = (e - b + 63) >> 6;
  const int p = b + lane;
  const WinPay W = win_pay(A, Z0);
  // The slices rotate through c[] by unrolling D steps (a register copy of a
  // chunk whose loads are in flight would wait for them).
  LvChunk<LV> c[D];
  // The scheduling barriers keep the issue order of the loads as written
  // (the scheduler would otherwise move a slice's loads, which wait on LDS
  // lookups, behind the next gathers).
#pragma unroll
  for (int i = 0; i < D - 1; ++i) {
    lv_load<LV>(W, L, p + 64 * i, e, c[i]);
    __builtin_amdgcn_sched_barrier(0);
  }
  lv_gather<LV, X32>(x, x32, c[0]);
  int k = 0;
  for (; k + D <= nsl; k += D) {
    const int q = p + 64 * (k + D - 1);
    level_step<LV, X32, D, 0>(W, L, q, e, x, x32, c);
    level_step<LV, X32, D, 1>(W, L, q, e, x, x32, c);
    level_step<LV, X32, D, 2>(W, L, q, e, x, x32, c);
    if constexpr (D == 4) level_step<LV, X32, D, 3>(W, L, q, e, x, x32, c);
  }
  // the last nsl - k (< D) slices
  const int q = p + 64 * (k + D - 1);
  const int rest = nsl - k;
  if (rest >= 1) level_step<LV, X32, D, 0>(W, L, q, e, x, x32, c);
  if (rest >= 2) level_step<LV, X32, D, 1>(W, L, q, e, x, x32, c);
  if constexpr (D == 4)
    if (rest >= 3) level_step<LV, X32, D, 2>(W, L, q, e, x, x32, c);
}

// The staged window's groups.  all64: every position (SKIP included) at
// FP64 with FP64 x; otherwise one level at a time, x32only making FP64 tiles
// multiply x32 (in FP64 arithmetic).
template <class LDS>
__device__ __forceinline__ void level_groups(const TileView &A, LDS &L, int Z0, const LvGroups &g,
                                             const double *__restrict__ x,
                                             const float *__restrict__ x32, bool all64,
                                             bool x32only, int lane) {
  if (all64) {
    if (g.eall > 0) level_group<MF_PREC_FP64, false>(A, L, Z0, 0, g.eall, x, x32, lane);
    return;
  }
  if (g.b1 > 0) {
    if (x32only)
      level_group<MF_PREC_FP64, true>(A, L, Z0, 0, g.b1, x, x32, lane);
    else
      level_group<MF_PREC_FP64, false>(A, L, Z0, 0, g.b1, x, x32, lane);
  }
  if (g.b2 > g.b1) level_group<MF_PREC_FP32, true>(A, L, Z0, g.b1, g.b2, x, x32, lane);
  if (g.b3 > g.b2) level_group<MF_PREC_FP16, true>(A, L, Z0, g.b2, g.b3, x, x32, lane);
  if (g.e > g.b3) level_group<MF_PREC_INT8, true>(A, L, Z0, g.b3, g.e, x, x32, lane);
}

// ---------------------------------------------------------------------------
// Resident windows of the persistent solvers.
//
// A solver wave whose SpMV share is a tile-row range of at most kResRows
// tile-rows, kResTiles tiles and kResChunks chunks (no heavy tile-row) keeps
// that whole share staged in its LDS for the length of the solve: the tile
// metadata is read from HBM once, and again only when the tile levels change
// (a policy update or the FP64 fallback), when the level groups are rebuilt.
// An SpMV is then only the payload stream of one window: no descriptor, no
// metadata loads and no window cuts on the critical path.
// ---------------------------------------------------------------------------
constexpr int kResTiles = 256;
constexpr int kResRows = 40;
constexpr int kResChunks = 1024;

struct ResLds {
  int tcol[kResTiles];        // 16 * tile column
  float tscale[kResTiles];    // FP16 / INT8 scale of the tile's current level
  int tend[kResTiles];        // entry end | position offset (level groups)
  uint8_t trow[kResTiles];    // tile-row relative to the wave's first
  uint8_t ctile[kResChunks];  // tile of each group position
  // Row sums, 17 per tile-row: rows of equal index in different tile-rows
  // (which the lanes of one slice often add to at once) fall in different
  // LDS banks.
  static constexpr int kAccStride = 17;
  double acc[kResRows * kAccStride];
};

// Accumulator of window row i (relative to the window's first row).
__device__ __forceinline__ int res_acc_index(int i) { return (i >> 4) * ResLds::kAccStride + (i & 15); }

// The wave's resident window: tile-rows [tr0, tr1), tiles [ta, ta + nt),
// payload from slot Z0, its level groups and per-level tile counts (for the
// precision histogram).
struct ResWin {
  int tr0, tr1, ta, nt, Z0;
  LvGroups g;
  unsigned n64, n32, n16, n8, nskip;
};

// Whether tile-rows [tr0, tr1) fit a resident window (host and device).
__host__ __device__ __forceinline__ bool res_fits(int rows, int tiles, int chunks) {
  return rows <= kResRows && tiles <= kResTiles && chunks <= kResChunks;
}

// Stages (or re-stages) the resident window of tile-rows [tr0, tr1) with the
// tiles' current levels (FP64 for all when all64).  Tiles lane + 64 q,
// q = 0..3, are taken in that order.
template <int MODE>
__device__ __forceinline__ void res_stage(const TileView &A, ResLds &L, ResWin &w, int tr0, int tr1,
                                          const uint8_t *__restrict__ seglev, bool all64, int lane) {
  const int ta = __builtin_amdgcn_readfirstlane(A.tile_ptr[tr0]);
  const int nt = __builtin_amdgcn_readfirstlane(A.tile_ptr[tr1]) - ta;
  const int Z0 = __builtin_amdgcn_readfirstlane(A.tile_off[ta]);
  w.tr0 = tr0;
  w.tr1 = tr1;
  w.ta = ta;
  w.nt = nt;
  w.Z0 = Z0;
  if (nt <= 0) {  // no tiles (a wave past the last tile-row): nothing to load
    w.g = LvGroups{0, 0, 0, 0, 0};
    w.n64 = w.n32 = w.n16 = w.n8 = w.nskip = 0;
    return;
  }
  int o[4], e[4], col[4], row[4], nz[4], lv[4];
  float sc[4];
  LvCount c[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = lane + 64 * q;
    const bool v = j < nt;
    const int t = ta + (v ? j : nt - 1);  // lanes past the last tile reload it
    o[q] = A.tile_off[t] - Z0;
    e[q] = A.tile_off[t + 1] - Z0;
    col[q] = A.tile_col[t];
    row[q] = A.tile_row[t] - tr0;
    nz[q] = A.tile_nz[t];
    lv[q] = MODE == kAll64 || all64 ? MF_PREC_FP64
          : MODE == kPerTile ? A.prec[t] : seglev[col[q]];
    sc[q] = lv[q] == MF_PREC_FP16 ? A.s16[t] : A.s8[t];
    c[q] = lv_pack(lv[q], v ? (e[q] - o[q]) >> 2 : 0);
  }
  LvCount tot{0u, 0u, 0u}, pre[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    LvCount tq;
    pre[q] = lv_scan(c[q], lane, tq) - c[q] + tot;
    tot = tot + tq;
  }
  w.g = lv_groups(tot);
  unsigned n[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = lane + 64 * q;
    if (j < nt)
      stage_tile_lv<false>(L, j, o[q], e[q], col[q], row[q], nz[q], lv[q], sc[q],
                           lv_base(w.g, lv[q]) + pre[q].field(lv[q]), nullptr);
#pragma unroll
    for (int l = 0; l < 5; ++l)
      n[l] += (unsigned)__builtin_popcountll(__ballot(j < nt && lv[q] == l));
  }
  w.n64 = n[0];
  w.n32 = n[1];
  w.n16 = n[2];
  w.n8 = n[3];
  w.nskip = n[4];
  __builtin_amdgcn_wave_barrier();
}

// One SpMV of the resident window through epi(row, value).
template <bool COUNT, typename Epi>
__device__ __forceinline__ void res_spmv(const TileView &A, ResLds &L, const ResWin &w,
                                         const double *__restrict__ x,
                                         const float *__restrict__ x32, bool all64, bool x32only,
                                         int lane, PrecCount *cnt, Epi epi) {
  const int nr = (w.tr1 - w.tr0) * 16;
  for (int i = lane; i < (w.tr1 - w.tr0) * ResLds::kAccStride; i += 64) L.acc[i] = 0.0;
  // (the counts are the wave's, alike in every lane: kept in scalar
  // registers, and the histogram flush takes lane 0's)
  if (COUNT && all64) {
    cnt->c0 += (unsigned)w.nt;
  } else if (COUNT) {
    cnt->c0 += w.n64;
    cnt->c1 += w.n32;
    cnt->c2 += w.n16;
    cnt->c3 += w.n8;
    cnt->c4 += w.nskip;
  }
  __builtin_amdgcn_wave_barrier();
  level_groups(A, L, w.Z0, w.g, x, x32, all64, x32only, lane);
  __builtin_amdgcn_wave_barrier();
  for (int i = lane; i < nr; i += 64) epi(16 * w.tr0 + i, L.acc[res_acc_index(i)]);
  __builtin_amdgcn_wave_barrier();
}

// Descriptor of tile-rows [tr0, tr1) from the tile pointers (callers without
// a precomputed one).
__device__ __forceinline__ WaveDesc wave_desc(const TileView &A, int tr0, int tr1) {
  WaveDesc d;
  d.tr0 = tr0;
  d.tr1 = tr1;
  d.ta = A.tile_ptr[tr0];
  d.T1 = A.tile_ptr[tr1];
  d.Z0 = A.tile_off[d.ta];
  d.Z1 = A.tile_off[d.T1];
  d.flags = d.aux = 0;
  return d;
}

// Tile j of the window into LDS: column, tile-row, end of its real entries,
// level and scale, and j in every chunk of its payload.  o, e: the tile's
// payload begin and end relative to the window's first slot.
template <int MODE, bool COUNT>
__device__ __forceinline__ void stage_tile(WaveLds &L, int j, int o, int e, int col, int trow,
                                           int nz, int lvl, float sc, PrecCount *cnt) {
  L.tcol[j] = col << 4;
  L.trow[j] = (uint8_t)trow;
  L.tend[j] = o + nz + 1;
  for (int q = o >> 2; q < (e >> 2); ++q) L.ctile[q] = (uint8_t)j;
  if (MODE != kAll64) {
    L.tprec[j] = (uint8_t)lvl;
    L.tscale[j] = sc;
    if (COUNT) cnt->add(lvl);
  }
}

// The staged window's nch chunks from slot Z0 on, in slices of 64 chunks,
// with the loads of later slices in flight while a slice is gathered and
// consumed.  FLAT: per-tile levels consumed without branches on slot data.  On entry cx holds the first slice's (row,col) bytes (and FP64
// payload) from chunk_load_rc; the products end in L.acc.
template <int MODE, int DEPTH, bool FLAT>
__device__ __forceinline__ void window_slices(const TileView &A, WaveLds &L, int Z0, int nch,
                                              const double *__restrict__ x,
                                              const float *__restrict__ x32, bool all64,
                                              bool x32only, int lane, Chunk &cx, Chunk &cy) {
  const int nsl = (nch + 63) >> 6;
  if (DEPTH == 3 && MODE == kAll64) {
    // Three-deep (FP64 payload only): slice m+2's loads go out after
    // slice m's gathers, so waiting for those gathers never waits for
    // them.
    Chunk cz;
    if (nsl > 1) chunk_load64_fixed(A, Z0, lane + 64, nch, cy);
    for (int m = 0; m < nsl; m += 3) {
      const int c0 = 64 * m + lane;
      chunk_locate(L, c0, nch, true, cx);
      chunk_gather64_fixed(L, x, cx);
      if (m + 2 < nsl) chunk_load64_fixed(A, Z0, c0 + 128, nch, cz);
      chunk_consume(L, cx, true);
      if (m + 1 >= nsl) break;
      chunk_locate(L, c0 + 64, nch, true, cy);
      chunk_gather64_fixed(L, x, cy);
      if (m + 3 < nsl) chunk_load64_fixed(A, Z0, c0 + 192, nch, cx);
      chunk_consume(L, cy, true);
      if (m + 2 >= nsl) break;
      chunk_locate(L, c0 + 128, nch, true, cz);
      chunk_gather64_fixed(L, x, cz);
      if (m + 4 < nsl) chunk_load64_fixed(A, Z0, c0 + 256, nch, cy);
      chunk_consume(L, cz, true);
    }
  } else if (DEPTH == 3 && MODE != kAll64) {
    // Three-deep, per-tile levels, fixed-width loads: slice m+2 is
    // located (LDS) and its loads issued after slice m's gathers.
    Chunk cz;
    chunk_locate(L, lane, nch, all64, cx);
    chunk_load_pay_fixed(A, Z0, lane, nch, cx);
    if (nsl > 1) {
      chunk_locate(L, lane + 64, nch, all64, cy);
      chunk_load_lv_fixed(A, Z0, lane + 64, nch, cy);
    }
    for (int m = 0; m < nsl; m += 3) {
      const int c0 = 64 * m + lane;
      chunk_gather_lv_fixed(L, x, x32, cx, x32only);
      if (m + 2 < nsl) {
        chunk_locate(L, c0 + 128, nch, all64, cz);
        chunk_load_lv_fixed(A, Z0, c0 + 128, nch, cz);
      }
      if (FLAT) chunk_consume_lv_flat(L, cx, x32only); else chunk_consume_lv_fixed(L, cx, x32only);
      if (m + 1 >= nsl) break;
      chunk_gather_lv_fixed(L, x, x32, cy, x32only);
      if (m + 3 < nsl) {
        chunk_locate(L, c0 + 192, nch, all64, cx);
        chunk_load_lv_fixed(A, Z0, c0 + 192, nch, cx);
      }
      if (FLAT) chunk_consume_lv_flat(L, cy, x32only); else chunk_consume_lv_fixed(L, cy, x32only);
      if (m + 2 >= nsl) break;
      chunk_gather_lv_fixed(L, x, x32, cz, x32only);
      if (m + 4 < nsl) {
        chunk_locate(L, c0 + 256, nch, all64, cy);
        chunk_load_lv_fixed(A, Z0, c0 + 256, nch, cy);
      }
      if (FLAT) chunk_consume_lv_flat(L, cz, x32only); else chunk_consume_lv_fixed(L, cz, x32only);
    }
  } else if (all64) {
    // Two-deep: slice m+1's loads need no metadata and go out right after
    // slice m is located.
    for (int m = 0; m < nsl; m += 2) {
      const int c0 = 64 * m + lane;
      chunk_locate(L, c0, nch, true, cx);
      if (m + 1 < nsl) chunk_load_rc(A, Z0, c0 + 64, nch, true, cy);
      chunk_gather(L, x, x32, true, cx);
      chunk_consume(L, cx, true);
      if (m + 1 >= nsl) break;
      chunk_locate(L, c0 + 64, nch, true, cy);
      if (m + 2 < nsl) chunk_load_rc(A, Z0, c0 + 128, nch, true, cx);
      chunk_gather(L, x, x32, true, cy);
      chunk_consume(L, cy, true);
    }
  } else {
    // Two-deep, per-tile levels: slice m+1 is located and its loads issued
    // before slice m's gathers.
    chunk_locate(L, lane, nch, false, cx);
    chunk_load_payload(A, Z0, lane, cx);
    for (int m = 0; m < nsl; m += 2) {
      const int c1 = 64 * (m + 1) + lane;
      if (m + 1 < nsl) {
        chunk_load_rc(A, Z0, c1, nch, false, cy);
        chunk_locate(L, c1, nch, false, cy);
        chunk_load_payload(A, Z0, c1, cy);
      }
      chunk_gather(L, x, x32, false, cx);
      chunk_consume(L, cx, false);
      if (m + 1 >= nsl) break;
      if (m + 2 < nsl) {
        chunk_load_rc(A, Z0, c1 + 64, nch, false, cx);
        chunk_locate(L, c1 + 64, nch, false, cx);
        chunk_load_payload(A, Z0, c1 + 64, cx);
      }
      chunk_gather(L, x, x32, false, cy);
      chunk_consume(L, cy, false);
    }
  }
}

// y rows of tile-rows [desc.tr0, desc.tr1) through epi(row, value).
//
// A descriptor whose tiles fit one window (kDescOneWindow, set by the host
// partition: at most kWinRows tile-rows, kWinTiles tiles and 4 * kWinChunks
// slots) is staged with every load it needs in flight at once: each tile's
// offset, column, tile-row, entry count, level and both scales, and the first
// slice's (row,col) bytes.  Other descriptors walk their tile-rows in windows
// that the limits cut, a tile-row's partial sum carried across a cut.
// ONEWIN = false compiles the walk only, for callers whose descriptors never
// carry the flag.  FLAT selects the branch-free per-level consume; the
// branching one keeps fewer values live, for kernels short of registers.
template <int MODE, bool COUNT, int DEPTH = 3, bool ONEWIN = true, bool FLAT = ONEWIN,
          typename Epi>
__device__ __forceinline__ void wave_spmv(const TileView &A, const WaveDesc &desc,
                                          const double *__restrict__ x,
                                          const float *__restrict__ x32,
                                          const uint8_t *__restrict__ seglev, int force64,
                                          WaveLds &L, int lane, PrecCount *cnt, Epi epi) {
  static_assert(kWinTiles == 128, "window staging covers tiles lane and lane + 64");
  const int tr0 = __builtin_amdgcn_readfirstlane(desc.tr0);
  const int tr1 = __builtin_amdgcn_readfirstlane(desc.tr1);
  if (tr0 >= tr1) return;
  const int T1 = __builtin_amdgcn_readfirstlane(desc.T1);
  const bool all64 = MODE == kAll64 || (force64 & 1);
  const bool x32only = MODE != kAll64 && !all64 && (force64 & 2);
  for (int i = lane; i < kWinRows * 16; i += 64) L.acc[i] = 0.0;
  int ta = __builtin_amdgcn_readfirstlane(desc.ta);
  int Za = __builtin_amdgcn_readfirstlane(desc.Z0);
  const int ja = lane, jb = lane + 64;
  if (ONEWIN && (__builtin_amdgcn_readfirstlane(desc.flags) & kDescOneWindow)) {
    const int nt = T1 - ta;
    const int Z0 = Za, Z1 = __builtin_amdgcn_readfirstlane(desc.Z1);
    const int nch = (Z1 - Z0) >> 2;
    int oa = 0, ob = 0, ca = 0, cb = 0, wa = 0, wb = 0, na = 0, nb = 0, la = 0, lb = 0;
    float ha = 0.0f, hb = 0.0f, qa = 0.0f, qb = 0.0f;
    if (ja < nt) {
      oa = A.tile_off[ta + ja];
      ca = A.tile_col[ta + ja];
      wa = A.tile_row[ta + ja];
      na = A.tile_nz[ta + ja];
      if (MODE != kAll64) {
        la = tile_level<MODE>(A, seglev, ta + ja, force64);
        ha = A.s16[ta + ja];
        qa = A.s8[ta + ja];
      }
    }
    if (jb < nt) {
      ob = A.tile_off[ta + jb];
      cb = A.tile_col[ta + jb];
      wb = A.tile_row[ta + jb];
      nb = A.tile_nz[ta + jb];
      if (MODE != kAll64) {
        lb = tile_level<MODE>(A, seglev, ta + jb, force64);
        hb = A.s16[ta + jb];
        qb = A.s8[ta + jb];
      }
    }
    // Tile j's payload ends where tile j + 1's begins (lane + 1's offset;
    // for lane 63's first tile, lane 0's second), the last tile's at Z1.
    const int oa_next = __shfl(oa, (lane + 1) & 63);
    const int ob_next = __shfl(ob, (lane + 1) & 63);
    const int ob_first = __shfl(ob, 0);
    const int ea = ja + 1 < nt ? (lane < 63 ? oa_next : ob_first) : Z1;
    const int eb = jb + 1 < nt ? ob_next : Z1;
    if constexpr (MODE != kAll64) {
      const LvGroups g = stage_levels<COUNT>(L, lane, nt, oa - Z0, ea - Z0, ca, wa - tr0, na, la,
                                             la == MF_PREC_FP16 ? ha : qa, ob - Z0, eb - Z0, cb,
                                             wb - tr0, nb, lb, lb == MF_PREC_FP16 ? hb : qb, cnt);
      __builtin_amdgcn_wave_barrier();
      level_groups(A, L, Z0, g, x, x32, all64, x32only, lane);
    } else {
      Chunk cx, cy;
      chunk_load_rc(A, Z0, lane, nch, all64, cx);
      if (ja < nt) stage_tile<MODE, COUNT>(L, ja, oa - Z0, ea - Z0, ca, wa - tr0, na, la, 0.0f, cnt);
      if (jb < nt) stage_tile<MODE, COUNT>(L, jb, ob - Z0, eb - Z0, cb, wb - tr0, nb, lb, 0.0f, cnt);
      __builtin_amdgcn_wave_barrier();
      if (nch > 0) window_slices<MODE, DEPTH, FLAT>(A, L, Z0, nch, x, x32, all64, x32only, lane, cx, cy);
    }
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < (tr1 - tr0) * 16; i += 64) epi(tr0 * 16 + i, L.acc[i]);
    __builtin_amdgcn_wave_barrier();
    return;
  }
  int ra = tr0;
  while (ra < tr1) {
    const int rlim = ra + kWinRows < tr1 ? ra + kWinRows : tr1;
    int tb = ta + kWinTiles < T1 ? ta + kWinTiles : T1;
    if (rlim < tr1) {
      const int tl = __builtin_amdgcn_readfirstlane(A.tile_ptr[rlim]);
      if (tl < tb) tb = tl;
    }
    int nt = tb - ta;
    int Zb = Za;
    if (nt > 0) {
      const int Z0 = Za;
      // Window metadata into registers: payload offsets of tiles lane,
      // lane + 64 and 128 (one past the last tile when the window is full).
      int oa = 0, ob = 0, oc = 0;
      if (ja <= nt) oa = A.tile_off[ta + ja];
      if (jb <= nt) ob = A.tile_off[ta + jb];
      if (nt == kWinTiles) oc = A.tile_off[ta + kWinTiles];
      // Keep the tiles whose payload ends within kWinChunks chunks (always
      // at least one: a tile holds at most 64 chunks).
      const int lim = Z0 + 4 * kWinChunks;
      const int fit = __builtin_popcountll(__ballot(ja >= 1 && ja <= nt && oa <= lim)) +
                      __builtin_popcountll(__ballot(jb <= nt && ob <= lim)) +
                      (nt == kWinTiles && oc <= lim ? 1 : 0);
      if (fit < nt) {
        nt = fit;
        tb = ta + nt;
      }
      const int Z1 = tb == T1 ? __builtin_amdgcn_readfirstlane(desc.Z1)
                              : __builtin_amdgcn_readfirstlane(A.tile_off[tb]);
      Zb = Z1;
      const int nch = (Z1 - Z0) >> 2;
      int ca = 0, cb = 0, wa = 0, wb = 0, na = 0, nb = 0, la = 0, lb = 0;
      float sa = 0.0f, sb = 0.0f;
      if (ja < nt) {
        ca = A.tile_col[ta + ja];
        wa = A.tile_row[ta + ja];
        na = A.tile_nz[ta + ja];
        if (MODE != kAll64) {
          la = tile_level<MODE>(A, seglev, ta + ja, force64);
          sa = la == MF_PREC_FP16 ? A.s16[ta + ja] : A.s8[ta + ja];
        }
      }
      if (jb < nt) {
        cb = A.tile_col[ta + jb];
        wb = A.tile_row[ta + jb];
        nb = A.tile_nz[ta + jb];
        if (MODE != kAll64) {
          lb = tile_level<MODE>(A, seglev, ta + jb, force64);
          sb = lb == MF_PREC_FP16 ? A.s16[ta + jb] : A.s8[ta + jb];
        }
      }
      // Metadata into LDS.  The payload end of tile j is the offset of tile
      // j + 1 (shuffled in uniform control flow: lane + 1's oa, for lane 63
      // lane 0's ob; lane + 1's ob, for lane 63 oc).
      const int na_next = __shfl(oa, (lane + 1) & 63);
      const int nb_next = __shfl(ob, (lane + 1) & 63);
      const int ob_first = __shfl(ob, 0);
      const int ea = lane < 63 ? na_next : ob_first;
      const int eb = lane < 63 ? nb_next : oc;
      if constexpr (MODE != kAll64) {
        const LvGroups g = stage_levels<COUNT>(L, lane, nt, oa - Z0, ea - Z0, ca, wa - ra, na, la, sa,
                                               ob - Z0, eb - Z0, cb, wb - ra, nb, lb, sb, cnt);
        __builtin_amdgcn_wave_barrier();
        level_groups(A, L, Z0, g, x, x32, all64, x32only, lane);
      } else {
        // The first slice's loads (no metadata needed) go out first.
        Chunk cx, cy;
        chunk_load_rc(A, Z0, lane, nch, all64, cx);
        if (ja < nt) stage_tile<MODE, COUNT>(L, ja, oa - Z0, ea - Z0, ca, wa - ra, na, la, sa, cnt);
        if (jb < nt) stage_tile<MODE, COUNT>(L, jb, ob - Z0, eb - Z0, cb, wb - ra, nb, lb, sb, cnt);
        __builtin_amdgcn_wave_barrier();
        window_slices<MODE, DEPTH, FLAT>(A, L, Z0, nch, x, x32, all64, x32only, lane, cx, cy);
      }
      __builtin_amdgcn_wave_barrier();
    }
    // Tile-rows [ra, R) are complete; tile-row R has started iff tile_ptr[R] < tb.
    int R;
    if (tb < T1) {
      R = __builtin_amdgcn_readfirstlane(A.tile_row[tb]);
      if (R > rlim) R = rlim;
    } else {
      R = rlim;
    }
    for (int i = lane; i < (R - ra) * 16; i += 64) epi(ra * 16 + i, L.acc[i]);
    const bool partial = R < tr1 && A.tile_ptr[R] < tb;
    double carry = 0.0;
    if (partial && lane < 16) carry = L.acc[(R - ra) * 16 + lane];
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < kWinRows * 16; i += 64) L.acc[i] = 0.0;
    __builtin_amdgcn_wave_barrier();
    if (partial && lane < 16) L.acc[lane] = carry;
    __builtin_amdgcn_wave_barrier();
    ra = R;
    ta = tb;
    Za = Zb;
  }
}

// Tile-rows [tr0, tr1) (descriptor loaded from the tile pointers).
template <int MODE, bool COUNT, int DEPTH = 3, typename Epi>
__device__ __forceinline__ void wave_spmv(const TileView &A, int tr0, int tr1,
                                          const double *__restrict__ x,
                                          const float *__restrict__ x32,
                                          const uint8_t *__restrict__ seglev, int force64,
                                          WaveLds &L, int lane, PrecCount *cnt, Epi epi) {
  if (tr0 >= tr1) return;
  wave_spmv<MODE, COUNT, DEPTH, false>(A, wave_desc(A, tr0, tr1), x, x32, seglev, force64, L, lane,
                                       cnt, epi);
}

// Heavy tile-rows cut into pieces: 16 partial row sums per piece, and per
// heavy tile-row an arrival counter (zero between SpMVs), the piece count and
// the first piece's slot.
struct HeavyView {
  double *part;
  unsigned *cnt;
  const int *n;
  const int *base;
};

// One descriptor: a tile-row range through epi(row, value), or a piece of a
// heavy tile-row.  Piece k of heavy tile-row h stores its 16 partial row sums
// in its own slot (agent-scope atomic stores, visible across XCDs) and counts
// itself done; the piece that completes the count (the last to finish, so no
// wave ever waits for another) adds the slots in a fixed order
// (deterministic), resets the counter for the next SpMV and runs epi for the
// tile-row's 16 rows.
template <int MODE, bool COUNT, int DEPTH, bool ONEWIN, bool FLAT, typename Epi>
__device__ __forceinline__ void wave_spmv_piece(const TileView &A, const WaveDesc &desc,
                                                const HeavyView &hv,
                                                const double *__restrict__ x,
                                                const float *__restrict__ x32,
                                                const uint8_t *__restrict__ seglev, int force64,
                                                WaveLds &L, int lane, PrecCount *cnt, Epi epi) {
  const int flags = __builtin_amdgcn_readfirstlane(desc.flags);
  const bool piece = flags & kDescPiece;
  const int h = piece ? __builtin_amdgcn_readfirstlane(desc.aux) : 0;
  const int base = piece ? __builtin_amdgcn_readfirstlane(hv.base[h]) : 0;
  double *slot = hv.part + 16 * (size_t)(base + (flags >> 8));
  wave_spmv<MODE, COUNT, DEPTH, ONEWIN, FLAT>(A, desc, x, x32, seglev, force64, L, lane, cnt,
    [&](const int i, const double v) {
      if (piece)
        __hip_atomic_store(slot + (i & 15), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      else
        epi(i, v);
    });
  if (!piece) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  unsigned done = 0;
  if (lane == 0)
    done = __hip_atomic_fetch_add(hv.cnt + h, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
  done = __builtin_amdgcn_readfirstlane(done);
  const int np = __builtin_amdgcn_readfirstlane(hv.n[h]);
  if (done != (unsigned)np) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  // lane (q, row): pieces q, q+4, ... of the row, then the four partial sums
  // in q order
  const int row = lane & 15, q = lane >> 4;
  double sum = 0.0;
  for (int k = q; k < np; k += 4)
    sum += __hip_atomic_load(hv.part + 16 * (size_t)(base + k) + row, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  const double s1 = __shfl(sum, row + 16), s2 = __shfl(sum, row + 32), s3 = __shfl(sum, row + 48);
  const int R = __builtin_amdgcn_readfirstlane(desc.tr0);
  if (lane < 16) epi(16 * R + lane, ((sum + s1) + s2) + s3);
  if (lane == 0) __hip_atomic_store(hv.cnt + h, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Row-group tiled SpMV.
//
// Lane (g, r) = (lane / 16, lane % 16) computes row r of one tile-row R (the
// caller's choice; normally four consecutive tile-rows per wave, so lane L
// holds row 64k + L of a 64-row block).  The lane walks the tiles of its
// tile-row in batches of kRgBatch: the batch's per-tile data (payload offset,
// column, the row's start and end bytes, level) is loaded first, then the
// first kRgUnroll entries of the row in every tile of the batch (payload and
// column nibble, then the x gathers), so up to kRgBatch*kRgUnroll independent
// load chains are in flight per lane; longer rows finish in a loop.  The row
// sum stays in a register (no LDS, no atomics) and is added in a fixed order.
// Lanes of tile-rows with fewer tiles idle while the wave finishes the
// longest tile-row of the four: the kernel for matrices whose tile-rows hold
// similar tile counts (stencils, meshes); irregular ones use wave_spmv.
// ---------------------------------------------------------------------------
constexpr int kRgBatch = 4;
constexpr int kRgUnroll = 2;

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m));
  return v;
}

template <int MODE, bool COUNT>
__device__ __forceinline__ double rowgroup_row(const TileView &A, int R, bool valid, int r,
                                               const double *__restrict__ x,
                                               const float *__restrict__ x32,
                                               const uint8_t *__restrict__ seglev, int force64,
                                               PrecCount *cnt) {
  int t0 = 0, nt = 0;
  if (valid) {
    t0 = A.tile_ptr[R];
    nt = A.tile_ptr[R + 1] - t0;
  }
  const int ntw = wave_max(nt);
  double acc = 0.0;
  for (int i = 0; i < ntw; i += kRgBatch) {
    int off[kRgBatch], xo[kRgBatch], b[kRgBatch], e[kRgBatch], lv[kRgBatch];
    float sc[kRgBatch];
#pragma unroll
    for (int u = 0; u < kRgBatch; ++u) {
      off[u] = xo[u] = b[u] = e[u] = 0;
      lv[u] = MF_PREC_FP64;
      sc[u] = 1.0f;
      if (i + u < nt) {
        const int t = t0 + i + u;
        off[u] = A.tile_off[t];
        const int nz = A.tile_nz[t] + 1;
        xo[u] = A.tile_col[t] << 4;
        b[u] = A.rp16[(size_t)t * 16 + r];
        e[u] = r < 15 ? A.rp16[(size_t)t * 16 + r + 1] : nz;
        if (MODE != kAll64) {
          lv[u] = tile_level<MODE>(A, seglev, t, force64);
          sc[u] = lv[u] == MF_PREC_FP16 ? A.s16[t] : A.s8[t];
          if (COUNT && r == 0) cnt->add(lv[u]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kRgUnroll; ++q)
#pragma unroll
      for (int u = 0; u < kRgBatch; ++u)
        if (b[u] + q < e[u]) {
          const int k = off[u] + b[u] + q;
          acc += entry_product<MODE>(A, lv[u], k, xo[u] + (A.rc[k] & 15), x, x32, sc[u]);
        }
#pragma unroll
    for (int u = 0; u < kRgBatch; ++u)
      for (int k = off[u] + b[u] + kRgUnroll; k < off[u] + e[u]; ++k)
        acc += entry_product<MODE>(A, lv[u], k, xo[u] + (A.rc[k] & 15), x, x32, sc[u]);
  }
  return acc;
}

// K passes of four tile-rows: w[k] = row 16*tr0 + 64k + lane of A x (0 past
// tr1).  The wave's state layout (lane L holds rows L + 64k) is the pass
// layout, so the products land in the registers that need them.
template <int MODE, bool COUNT, int K>
__device__ __forceinline__ void rowgroup_passes(const TileView &A, int tr0, int tr1, int lane,
                                                const double *__restrict__ x,
                                                const float *__restrict__ x32,
                                                const uint8_t *__restrict__ seglev, int force64,
                                                PrecCount *cnt, double (&w)[K]) {
  const int g = lane >> 4, r = lane & 15;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    w[k] = 0.0;
    if (tr0 + 4 * k < tr1) {  // wave-uniform
      const int R = tr0 + 4 * k + g;
      w[k] = rowgroup_row<MODE, COUNT>(A, R, R < tr1, r, x, x32, seglev, force64, cnt);
    }
  }
}

// Learned tile levels for tiles [t0, t1): the MLP runs on 16 tiles per MFMA
// batch and the levels go to ic.prec_out, where the next per-tile SpMV pass
// of the same wave reads them (the closing workgroup fence orders the stores
// before those loads).  Callers infer exactly the tiles their own SpMV reads,
// so no level is written while another wave may read it.  Returns (alike in
// every lane) whether any of the tiles changed level.
__device__ __forceinline__ bool wave_infer_tiles(const TileView &A, const InferCtx &ic, int t0, int t1,
                                                 int lane) {
  bool changed = false;
  if (t0 >= t1) return false;
  int narrow = MF_PREC_SKIP;  // most precise FP32 / FP16 / INT8 class seen
  MlpRegs wr;
  mlp_load(ic.w, lane, wr);
  // features of the next batch are loaded while this one is evaluated
  int t = t0 + (lane & 15);
  bool valid = t < t1;
  half4_t f = infer_features(ic, valid ? t : t0, (lane >= 16 && valid) ? A.tile_col[t] : 0, lane, valid);
  for (int g = t0; g < t1; g += 16) {
    const int tn = g + 16 + (lane & 15);
    const bool vn = tn < t1;
    const half4_t fn = infer_features(ic, vn ? tn : t0, (lane >= 16 && vn) ? A.tile_col[tn] : 0, lane, vn);
    const float4_t z = mlp_eval16r(wr, f);
    f = fn;
    if (lane < 16 && valid) {
      // the class and the tile's previous level, packed until the second pass
      const int cls = mlp_argmax(z);
      if (cls >= MF_PREC_FP32 && cls <= MF_PREC_INT8) narrow = min(narrow, cls);
      ic.prec_out[t] = (uint8_t)(cls | (ic.prec_out[t] << 4));
    }
    t = tn;
    valid = vn;
  }
  // Every FP32 / FP16 / INT8 class of the wave's tiles runs at the most
  // precise of those classes: one level group per window instead of two or
  // three.  Measured on MI355X, CG on the 3D 7-point 128^3 Laplacian to
  // 1e-10 (448 iterations each): every tile FP32 19.3 ms, the FP32 diagonal /
  // INT8 off-diagonal split that the classes give 21.7 ms, FP64 24.0 ms; on
  // the 100^3 convection-diffusion matrix, whose tiles are all INT8, BiCGSTAB
  // with every tile INT8 12.3 ms and every tile FP32 13.2 ms.  FP64 and SKIP
  // run as predicted.
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) narrow = min(narrow, __shfl_xor(narrow, m));
  __threadfence_block();  // the first pass's level bytes, for every lane
  for (int u = t0 + lane; u < t1; u += kWave) {
    const int v = ic.prec_out[u];
    const int cls = v & 15, old = v >> 4;
    const int lv = cls >= MF_PREC_FP32 && cls <= MF_PREC_INT8 ? narrow : cls;
    changed = changed || lv != old;
    ic.prec_out[u] = (uint8_t)lv;
  }
  __threadfence_block();
  return __ballot(changed) != 0;
}

// Learned tile levels for the tiles of tile-rows [tr0, tr1).
__device__ __forceinline__ bool wave_infer(const TileView &A, const InferCtx &ic, int tr0, int tr1,
                                           int lane) {
  if (tr0 >= tr1) return false;
  return wave_infer_tiles(A, ic, A.tile_ptr[tr0], A.tile_ptr[tr1], lane);
}

// Deterministic sum over the 64 lanes of a wave (fixed xor tree); every lane
// gets the result.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

// A wave-uniform double moved to scalar registers, so that loop state which
// every lane holds alike does not occupy vector registers.
__device__ __forceinline__ double uniform_double(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Deterministic sum over the block: wave sums, then waves in index order.
// `lds` must hold blockDim.x/64 doubles.  Every thread gets the result.
__device__ __forceinline__ double block_sum(double v, double *lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; ++i) s += lds[i];
  return uniform_double(s);  // every lane summed the same values in the same order
}

// Grid synchronisation for a co-resident grid.
//
// Every sync point has an ordinal (epoch, counted identically by all
// blocks).  Publication: every wave first drains its own outstanding memory
// operations (s_waitcnt vmcnt(0)), the block meets at s_barrier, and thread 0
// alone issues the agent-scope release (one L2 write-back per block) and
// stores the epoch in the block's own arrival word.  Block 0 watches all
// arrival words (one lane per block), then publishes the epoch in `flag`;
// every other block's thread 0 polls that one word and issues the single
// agent-scope acquire (L1/L2 invalidate) for its block before the block
// barrier releases the other waves.  Measured on MI355X with one 1024-thread
// block per CU: about 5 us per sync, against 30-120 us when every wave
// fences and 4-7 us more when all blocks hit one atomic counter.
//
// Every wait is bounded: after kSpinLimit polls a waiter raises `abort` and
// returns, so a block that is not resident (or any lost arrival) ends the
// kernel with an error the host reports, never a hang.
constexpr int kMaxBlocks = 4096;

struct GridSync {
  unsigned flag;
  unsigned pad0[63];
  unsigned abort;
  unsigned pad1[63];
  unsigned arrive[kMaxBlocks];
};

constexpr unsigned kSpinLimit = 1u << 22;

__device__ __forceinline__ bool wait_word(unsigned *w, GridSync *gs, unsigned epoch) {
  unsigned n = 0;
  while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
    if (++n > kSpinLimit) {
      __hip_atomic_store(&gs->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    if ((n & 1023u) == 0 &&
        __hip_atomic_load(&gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return false;
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

__device__ __forceinline__ void drain_own_memory_ops() {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// Arrival + master wait + release of the flag.  On return the whole grid has
// passed the sync point and this block has acquired.  `master_work` runs in
// block 0 after every arrival has been seen and before the flag is
// published (used by grid_sum to add the partials).
template <typename F>
__device__ __forceinline__ void grid_arrive_wait(GridSync *gs, unsigned &epoch, F master_work) {
  drain_own_memory_ops();
  __syncthreads();
  ++epoch;
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_store(&gs->arrive[blockIdx.x], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (blockIdx.x == 0) {
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
      if (!wait_word(&gs->arrive[b], gs, epoch)) break;
    if (threadIdx.x < (int)gridDim.x) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __syncthreads();
    master_work();
    drain_own_memory_ops();
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_store(&gs->flag, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  } else if (threadIdx.x == 0) {
    wait_word(&gs->flag, gs, epoch);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Plain grid barrier.
__device__ __forceinline__ void grid_sync(GridSync *gs, unsigned &epoch, int *lds_i) {
  (void)lds_i;
  grid_arrive_wait(gs, epoch, [] {});
}

// Sum over the block of NV values per thread: fixed-order wave trees, then
// the waves' sums in wave order.  `lds` must hold NV * blockDim.x/64
// doubles.  Every thread gets the NV totals.
template <int NV>
__device__ __forceinline__ void block_sum_n(double (&v)[NV], double *lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) lds[i * nw + w] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
    for (int k = 0; k < nw; ++k) s += lds[i * nw + k];
    v[i] = uniform_double(s);  // every lane summed the same values in the same order
  }
}

// Barrier of the block's waves for LDS traffic only: it waits for the
// wave's own LDS operations, not for its outstanding global stores (which a
// full __syncthreads would drain).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// block_sum_n with LDS-only barriers (callers whose global stores need no
// ordering against the sum).
template <int NV>
__device__ __forceinline__ void block_sum_lds(double (&v)[NV], double *lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  lds_barrier();
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) lds[i * nw + w] = v[i];
  }
  lds_barrier();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
    for (int k = 0; k < nw; ++k) s += lds[i * nw + k];
    v[i] = uniform_double(s);
  }
}

// Grid-wide deterministic sum of NV values per thread, also a full barrier.
// Each block writes its NV block sums to partials[slot][block][i] and
// arrives; once block 0 has seen every arrival and published the barrier,
// every block reads all the partials and adds them in block order with the
// same fixed-order tree, so every block gets the same totals without a
// second broadcast.  The two slots alternate so that storage is never
// reused while a slower block may still read it (two sync points always
// separate two uses of one slot).  `lds` must hold NV * blockDim.x/64
// doubles.
constexpr int kMaxGridSum = 4;

template <int NV>
__device__ __forceinline__ void grid_sum(double (&v)[NV], double *partials, GridSync *gs,
                                         unsigned &epoch, int &slot, double *lds, int *lds_i) {
  static_assert(NV <= kMaxGridSum, "grid_sum width");
  (void)lds_i;
  const int G = gridDim.x;
  block_sum_n<NV>(v, lds);
  double *buf = partials + (size_t)slot * kMaxGridSum * G;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) buf[blockIdx.x * NV + i] = v[i];
  }
  grid_arrive_wait(gs, epoch, [] {});
  double s[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) s[i] = 0.0;
  for (int b = threadIdx.x; b < G; b += blockDim.x) {
    double t[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) t[i] = buf[b * NV + i];
#pragma unroll
    for (int i = 0; i < NV; ++i) s[i] += t[i];
  }
  block_sum_n<NV>(s, lds);
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = s[i];
  slot ^= 1;
}

// ---------------------------------------------------------------------------
// Value-only grid sum with epoch-tagged mailboxes.
//
// Every 64-bit mailbox word carries the sync point's epoch in its high half
// and 32 bits of a value in its low half (64-bit agent-scope atomic stores
// and loads), so a reader that sees the epoch in every word of a mailbox
// holds the whole value, with no memory fence, no drain of other stores and
// no second round trip.  Each block posts its NV block sums in its mailbox;
// block 0 reads every mailbox and adds them in block order with a fixed
// tree, and posts the totals in the result mailbox, which every other
// block's thread 0 polls.  This orders no other memory: it is used where the
// next phase reads only the calling wave's own rows (the rows its SpMV
// epilogue wrote), and a full barrier (grid_sync, grid_sum) precedes every
// phase that reads other waves' data.  Waits are bounded as in wait_word.
// ---------------------------------------------------------------------------
constexpr int kMboxWords = 2 * kMaxGridSum;  // one 64-byte line per block

template <int NV>
__device__ __forceinline__ void mbox_post(unsigned long long *mb, unsigned epoch, const double (&v)[NV]) {
  const unsigned long long e = (unsigned long long)epoch << 32;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v[i]);
    __hip_atomic_store(mb + 2 * i, e | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(mb + 2 * i + 1, e | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <int NV>
__device__ __forceinline__ bool mbox_take(const unsigned long long *mb, unsigned epoch, GridSync *gs,
                                          double (&v)[NV]) {
  unsigned long long w[2 * NV];
  unsigned n = 0;
  for (;;) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 2 * NV; ++k) {
      w[k] = __hip_atomic_load(mb + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ok = ok && (unsigned)(w[k] >> 32) == epoch;
    }
    if (ok) break;
    if (++n > kSpinLimit) {
      __hip_atomic_store(&gs->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    if ((n & 1023u) == 0 && __hip_atomic_load(&gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return false;
    __builtin_amdgcn_s_sleep(1);
  }
#pragma unroll
  for (int i = 0; i < NV; ++i)
    v[i] = __builtin_bit_cast(double, ((w[2 * i + 1] & 0xffffffffull) << 32) | (w[2 * i] & 0xffffffffull));
  return true;
}

// mbox: (gridDim.x + 1) * kMboxWords words, zero before the launch.
template <int NV>
__device__ __forceinline__ void grid_sum_values(double (&v)[NV], unsigned long long *mbox, GridSync *gs,
                                                unsigned &epoch, double *lds) {
  static_assert(NV <= kMaxGridSum, "grid_sum_values width");
  __shared__ double bcast[kMaxGridSum];
  const int G = gridDim.x;
  block_sum_lds<NV>(v, lds);
  ++epoch;
  if (threadIdx.x == 0) mbox_post<NV>(mbox + (size_t)blockIdx.x * kMboxWords, epoch, v);
  unsigned long long *res = mbox + (size_t)G * kMboxWords;
  if (blockIdx.x == 0) {
    double s[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) s[i] = 0.0;
    for (int b = threadIdx.x; b < G; b += blockDim.x) {
      double t[NV];
      if (!mbox_take<NV>(mbox + (size_t)b * kMboxWords, epoch, gs, t)) break;
#pragma unroll
      for (int i = 0; i < NV; ++i) s[i] += t[i];
    }
    block_sum_lds<NV>(s, lds);
    if (threadIdx.x == 0) mbox_post<NV>(res, epoch, s);
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = s[i];
  } else {
    if (threadIdx.x == 0) {
      double t[NV];
#pragma unroll
      for (int i = 0; i < NV; ++i) t[i] = 0.0;
      mbox_take<NV>(res, epoch, gs, t);
#pragma unroll
      for (int i = 0; i < NV; ++i) bcast[i] = t[i];
    }
    lds_barrier();
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = uniform_double(bcast[i]);
  }
  lds_barrier();  // bcast and lds are free again
}

}  // namespace mf
