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
    Chunk cx, cy;
    chunk_load_rc(A, Z0, lane, nch, all64, cx);
    // Tile j's payload ends where tile j + 1's begins (lane + 1's offset;
    // for lane 63's first tile, lane 0's second), the last tile's at Z1.
    const int oa_next = __shfl(oa, (lane + 1) & 63);
    const int ob_next = __shfl(ob, (lane + 1) & 63);
    const int ob_first = __shfl(ob, 0);
    if (ja < nt) {
      const int e = ja + 1 < nt ? (lane < 63 ? oa_next : ob_first) : Z1;
      stage_tile<MODE, COUNT>(L, ja, oa - Z0, e - Z0, ca, wa - tr0, na, la,
                              la == MF_PREC_FP16 ? ha : qa, cnt);
    }
    if (jb < nt) {
      const int e = jb + 1 < nt ? ob_next : Z1;
      stage_tile<MODE, COUNT>(L, jb, ob - Z0, e - Z0, cb, wb - tr0, nb, lb,
                              lb == MF_PREC_FP16 ? hb : qb, cnt);
    }
    __builtin_amdgcn_wave_barrier();
    if (nch > 0) window_slices<MODE, DEPTH, FLAT>(A, L, Z0, nch, x, x32, all64, x32only, lane, cx, cy);
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
      // The first slice's loads that need no metadata go out now.
      Chunk cx, cy;
      chunk_load_rc(A, Z0, lane, nch, all64, cx);
      // Metadata into LDS.  The payload end of tile j is the offset of tile
      // j + 1 (shuffled in uniform control flow: lane + 1's oa, for lane 63
      // lane 0's ob; lane + 1's ob, for lane 63 oc).
      const int na_next = __shfl(oa, (lane + 1) & 63);
      const int nb_next = __shfl(ob, (lane + 1) & 63);
      const int ob_first = __shfl(ob, 0);
      if (ja < nt)
        stage_tile<MODE, COUNT>(L, ja, oa - Z0, (lane < 63 ? na_next : ob_first) - Z0, ca,
                                wa - ra, na, la, sa, cnt);
      if (jb < nt)
        stage_tile<MODE, COUNT>(L, jb, ob - Z0, (lane < 63 ? nb_next : oc) - Z0, cb, wb - ra, nb,
                                lb, sb, cnt);
      __builtin_amdgcn_wave_barrier();
      window_slices<MODE, DEPTH, FLAT>(A, L, Z0, nch, x, x32, all64, x32only, lane, cx, cy);
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

// Learned tile levels for the wave's tile-rows [tr0, tr1): the MLP runs on 16
// tiles per MFMA batch and the levels go to ic.prec_out, where the next
// per-tile SpMV pass of the same wave reads them (the closing workgroup
// fence orders the stores before those loads).
__device__ __forceinline__ void wave_infer(const TileView &A, const InferCtx &ic, int tr0, int tr1,
                                           int lane) {
  if (tr0 >= tr1) return;
  const int t0 = A.tile_ptr[tr0], t1 = A.tile_ptr[tr1];
  for (int g = t0; g < t1; g += 16) {
    const int t = g + (lane & 15);
    const bool valid = t < t1;
    const int tt = valid ? t : t0;
    const half4_t f = infer_features(ic, tt, (lane >= 16 && valid) ? A.tile_col[tt] : 0, lane, valid);
    const float4_t z = mlp_eval16(ic.w, lane, f);
    if (lane < 16 && valid) ic.prec_out[t] = (uint8_t)mlp_argmax(z);
  }
  __threadfence_block();
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
  double result[2][8];
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

// Grid-wide deterministic sum of NV values per thread, also a full barrier.
// Each block writes its block sums to partials[slot][i][block]; block 0,
// once every arrival is in, adds them in block order with one wave (strided
// over the lanes, then a fixed xor tree) and writes result[slot]; every block
// then reads the same totals.  The two slots alternate so that storage is
// never reused while a slower block may still read it (two sync points
// always separate two uses of one slot).
constexpr int kMaxGridSum = 4;

template <int NV>
__device__ __forceinline__ void grid_sum(double (&v)[NV], double *partials, GridSync *gs,
                                         unsigned &epoch, int &slot, double *lds, int *lds_i) {
  static_assert(NV <= kMaxGridSum, "grid_sum width");
  (void)lds_i;
  const int G = gridDim.x;
  double bs[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) bs[i] = block_sum(v[i], lds);
  double *buf = partials + (size_t)slot * kMaxGridSum * G;
  double *res = gs->result[slot];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) buf[i * G + blockIdx.x] = bs[i];
  }
  grid_arrive_wait(gs, epoch, [&] {
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        double s = 0.0;
        for (int b = lane; b < G; b += 64) s += buf[i * G + b];
        s = wave_sum(s);
        if (lane == 0) res[i] = s;
      }
    }
  });
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = uniform_double(res[i]);
  slot ^= 1;
}

}  // namespace mf
