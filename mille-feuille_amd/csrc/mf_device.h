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
  const uint8_t *rc;    // nnz: (row << 4) | col inside the tile
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
  return TileView{d.tile_ptr, d.tile_col, d.tile_off, d.tile_row, d.rc,   d.rp16,
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
// tiles and of payload entries (tiles are stored tile-row by tile-row).  It
// walks that range in windows of at most kWinTiles tiles spanning at most
// kWinRows tile-rows.  Per window the tile metadata (payload offset, column,
// tile-row, precision, scale) is staged in wave-private LDS, then the window's
// entries are streamed: lane L takes the four consecutive entries
// base+4L..base+4L+3, so one wave instruction covers 256 entries and the
// values / (row,col) bytes arrive as 32-, 16- or 4-byte vector loads.  Each
// lane finds the tile of its first entry by a binary search over the staged
// offsets (overlapping the HBM loads), gathers x, and adds its products into
// per-row FP64 accumulators in LDS, merging consecutive entries of one row in
// registers first.  When a window ends, every tile-row it completed is
// handed to the epilogue row by row (one lane per row); an unfinished last
// tile-row keeps its accumulators for the next window.
//
// Mixed precision: FP64 tiles multiply x; FP32 / FP16 / INT8 tiles multiply
// the FP32 copy x32 with FP32 products (FP16/INT8 scaled by the tile's power
// of two) that are accumulated in FP64; SKIP tiles contribute nothing.
// force64 != 0 treats every tile as FP64 (the solver's fallback).
// ---------------------------------------------------------------------------

struct WaveLds {
  int toff[kWinTiles + 1];  // payload offset of each tile, relative to the window
  int tcol[kWinTiles];      // 16 * tile column
  float tscale[kWinTiles];  // FP16 / INT8 scale of the tile's current level
  uint8_t trow[kWinTiles];  // tile-row relative to the window's first tile-row
  uint8_t tprec[kWinTiles];
  double acc[kWinRows * 16];
};

template <int MODE>
__device__ __forceinline__ int tile_level(const TileView &A, const uint8_t *seglev, int t,
                                          int force64) {
  if (MODE == kAll64 || force64) return MF_PREC_FP64;
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

// One lane's share of a slice: the four consecutive entries k0..k0+3 of the
// window.  The tile of every entry is found from the staged offsets alone
// (before any payload arrives), so the payload is loaded at the right level
// at once: as one vector load when the valid entries share one level (FP64:
// 8 words, FP32: 4, FP16: 2, INT8: 1, packed), else entry by entry into
// words 2u, 2u+1 ("split").  x is gathered once the (row,col) bytes are in.
struct Chunk {
  uint32_t rc4;
  uint32_t jp;    // window-relative tile of entry u in byte u
  int vmask;      // bit u: entry u lies in the window; 0: no load issued
  int lvl;        // common level of the valid entries; -1: split
  uint4 a, b;     // raw payload words 0-3 and 4-7
  double x0, x1, x2, x3;  // x (FP64 entries) or x32 (other levels) at each entry's column
};

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

__device__ __forceinline__ void chunk_set_words(Chunk &c, int u, uint32_t lo, uint32_t hi) {
  switch (u) {
    case 0: c.a.x = lo; c.a.y = hi; break;
    case 1: c.a.z = lo; c.a.w = hi; break;
    case 2: c.b.x = lo; c.b.y = hi; break;
    default: c.b.z = lo; c.b.w = hi; break;
  }
}

// Does the chunk of lane `lane` in the slice at `base` hold a window entry?
__device__ __forceinline__ bool chunk_live(int base, int lane, int Z0, int Z1) {
  const int k0 = base + 4 * lane;
  return k0 + 3 >= Z0 && k0 < Z1;
}

// Stage 1a (FP64 payload): the (row,col) bytes and the payload, which need no
// window metadata, so they can be issued before the window is staged.
__device__ __forceinline__ void chunk_load64(const TileView &A, int base, int lane, int Z0,
                                             int Z1, Chunk &c) {
  if (!chunk_live(base, lane, Z0, Z1)) return;
  const int k0 = base + 4 * lane;
  c.rc4 = *(const uint32_t *)(A.rc + k0);
  const uint4 *q = (const uint4 *)(A.v64 + k0);
  c.a = q[0];
  c.b = q[1];
}

// Stage 1b: the tile of every entry, from the staged offsets alone.
__device__ __forceinline__ void chunk_locate(const WaveLds &L, int base, int lane, int Z0,
                                             int Z1, int nt, Chunk &c) {
  c.vmask = 0;
  c.lvl = MF_PREC_FP64;
  if (!chunk_live(base, lane, Z0, Z1)) return;
  const int k0 = base + 4 * lane;
  // tile of the first valid entry: binary search over the staged offsets
  const int kf = (k0 > Z0 ? k0 : Z0) - Z0;
  int lo = 0, hi = nt - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.toff[mid] <= kf) lo = mid;
    else hi = mid - 1;
  }
  int j = lo;
  uint32_t jp = 0;
  int vm = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int k = k0 + u - Z0;
    if (k >= 0 && k < Z1 - Z0) {
      while (L.toff[j + 1] <= k) ++j;
      vm |= 1 << u;
    }
    jp |= (uint32_t)j << (8 * u);
  }
  c.jp = jp;
  c.vmask = vm;
}

// Stage 1a (per-tile levels): the (row,col) bytes, which need no metadata.
__device__ __forceinline__ void chunk_load_rc(const TileView &A, int base, int lane, int Z0,
                                              int Z1, Chunk &c) {
  if (!chunk_live(base, lane, Z0, Z1)) return;
  c.rc4 = *(const uint32_t *)(A.rc + base + 4 * lane);
}

// Stage 1c (per-tile levels, after chunk_locate): the payload at the level
// of the chunk's tiles.
__device__ __forceinline__ void chunk_load_levels(const TileView &A, const WaveLds &L, int base,
                                                  int lane, Chunk &c) {
  const int vm = c.vmask;
  if (!vm) return;
  const int k0 = base + 4 * lane;
  const uint32_t jp = c.jp;
  // common level of the valid entries
  int lv = -2;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (vm >> u & 1) {
      const int l = L.tprec[(jp >> (8 * u)) & 255];
      lv = lv == -2 ? l : (lv == l ? lv : -1);
    }
  }
  c.lvl = lv;
  if (lv == MF_PREC_FP64) {
    const uint4 *q = (const uint4 *)(A.v64 + k0);
    c.a = q[0];
    c.b = q[1];
  } else if (lv == MF_PREC_FP32) {
    c.a = *(const uint4 *)(A.v32 + k0);
  } else if (lv == MF_PREC_FP16) {
    const uint2 h = *(const uint2 *)(A.v16 + k0);
    c.a.x = h.x;
    c.a.y = h.y;
  } else if (lv == MF_PREC_INT8) {
    c.a.x = *(const uint32_t *)(A.v8 + k0);
  } else if (lv == -1) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!(vm >> u & 1)) continue;
      const int l = L.tprec[(jp >> (8 * u)) & 255];
      const int k = k0 + u;
      uint32_t w0 = 0, w1 = 0;
      if (l == MF_PREC_FP64) {
        const uint2 v = *(const uint2 *)(A.v64 + k);
        w0 = v.x;
        w1 = v.y;
      } else if (l == MF_PREC_FP32) {
        w0 = __builtin_bit_cast(uint32_t, A.v32[k]);
      } else if (l == MF_PREC_FP16) {
        w0 = A.v16[k];
      } else if (l == MF_PREC_INT8) {
        w0 = (uint32_t)(uint8_t)A.v8[k];
      }
      chunk_set_words(c, u, w0, w1);
    }
  }
}

// Stage 2 (needs the (row,col) bytes): gather x at every valid entry's column.
__device__ __forceinline__ void chunk_gather(const WaveLds &L, const double *__restrict__ x,
                                             const float *__restrict__ x32, bool all64,
                                             Chunk &c) {
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (!(c.vmask >> u & 1)) continue;
    const int j = (c.jp >> (8 * u)) & 255;
    const int xi = L.tcol[j] + ((c.rc4 >> (8 * u)) & 15);
    if (all64) {
      chunk_set_x(c, u, x[xi]);
    } else {
      const int l = L.tprec[j];
      if (l == MF_PREC_FP64) chunk_set_x(c, u, x[xi]);
      else if (l != MF_PREC_SKIP) chunk_set_x(c, u, (double)x32[xi]);
    }
  }
}

// Stage 3 (needs payload and x): products, merged per row in registers and
// added into the window's LDS row accumulators.
__device__ __forceinline__ void chunk_consume(WaveLds &L, const Chunk &c, bool all64) {
  int key = -1;
  double run = 0.0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (!(c.vmask >> u & 1)) continue;
    const int j = (c.jp >> (8 * u)) & 255;
    const int kk = L.trow[j] * 16 + ((c.rc4 >> (8 * u + 4)) & 15);
    double prod;
    if (all64) {
      prod = __builtin_bit_cast(double, ((uint64_t)chunk_word(c, 2 * u + 1) << 32) |
                                            chunk_word(c, 2 * u)) * chunk_x(c, u);
    } else {
      const int l = L.tprec[j];
      const bool split = c.lvl < 0;
      if (l == MF_PREC_FP64) {
        prod = __builtin_bit_cast(double, ((uint64_t)chunk_word(c, 2 * u + 1) << 32) |
                                              chunk_word(c, 2 * u)) * chunk_x(c, u);
      } else if (l == MF_PREC_FP32) {
        const uint32_t w = split ? chunk_word(c, 2 * u) : chunk_word(c, u);
        prod = (double)(__builtin_bit_cast(float, w) * (float)chunk_x(c, u));
      } else if (l == MF_PREC_FP16) {
        const uint16_t h = split ? (uint16_t)chunk_word(c, 2 * u)
                                 : (uint16_t)(chunk_word(c, u >> 1) >> (16 * (u & 1)));
        prod = (double)(half_bits_to_float(h) * (float)chunk_x(c, u) * L.tscale[j]);
      } else if (l == MF_PREC_INT8) {
        const int8_t q = split ? (int8_t)chunk_word(c, 2 * u) : (int8_t)(c.a.x >> (8 * u));
        prod = (double)((float)q * (float)chunk_x(c, u) * L.tscale[j]);
      } else {
        continue;
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
  if (key >= 0) atomicAdd(&L.acc[key], run);
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
  d.pad0 = d.pad1 = 0;
  return d;
}

template <int MODE, bool COUNT, typename Epi>
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
  const bool all64 = MODE == kAll64 || force64;
  for (int i = lane; i < kWinRows * 16; i += 64) L.acc[i] = 0.0;
  int ta = __builtin_amdgcn_readfirstlane(desc.ta);
  int Za = __builtin_amdgcn_readfirstlane(desc.Z0);
  int ra = tr0;
  while (ra < tr1) {
    const int rlim = ra + kWinRows < tr1 ? ra + kWinRows : tr1;
    int tb = ta + kWinTiles < T1 ? ta + kWinTiles : T1;
    if (rlim < tr1) {
      const int tl = __builtin_amdgcn_readfirstlane(A.tile_ptr[rlim]);
      if (tl < tb) tb = tl;
    }
    const int nt = tb - ta;
    int Zb = Za;
    if (nt > 0) {
      const int Z0 = Za;
      const int Z1 = tb == T1 ? __builtin_amdgcn_readfirstlane(desc.Z1)
                              : __builtin_amdgcn_readfirstlane(A.tile_off[tb]);
      Zb = Z1;
      const int zb = Z0 & ~3;
      const int nsl = (Z1 - zb + 255) >> 8;
      // Window metadata into registers (tiles lane and lane + 64; offsets
      // also of tile 128), then the first slice's loads that need none of
      // it ((row,col) bytes; FP64 payload), then the metadata into LDS.
      const int ja = lane, jb = lane + 64;
      int oa = 0, ob = 0, oc = 0, ca = 0, cb = 0, wa = 0, wb = 0, la = 0, lb = 0;
      float sa = 0.0f, sb = 0.0f;
      if (ja <= nt) oa = A.tile_off[ta + ja];
      if (jb <= nt) ob = A.tile_off[ta + jb];
      if (lane == 0 && nt == kWinTiles) oc = A.tile_off[ta + kWinTiles];
      if (ja < nt) {
        ca = A.tile_col[ta + ja];
        wa = A.tile_row[ta + ja];
        if (MODE != kAll64) {
          la = tile_level<MODE>(A, seglev, ta + ja, force64);
          sa = la == MF_PREC_FP16 ? A.s16[ta + ja] : A.s8[ta + ja];
        }
      }
      if (jb < nt) {
        cb = A.tile_col[ta + jb];
        wb = A.tile_row[ta + jb];
        if (MODE != kAll64) {
          lb = tile_level<MODE>(A, seglev, ta + jb, force64);
          sb = lb == MF_PREC_FP16 ? A.s16[ta + jb] : A.s8[ta + jb];
        }
      }
      Chunk cx, cy;
      if (all64) chunk_load64(A, zb, lane, Z0, Z1, cx);
      else chunk_load_rc(A, zb, lane, Z0, Z1, cx);
      if (ja <= nt) L.toff[ja] = oa - Z0;
      if (jb <= nt) L.toff[jb] = ob - Z0;
      if (lane == 0 && nt == kWinTiles) L.toff[kWinTiles] = oc - Z0;
      if (ja < nt) {
        L.tcol[ja] = ca << 4;
        L.trow[ja] = (uint8_t)(wa - ra);
        if (MODE != kAll64) {
          L.tprec[ja] = (uint8_t)la;
          L.tscale[ja] = sa;
          if (COUNT) cnt->add(la);
        }
      }
      if (jb < nt) {
        L.tcol[jb] = cb << 4;
        L.trow[jb] = (uint8_t)(wb - ra);
        if (MODE != kAll64) {
          L.tprec[jb] = (uint8_t)lb;
          L.tscale[jb] = sb;
          if (COUNT) cnt->add(lb);
        }
      }
      __builtin_amdgcn_wave_barrier();
      // Slices of 256 entries, with the loads of slice m+1 in flight while
      // slice m is gathered and consumed.  FP64 payload: slice m+1's loads
      // need no metadata and go out right after slice m is located.  Per-tile
      // levels: slice m+1 is located and its loads issued before slice m's
      // gathers.  Two chunk buffers in fixed roles, so no register copy
      // waits for loads in flight.
      if (all64) {
        for (int m = 0; m < nsl; m += 2) {
          const int b0 = zb + 256 * m;
          chunk_locate(L, b0, lane, Z0, Z1, nt, cx);
          if (m + 1 < nsl) chunk_load64(A, b0 + 256, lane, Z0, Z1, cy);
          chunk_gather(L, x, x32, true, cx);
          chunk_consume(L, cx, true);
          if (m + 1 >= nsl) break;
          chunk_locate(L, b0 + 256, lane, Z0, Z1, nt, cy);
          if (m + 2 < nsl) chunk_load64(A, b0 + 512, lane, Z0, Z1, cx);
          chunk_gather(L, x, x32, true, cy);
          chunk_consume(L, cy, true);
        }
      } else {
        chunk_locate(L, zb, lane, Z0, Z1, nt, cx);
        chunk_load_levels(A, L, zb, lane, cx);
        for (int m = 0; m < nsl; m += 2) {
          const int b1 = zb + 256 * (m + 1);
          if (m + 1 < nsl) {
            chunk_locate(L, b1, lane, Z0, Z1, nt, cy);
            chunk_load_rc(A, b1, lane, Z0, Z1, cy);
            chunk_load_levels(A, L, b1, lane, cy);
          }
          chunk_gather(L, x, x32, false, cx);
          chunk_consume(L, cx, false);
          if (m + 1 >= nsl) break;
          if (m + 2 < nsl) {
            chunk_locate(L, b1 + 256, lane, Z0, Z1, nt, cx);
            chunk_load_rc(A, b1 + 256, lane, Z0, Z1, cx);
            chunk_load_levels(A, L, b1 + 256, lane, cx);
          }
          chunk_gather(L, x, x32, false, cy);
          chunk_consume(L, cy, false);
        }
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
template <int MODE, bool COUNT, typename Epi>
__device__ __forceinline__ void wave_spmv(const TileView &A, int tr0, int tr1,
                                          const double *__restrict__ x,
                                          const float *__restrict__ x32,
                                          const uint8_t *__restrict__ seglev, int force64,
                                          WaveLds &L, int lane, PrecCount *cnt, Epi epi) {
  if (tr0 >= tr1) return;
  wave_spmv<MODE, COUNT>(A, wave_desc(A, tr0, tr1), x, x32, seglev, force64, L, lane, cnt, epi);
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
        const int nz = A.tile_off[t + 1] - off[u];
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
  return s;
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
  for (int i = 0; i < NV; ++i) v[i] = res[i];
  slot ^= 1;
}

}  // namespace mf
