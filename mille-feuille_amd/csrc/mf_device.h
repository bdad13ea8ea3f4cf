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
constexpr int kWinTiles = 128;
constexpr int kWinRows = 16;

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

// One lane's share of a slice: four consecutive entries k0..k0+3, their
// (row,col) bytes and the raw payload words at the level of the tile that
// holds the first valid entry (FP64: 8 words, FP32: 4, FP16: 2, INT8: 1).
struct SliceBuf {
  int k0;
  int j;    // window-relative tile of the first valid entry, -1: none valid
  int lvl;  // level of tile j (the level the payload words were loaded at)
  uint32_t rc4;
  uint4 a, b;  // raw payload words 0-3 and 4-7
};

// Word i of the payload (i is a compile-time constant after unrolling).
__device__ __forceinline__ uint32_t slice_word(const SliceBuf &s, int i) {
  switch (i) {
    case 0: return s.a.x;
    case 1: return s.a.y;
    case 2: return s.a.z;
    case 3: return s.a.w;
    case 4: return s.b.x;
    case 5: return s.b.y;
    case 6: return s.b.z;
    default: return s.b.w;
  }
}

template <int MODE>
__device__ __forceinline__ void slice_load(const TileView &A, const WaveLds &L, int base,
                                           int lane, int Z0, int Z1, int nt, bool all64,
                                           SliceBuf &s) {
  s.k0 = base + 4 * lane;
  s.j = -1;
  s.lvl = MF_PREC_FP64;
  if (s.k0 + 3 < Z0 || s.k0 >= Z1) return;
  s.rc4 = *(const uint32_t *)(A.rc + s.k0);
  if (all64) {
    const uint4 *q = (const uint4 *)(A.v64 + s.k0);
    s.a = q[0];
    s.b = q[1];
  }
  // tile of the first valid entry: binary search over the staged offsets
  const int kf = (s.k0 > Z0 ? s.k0 : Z0) - Z0;
  int lo = 0, hi = nt - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.toff[mid] <= kf) lo = mid;
    else hi = mid - 1;
  }
  s.j = lo;
  if (!all64) {
    s.lvl = L.tprec[lo];
    if (s.lvl == MF_PREC_FP64) {
      const uint4 *q = (const uint4 *)(A.v64 + s.k0);
      s.a = q[0];
      s.b = q[1];
    } else if (s.lvl == MF_PREC_FP32) {
      s.a = *(const uint4 *)(A.v32 + s.k0);
    } else if (s.lvl == MF_PREC_FP16) {
      const uint2 h = *(const uint2 *)(A.v16 + s.k0);
      s.a.x = h.x;
      s.a.y = h.y;
    } else if (s.lvl == MF_PREC_INT8) {
      s.a.x = *(const uint32_t *)(A.v8 + s.k0);
    }
  }
}

template <int MODE>
__device__ __forceinline__ void slice_consume(const TileView &A, WaveLds &L, const SliceBuf &s,
                                              int Z0, int Z1, bool all64,
                                              const double *__restrict__ x,
                                              const float *__restrict__ x32) {
  if (s.j < 0) return;
  int j = s.j;
  int key = -1;
  double run = 0.0;
  // Fast path: all four entries valid and inside tile j (the common case:
  // tiles hold 12-18 entries on the stencil workloads).
  if (s.k0 >= Z0 && s.k0 + 3 < Z1 && L.toff[j + 1] > s.k0 + 3 - Z0) {
    const int xo = L.tcol[j];
    const int ro = L.trow[j] * 16;
    double prod[4];
    const int lvl = all64 ? MF_PREC_FP64 : s.lvl;
    if (lvl == MF_PREC_FP64) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        prod[u] = __builtin_bit_cast(double, ((uint64_t)slice_word(s, 2 * u + 1) << 32) |
                                                 slice_word(s, 2 * u)) *
                  x[xo + ((s.rc4 >> (8 * u)) & 15)];
    } else if (lvl == MF_PREC_FP32) {
#pragma unroll
      for (int u = 0; u < 4; ++u)
        prod[u] = (double)(__builtin_bit_cast(float, slice_word(s, u)) *
                           x32[xo + ((s.rc4 >> (8 * u)) & 15)]);
    } else if (lvl == MF_PREC_FP16) {
      const float sc = L.tscale[j];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint16_t h = (uint16_t)(slice_word(s, u >> 1) >> (16 * (u & 1)));
        prod[u] = (double)(half_bits_to_float(h) * x32[xo + ((s.rc4 >> (8 * u)) & 15)] * sc);
      }
    } else if (lvl == MF_PREC_INT8) {
      const float sc = L.tscale[j];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int8_t q = (int8_t)(s.a.x >> (8 * u));
        prod[u] = (double)((float)q * x32[xo + ((s.rc4 >> (8 * u)) & 15)] * sc);
      }
    } else {
      return;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kk = ro + ((s.rc4 >> (8 * u + 4)) & 15);
      if (kk != key) {
        if (key >= 0) atomicAdd(&L.acc[key], run);
        key = kk;
        run = prod[u];
      } else {
        run += prod[u];
      }
    }
    atomicAdd(&L.acc[key], run);
    return;
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int k = s.k0 + u;
    if (k < Z0 || k >= Z1) continue;
    while (L.toff[j + 1] <= k - Z0) ++j;
    const int rc = (s.rc4 >> (8 * u)) & 0xff;
    const int xi = L.tcol[j] + (rc & 15);
    const int kk = L.trow[j] * 16 + (rc >> 4);
    double prod;
    if (all64) {
      prod = __builtin_bit_cast(double, ((uint64_t)slice_word(s, 2 * u + 1) << 32) | slice_word(s, 2 * u)) * x[xi];
    } else {
      const int lvl = L.tprec[j];
      if (lvl != s.lvl) {
        prod = entry_product<MODE>(A, lvl, k, xi, x, x32, L.tscale[j]);
      } else if (lvl == MF_PREC_FP64) {
        prod = __builtin_bit_cast(double, ((uint64_t)slice_word(s, 2 * u + 1) << 32) | slice_word(s, 2 * u)) * x[xi];
      } else if (lvl == MF_PREC_FP32) {
        prod = (double)(__builtin_bit_cast(float, slice_word(s, u)) * x32[xi]);
      } else if (lvl == MF_PREC_FP16) {
        const uint16_t h = (uint16_t)(slice_word(s, u >> 1) >> (16 * (u & 1)));
        prod = (double)(half_bits_to_float(h) * x32[xi] * L.tscale[j]);
      } else if (lvl == MF_PREC_INT8) {
        const int8_t q = (int8_t)(s.a.x >> (8 * u));
        prod = (double)((float)q * x32[xi] * L.tscale[j]);
      } else {
        prod = 0.0;
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

template <int MODE, bool COUNT, typename Epi>
__device__ __forceinline__ void wave_spmv(const TileView &A, int tr0, int tr1,
                                          const double *__restrict__ x,
                                          const float *__restrict__ x32,
                                          const uint8_t *__restrict__ seglev, int force64,
                                          WaveLds &L, int lane, PrecCount *cnt, Epi epi) {
  if (tr0 >= tr1) return;
  const bool all64 = MODE == kAll64 || force64;
  for (int i = lane; i < kWinRows * 16; i += 64) L.acc[i] = 0.0;
  const int T1 = A.tile_ptr[tr1];
  int ta = A.tile_ptr[tr0];
  int ra = tr0;
  while (ra < tr1) {
    const int rlim = ra + kWinRows < tr1 ? ra + kWinRows : tr1;
    int tb = ta + kWinTiles < T1 ? ta + kWinTiles : T1;
    if (rlim < tr1) {
      const int tl = A.tile_ptr[rlim];
      if (tl < tb) tb = tl;
    }
    const int nt = tb - ta;
    if (nt > 0) {
      const int Z0 = A.tile_off[ta], Z1 = A.tile_off[tb];
      for (int j = lane; j <= nt; j += 64) L.toff[j] = A.tile_off[ta + j] - Z0;
      for (int j = lane; j < nt; j += 64) {
        const int t = ta + j;
        L.tcol[j] = A.tile_col[t] << 4;
        L.trow[j] = (uint8_t)(A.tile_row[t] - ra);
        if (MODE != kAll64) {
          const int lv = tile_level<MODE>(A, seglev, t, force64);
          L.tprec[j] = (uint8_t)lv;
          L.tscale[j] = lv == MF_PREC_FP16 ? A.s16[t] : A.s8[t];
          if (COUNT) cnt->add(lv);
        }
      }
      __builtin_amdgcn_wave_barrier();
      // Slices of 256 entries; the loads of the next slice are issued before
      // the current one is consumed.
      const int zb = Z0 & ~3;
      const int nsl = (Z1 - zb + 255) >> 8;
      SliceBuf cur, nxt;
      slice_load<MODE>(A, L, zb, lane, Z0, Z1, nt, all64, cur);
      for (int m = 0; m < nsl; ++m) {
        if (m + 1 < nsl) slice_load<MODE>(A, L, zb + 256 * (m + 1), lane, Z0, Z1, nt, all64, nxt);
        slice_consume<MODE>(A, L, cur, Z0, Z1, all64, x, x32);
        cur = nxt;
      }
      __builtin_amdgcn_wave_barrier();
    }
    // Tile-rows [ra, R) are complete; tile-row R has started iff tile_ptr[R] < tb.
    int R;
    if (tb < T1) {
      R = A.tile_row[tb];
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
  }
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
