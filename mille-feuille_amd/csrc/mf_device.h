// mf_device.h -- device-side building blocks shared by the SpMV and solver
// kernels: the wave64 tile-row SpMV body, deterministic block / grid
// reductions and the epoch grid barrier.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mf_internal.h"

#define MF_HIP(call)                                                         \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return mf::fail(MF_ERR_HIP, std::string(#call) + ": " +               \
                                      hipGetErrorString(e_));                \
  } while (0)

namespace mf {

constexpr int kWave = 64;

// Pointers of the device layout passed by value into kernels.
struct TileView {
  const int *tile_ptr;
  const int *tile_col;
  const int *tile_off;
  const uint8_t *rowptr;
  const uint8_t *col;
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
  return TileView{d.tile_ptr, d.tile_col, d.tile_off, d.rowptr, d.col, d.v64,
                  d.v32,      d.v16,      d.v8,       d.s16,    d.s8,  d.prec,
                  t->tilem,   t->tilenum, d.npad};
}

__device__ __forceinline__ float half_bits_to_float(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}

// SpMV modes: every tile FP64; per-tile levels (A.prec); per-segment levels
// (seglev[tile column], the rule policy's flag indexed by the 16-entry segment
// of x that the tile multiplies, src/main-cg-mixed.cu:236, 1134).
enum SpmvMode { kAll64 = 0, kPerTile = 1, kPerSeg = 2 };

// Per-lane counters of tile visits per precision level (lane r == 0 of each
// group counts its tile).
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

// One wave computes the 16 row sums of tile-row `bi`.  Lanes are split into
// four groups of 16 (g = lane>>4); group g walks tiles t0+g, t0+g+4, ...,
// lane r of the group owns row r of the tile.  Each lane keeps one running
// FP64 sum across its tiles, and the four groups are folded with two xor
// shuffles in a fixed order, so the result is deterministic.  On return every
// lane holds the sum of row (lane & 15).
//
// Mixed precision: FP64 tiles multiply x (FP64); FP32 / FP16 / INT8 tiles
// multiply the FP32 copy x32 and accumulate the tile in FP32, then the
// per-tile scale is applied once per tile row; SKIP tiles are not read.
// force64 != 0 treats every tile as FP64 (the solver's fallback).
template <int MODE, bool COUNT = false>
__device__ __forceinline__ double tile_row_sum(const TileView &A, int bi, int lane,
                                               const double *__restrict__ x,
                                               const float *__restrict__ x32,
                                               const uint8_t *__restrict__ seglev = nullptr,
                                               int force64 = 0,
                                               PrecCount *cnt = nullptr) {
  const int g = lane >> 4, r = lane & 15;
  const int t0 = A.tile_ptr[bi], t1 = A.tile_ptr[bi + 1];
  double acc = 0.0;
  for (int t = t0 + g; t < t1; t += 4) {
    const int off = A.tile_off[t];
    const int nzt = A.tile_off[t + 1] - off;
    const int tc = A.tile_col[t];
    const int xo = tc << 4;
    const int s = A.rowptr[(size_t)t * 16 + r];
    int e = __shfl_down(s, 1, 16);
    if (r == 15) e = nzt;
    int p = MF_PREC_FP64;
    if (MODE == kPerTile) p = force64 ? MF_PREC_FP64 : A.prec[t];
    if (MODE == kPerSeg) p = force64 ? MF_PREC_FP64 : seglev[tc];
    if (COUNT && r == 0) cnt->add(p);
    if (p == MF_PREC_FP64) {
      const double *__restrict__ v = A.v64 + off;
      const uint8_t *__restrict__ c = A.col + off;
      double a = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (s + j < e) a += v[s + j] * x[xo + c[s + j]];
      for (int k = s + 4; k < e; ++k) a += v[k] * x[xo + c[k]];
      acc += a;
    } else if (p == MF_PREC_FP32) {
      const float *__restrict__ v = A.v32 + off;
      const uint8_t *__restrict__ c = A.col + off;
      float a = 0.0f;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (s + j < e) a += v[s + j] * x32[xo + c[s + j]];
      for (int k = s + 4; k < e; ++k) a += v[k] * x32[xo + c[k]];
      acc += (double)a;
    } else if (p == MF_PREC_FP16) {
      const uint16_t *__restrict__ v = A.v16 + off;
      const uint8_t *__restrict__ c = A.col + off;
      float a = 0.0f;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (s + j < e) a += half_bits_to_float(v[s + j]) * x32[xo + c[s + j]];
      for (int k = s + 4; k < e; ++k) a += half_bits_to_float(v[k]) * x32[xo + c[k]];
      acc += (double)(a * A.s16[t]);
    } else if (p == MF_PREC_INT8) {
      const int8_t *__restrict__ v = A.v8 + off;
      const uint8_t *__restrict__ c = A.col + off;
      float a = 0.0f;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (s + j < e) a += (float)v[s + j] * x32[xo + c[s + j]];
      for (int k = s + 4; k < e; ++k) a += (float)v[k] * x32[xo + c[k]];
      acc += (double)(a * A.s8[t]);
    }
  }
  acc += __shfl_xor(acc, 16);
  acc += __shfl_xor(acc, 32);
  return acc;
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
