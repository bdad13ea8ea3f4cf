// mf_solve.hip -- single-kernel (persistent, cooperative) CG and BiCGSTAB.
//
// One launch runs the whole solve.  The grid is sized to the occupancy of
// the kernel on every CU and launched cooperatively, so all blocks are
// co-resident and the epoch grid barrier cannot deadlock (the reference
// assumes co-residency and hardcodes the A100 warp count,
// src/main-cg.cu:675-696).  Every wave owns a contiguous, nnz-balanced range
// of tile-rows: it computes the SpMV rows of those tile-rows (lanes 0..15 hold
// the 16 row sums) and does the vector updates of the same rows, so the
// vector phases never read data written by another wave in the same phase.
// Dot products go through deterministic grid sums: the result is bitwise
// identical run to run and in every block (the reference's double atomics
// are order dependent, src/Mille-feuille-cg.h:489-509).
//
// FP64 is kept for the scalars (alpha, beta, omega), the solution update and
// the true residual r = b - A x, which is recomputed in FP64 whenever the
// recursive residual meets the tolerance (so the exit test is always on the
// FP64 residual) and, for the mixed policies, whenever the recursive residual
// fell by `refresh_ratio` since the last recomputation (reliable updates,
// i.e. iterative refinement inside one Krylov run).
#include <algorithm>
#include <queue>
#include <mutex>
#include <string>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <vector>
#include <type_traits>

#include "mf_device.h"

namespace mf {

// Packed MLP of a handle (mf_learned.hip), or null when none is installed.
const MlpDevice *policy_mlp(const mf_tiled *t);

constexpr int kBlock = 1024;  // 16 waves, one block per CU
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int kUpd = 4;  // 64-row groups per step of the vector loops (loads first)

enum StatSlot {
  kStIters = 0,
  kStRefresh,
  kStFallback,
  kStRecursive,
  kStTrue,
  kStBB,
  kStStatus,
  kStCount = 16
};

// Wave index of the calling wave in the solver's tile-row order.  Blocks are
// dispatched round-robin over the 8 XCDs (block b on XCD b % 8); with a grid
// that is a multiple of 8, XCD x takes the x-th eighth of the waves, so the
// tile-rows a wave's neighbours own (and whose vector entries its SpMV
// gathers) sit in the same XCD's L2.
__device__ __forceinline__ int solver_wave_id() {
  const int G = gridDim.x;
  const int lb = (G & 7) ? (int)blockIdx.x : (int)(blockIdx.x & 7) * (G >> 3) + (int)(blockIdx.x >> 3);
  return lb * kWavesPerBlock + (int)(threadIdx.x >> 6);
}

struct SolveArgs {
  TileView A;
  const double *b;
  double *x;
  double *r, *p, *q;          // CG (q doubles as v in BiCGSTAB)
  double *rh, *s, *tv;        // BiCGSTAB
  float *p32, *s32;
  float *r32, *q32;           // mixed CG: r and q in FP32 (p in p32)
  float *rh32, *t32;          // mixed BiCGSTAB: r^ and t in FP32 (v in q32)
  uint8_t *seg_p, *seg_s;     // per-segment levels (rule policy)
  const int *wave_tr;
  const int *wave_win;       // per wave, bounds of its SpMV descriptors
  const WaveDesc *win_desc;  // window descriptors and heavy tile-row pieces
  double *heavy_part;        // 16 partial row sums per heavy tile-row piece
  unsigned *heavy_cnt;       // pieces of each heavy tile-row done
  const int *heavy_n;        // pieces of each heavy tile-row
  const int *heavy_base;     // first piece slot of each heavy tile-row
  const int *heavy_tr;       // tile-row of each heavy tile-row
  const int *wave_heavy;     // per wave, bounds of the heavy tile-rows it owns
  double *partials;
  unsigned long long *mbox;   // value-only grid sums: (blocks + 1) mailboxes
  GridSync *gs;
  double *stats;
  unsigned long long *hist;   // 5 counters
  double tol2;                // tol^2
  int maxiter;
  int fixed_iters;
  int policy;
  int check_interval;
  double ratio2;              // refresh_ratio^2
  int max_refresh;
  int refresh_every;
  // learned policy (MODE kPerTile; mlp == nullptr otherwise)
  const MlpDevice *mlp;
  const uint16_t *feat16;
  float *seg2;                // per-segment ||v_seg||^2 of the multiplied vector
  uint8_t *prec_rw;           // tile levels written by the inference
  int nseg;
#ifdef MF_PHASE_TRACE
  unsigned long long *trace;  // per (iteration, phase, block) wall clock
#endif
};

#ifdef MF_PHASE_TRACE
// Profiling builds: thread 0 of every block records the wall clock (100 MHz)
// at the phase boundaries of the first kTraceIters CG iterations.
constexpr int kTraceIters = 64, kTracePhases = 5;
#define MF_MARK(ph)                                                                         \
  do {                                                                                     \
    if (a.trace && trace_it < kTraceIters) {                                               \
      __syncthreads();                                                                     \
      if (threadIdx.x == 0)                                                                \
        a.trace[((size_t)trace_it * kTracePhases + (ph)) * gridDim.x + blockIdx.x] =       \
            wall_clock64();                                                                \
    }                                                                                      \
  } while (0)
#else
#define MF_MARK(ph) \
  do {              \
  } while (0)
#endif

// Reference rule (src/main-cg-mixed.cu:1326-1363): level from the absolute
// change of a 16-entry segment of the search direction.
__device__ __forceinline__ int rule_level(double d) {
  if (d > 1e-5) return MF_PREC_FP64;
  if (d > 1e-6) return MF_PREC_FP32;
  if (d > 1e-7) return MF_PREC_FP16;
  if (d > 1e-8) return MF_PREC_INT8;
  return MF_PREC_SKIP;
}

// Max over the 16 lanes of each group.
__device__ __forceinline__ double group16_max(double v) {
#pragma unroll
  for (int m = 8; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m));
  return v;
}

// Sum over the 16 lanes of each group.
__device__ __forceinline__ double group16_sum(double v) {
#pragma unroll
  for (int m = 8; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}

// Inference context of the learned policy for a vector of squared norm v2.
__device__ __forceinline__ InferCtx infer_ctx(const SolveArgs &a, double v2, double rr, double bb) {
  InferCtx c;
  c.w = a.mlp;
  c.feat16 = a.feat16;
  c.seg2 = a.seg2;
  c.prec_out = a.prec_rw;
  c.seg_scale = v2 > 0.0 ? (float)(sqrt((double)a.nseg) / sqrt(v2)) : 0.0f;
  const double ratio = bb > 0.0 ? sqrt(rr / bb) : 1.0;
  c.progress = (float)(log10(fmax(ratio, 1e-16)) / 16.0);
  return c;
}

// uniform: the counters are per wave (alike in every lane; lane 0's count).
__device__ __forceinline__ void flush_hist(const PrecCount &c, unsigned long long *hist,
                                           double *lds, bool uniform = false) {
  const bool take = !uniform || (threadIdx.x & 63) == 0;
  double v0 = take ? c.c0 : 0, v1 = take ? c.c1 : 0, v2 = take ? c.c2 : 0, v3 = take ? c.c3 : 0,
         v4 = take ? c.c4 : 0;
  v0 = block_sum(v0, lds);
  v1 = block_sum(v1, lds);
  v2 = block_sum(v2, lds);
  v3 = block_sum(v3, lds);
  v4 = block_sum(v4, lds);
  if (threadIdx.x == 0) {
    atomicAdd(&hist[0], (unsigned long long)v0);
    atomicAdd(&hist[1], (unsigned long long)v1);
    atomicAdd(&hist[2], (unsigned long long)v2);
    atomicAdd(&hist[3], (unsigned long long)v3);
    atomicAdd(&hist[4], (unsigned long long)v4);
  }
}

// A solver wave's SpMV state.  RES: every wave's share is a resident window
// (ResLds, staged once); otherwise every wave streams its descriptors through
// a WaveLds.  The host picks the kernel: the resident one when every wave's
// share fits a window.
template <bool RES>
struct WaveCtx {
  static constexpr bool res = RES;
  typename std::conditional<RES, ResLds, WaveLds>::type *L;
  ResWin win;
};

// SpMV of the calling wave's share: its descriptors (whole tile-rows) with
// epi(row, value), then its pieces of heavy tile-rows (wave_spmv_piece).
template <int MODE, bool COUNT, bool RES, typename Epi>
__device__ __forceinline__ void solver_spmv(const SolveArgs &a, const double *__restrict__ x,
                                            const float *__restrict__ x32,
                                            const uint8_t *__restrict__ seglev, int force64,
                                            WaveCtx<RES> &ctx, int lane, PrecCount *cnt, Epi epi) {
  if constexpr (RES) {
    const bool all64 = MODE == kAll64 || (force64 & 1);
    res_spmv<COUNT>(a.A, *ctx.L, ctx.win, x, x32, all64, !all64 && (force64 & 2), lane, cnt, epi);
    return;
  }
  WaveLds &L = *reinterpret_cast<WaveLds *>(ctx.L);
  const int wid = solver_wave_id();
  const int d0 = __builtin_amdgcn_readfirstlane(a.wave_win[wid]);
  const int d1 = __builtin_amdgcn_readfirstlane(a.wave_win[wid + 1]);
  for (int d = d0; d < d1; ++d) {
    const WaveDesc desc = a.win_desc[d];
    const int flags = __builtin_amdgcn_readfirstlane(desc.flags);
    if (!(flags & kDescPiece)) {
      wave_spmv<MODE, COUNT, 3, false, MODE != kAll64>(a.A, desc, x, x32, seglev, force64, L, lane,
                                                       cnt, epi);
      continue;
    }
    // A piece of a heavy tile-row: its 16 partial row sums go to its slot
    // (agent-scope stores, visible across XCDs), then it counts itself done.
    const int h = __builtin_amdgcn_readfirstlane(desc.aux);
    double *slot = a.heavy_part + 16 * (size_t)(__builtin_amdgcn_readfirstlane(a.heavy_base[h]) + (flags >> 8));
    wave_spmv<MODE, COUNT, 3, false, MODE != kAll64>(a.A, desc, x, x32, seglev, force64, L, lane, cnt,
      [&](const int i, const double v) {
        __hip_atomic_store(slot + (i & 15), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      });
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0) __hip_atomic_fetch_add(a.heavy_cnt + h, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // The heavy tile-rows the wave owns: once every piece has counted itself
  // done, the owner adds the pieces' partial sums in piece order and runs the
  // epilogue, so the rows' values and their terms in the wave's dot-product
  // partials are the same on every run.  Pieces never wait, and a wave waits
  // only after its own descriptors (its pieces included), so the wait ends.
  const int h0 = __builtin_amdgcn_readfirstlane(a.wave_heavy[wid]);
  const int h1 = __builtin_amdgcn_readfirstlane(a.wave_heavy[wid + 1]);
  for (int h = h0; h < h1; ++h) {
    const unsigned np = (unsigned)__builtin_amdgcn_readfirstlane(a.heavy_n[h]);
    unsigned spins = 0;
    while (__hip_atomic_load(a.heavy_cnt + h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < np) {
      if (++spins > kSpinLimit) {
        __hip_atomic_store(&a.gs->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      if ((spins & 1023u) == 0 &&
          __hip_atomic_load(&a.gs->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        break;
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int base = __builtin_amdgcn_readfirstlane(a.heavy_base[h]);
    // lane (q, row): pieces q, q+4, ... of the row, then the four partial
    // sums in q order
    const int row = lane & 15, q = lane >> 4;
    double sum = 0.0;
    for (int k = q; k < (int)np; k += 4)
      sum += __hip_atomic_load(a.heavy_part + 16 * (size_t)(base + k) + row, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    const double s1 = __shfl(sum, row + 16), s2 = __shfl(sum, row + 32), s3 = __shfl(sum, row + 48);
    const int R = __builtin_amdgcn_readfirstlane(a.heavy_tr[h]);
    if (lane < 16) epi(16 * R + lane, ((sum + s1) + s2) + s3);
    // reset for the next SpMV (a grid barrier separates it from this one)
    if (lane == 0) __hip_atomic_store(a.heavy_cnt + h, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Learned levels for exactly the tiles the calling wave's SpMV reads: its
// resident window, or its descriptors (whole tile-rows, and the pieces of
// heavy tile-rows it holds; a heavy tile-row's tiles are inferred by the
// waves that multiply them, never by its owner).  Returns whether any of
// them changed level (a resident window is restaged only then).
template <bool RES>
__device__ __forceinline__ bool solver_infer(const SolveArgs &a, const InferCtx &ic, int tr0, int tr1,
                                             int lane) {
  if constexpr (RES) {
    return wave_infer(a.A, ic, tr0, tr1, lane);
  } else {
    const int wid = solver_wave_id();
    const int d0 = __builtin_amdgcn_readfirstlane(a.wave_win[wid]);
    const int d1 = __builtin_amdgcn_readfirstlane(a.wave_win[wid + 1]);
    bool changed = false;
    for (int d = d0; d < d1; ++d)
      changed = wave_infer_tiles(a.A, ic, __builtin_amdgcn_readfirstlane(a.win_desc[d].ta),
                                 __builtin_amdgcn_readfirstlane(a.win_desc[d].T1), lane) ||
                changed;
    return changed;
  }
}

// Dot products whose next phase reads only the wave's own rows.  A wave's
// SpMV epilogue writes exactly its own rows (the owner finishes its heavy
// tile-rows), so the values-only grid sum suffices: it orders no memory, and
// a phase that reads other waves' rows takes a grid_sync first.
template <bool RES, int NV>
__device__ __forceinline__ void grid_sum_own(const SolveArgs &a, double (&v)[NV], unsigned &epoch,
                                             int &slot, double *lds, int *lds_i) {
  (void)slot;
  (void)lds_i;
  grid_sum_values<NV>(v, a.mbox, a.gs, epoch, lds);
}

// (Re-)stages a resident wave's window with the current tile levels.
template <int MODE, bool RES>
__device__ __forceinline__ void solver_stage(const SolveArgs &a, WaveCtx<RES> &ctx, int tr0, int tr1,
                                             const uint8_t *__restrict__ seglev, int lane) {
  if constexpr (RES) res_stage<MODE>(a.A, *ctx.L, ctx.win, tr0, tr1, seglev, false, lane);
}

// True residual of the wave's SpMV share (its tile-rows, and the heavy
// tile-rows whose last piece it completes): r_new = b - A x (FP64), optionally
// stored in FP64 (store) or FP32 (store32).  Returns the wave's share of
// ||r_new||^2 (of the stored FP32 values when store32 is given).
template <bool RES>
__device__ __forceinline__ double true_residual_rows(const SolveArgs &a, int lane, double *store,
                                                     WaveCtx<RES> &ctx,
                                                     float *store32 = nullptr) {
  double acc = 0.0;
  solver_spmv<kAll64, false>(a, a.x, nullptr, nullptr, 0, ctx, lane, nullptr,
    [&](const int i, const double ax) {
      double rv = a.b[i] - ax;
      if (store) store[i] = rv;
      if (store32) {
        const float rf = (float)rv;
        store32[i] = rf;
        rv = rf;
      }
      acc += rv * rv;
    });
  return acc;
}

// Restart helper: r = b - A x in FP64 for the wave's SpMV share, p = r (and
// r^ = r for BiCGSTAB).  Returns the wave's share of ||r||^2.
// lowv: the mixed solvers' FP32 vectors r32 = p32 (= rh32) = (float) r
// instead.
template <bool RES>
__device__ __forceinline__ double restart_rows(const SolveArgs &a, int lane, bool mixed, bool bicg,
                                               WaveCtx<RES> &ctx, bool lowv = false) {
  double acc = 0.0;
  solver_spmv<kAll64, false>(a, a.x, nullptr, nullptr, 0, ctx, lane, nullptr,
    [&](const int i, const double ax) {
      double rv = a.b[i] - ax;
      if (lowv) {
        const float rf = (float)rv;
        a.r32[i] = rf;
        a.p32[i] = rf;
        if (bicg) a.rh32[i] = rf;
        rv = rf;
      } else {
        a.r[i] = rv;
        a.p[i] = rv;
        if (mixed) a.p32[i] = (float)rv;
        if (bicg) a.rh[i] = rv;
      }
      acc += rv * rv;
    });
  return acc;
}

// Memory-form CG in the Chronopoulos-Gear arrangement: one SpMV w = A r whose
// epilogue accumulates gamma = r.r and delta = w.r, one fused grid sum, then
// p = r + beta p, s = w + beta s, x += alpha p, r -= alpha s and a grid sync:
// two grid-wide synchronisations per iteration.  With a mixed policy r, w,
// p and s are FP32 (x, the scalars and every residual recomputation stay
// FP64; FP64 tiles multiply the FP32 r in FP64 arithmetic); after a
// fallback to FP64 tiles the recurrence continues in FP64 vectors.
// Vectors: FP64 r, q (w), p, s / FP32 r32, q32 (w), p32, s32.  With resident
// windows w never goes to memory: the update reads it from the wave's row
// accumulators, where the SpMV left it.
// Store of a vector pass's result that writes through the XCD's L2 (an
// agent-scope store): the grid barrier's release then has few dirty L2 lines
// to write back (at 128^3 the FP64 CG iteration went from 59.4 to 54.7 us).
template <typename T>
__device__ __forceinline__ void wt_store(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int MODE, bool RES>
__global__ __launch_bounds__(kBlock) void k_cg(SolveArgs a) {
  __shared__ double lds[kMaxGridSum * kWavesPerBlock];
  __shared__ int lds_i;
  __shared__ typename std::conditional<RES, ResLds, WaveLds>::type rl[kWavesPerBlock];
  const int lane = threadIdx.x & 63;
  const int wid = solver_wave_id();
  const int tr0 = __builtin_amdgcn_readfirstlane(a.wave_tr[wid]);
  const int tr1 = __builtin_amdgcn_readfirstlane(a.wave_tr[wid + 1]);
  const int row0 = tr0 * 16, row1 = tr1 * 16;
  constexpr bool mixed = MODE != kAll64;
  const bool learned = MODE == kPerTile && a.mlp != nullptr;
  unsigned epoch = 0;
  int slot = 0;
  PrecCount cnt;
  WaveCtx<RES> ctx;
  ctx.L = &rl[threadIdx.x >> 6];
  solver_stage<MODE>(a, ctx, tr0, tr1, a.seg_p, lane);

  // r = b - A x0 (FP64, and the FP32 copy the mixed recurrence starts from)
  double red[2] = {0.0, 0.0};
  solver_spmv<kAll64, false>(a, a.x, nullptr, nullptr, 0, ctx, lane, nullptr,
    [&](const int i, const double ax) {
      const double bv = a.b[i];
      double rv = bv - ax;
      a.r[i] = rv;
      if (mixed) {
        const float rf = (float)rv;
        a.r32[i] = rf;
        rv = rf;
      }
      red[0] += rv * rv;
      red[1] += bv * bv;
    });
  if (learned) {
    // segment norms of r0 for the first inference
    for (int base = row0; base < row1; base += 64) {
      const int i = base + lane;
      const double rv = i < row1 ? (double)a.r32[i] : 0.0;
      const double d2 = group16_sum(rv * rv);
      if ((lane & 15) == 0 && i < row1) a.seg2[i >> 4] = (float)d2;
    }
  }
  grid_sum<2>(red, a.partials, a.gs, epoch, slot, lds, &lds_i);
  double rr = red[0];
  const double bb = red[1];
  const double stop2 = uniform_double(a.tol2 * bb);
  double rr_ref = rr, gamma_old = 1.0, alpha_old = 1.0;
  bool first = true;
  int it = 0, it_ref = 0, refresh = 0, force64 = 0, status = 0;
  bool done = bb == 0.0 || (a.fixed_iters <= 0 && rr <= stop2);
  bool pending = false;  // recompute the FP64 residual before the next step
  bool infer = learned;  // learned policy: new levels in the next SpMV
  bool restage = false;  // rule policy: new segment levels in the next SpMV
  bool lowv = mixed;     // FP32 vectors (until a fallback to FP64 tiles)

  int trace_it = -1;
  (void)trace_it;
  while (!done) {
    double gamma = 0.0, delta = 0.0;
    if (!pending) {
      ++trace_it;
      MF_MARK(0);
      // w = A r with gamma = r.r and delta = w.r of the wave's rows; new
      // levels (inference, or a rule update before the last sync) are staged
      // into a resident window first
      bool relevel = false;
      if (infer && !force64) relevel = solver_infer<RES>(a, infer_ctx(a, rr, rr, bb), tr0, tr1, lane);
      if (relevel || restage) solver_stage<MODE>(a, ctx, tr0, tr1, a.seg_p, lane);
      infer = false;
      restage = false;
      double gd[2] = {0.0, 0.0};
      solver_spmv<MODE, mixed>(a, a.r, a.r32, a.seg_p, lowv ? 2 : force64, ctx,
                             lane, &cnt, [&](const int i, const double wv) {
        // (a resident window keeps w in its row accumulators until the
        // update; a streamed one stores it)
        double rv;
        if (lowv) {
          rv = a.r32[i];
          if constexpr (!RES) a.q32[i] = (float)wv;
        } else {
          rv = a.r[i];
          if constexpr (!RES) a.q[i] = wv;
        }
        gd[0] += rv * rv;
        gd[1] += wv * rv;
      });
      MF_MARK(1);
      grid_sum_own<RES>(a, gd, epoch, slot, lds, &lds_i);
      MF_MARK(2);
      gamma = gd[0];
      delta = gd[1];
      rr = gamma;
      if (a.fixed_iters <= 0) {
        const bool converged = gamma <= stop2;
        const bool periodic = mixed && !force64 &&
                              (gamma <= a.ratio2 * rr_ref || it - it_ref >= a.refresh_every);
        pending = converged || periodic;
        if (!pending && it >= a.maxiter) { status = MF_NOT_CONVERGED; break; }
      }
    }
    if (pending) {
      pending = false;
      // FP64 residual replacement r = b - A x (x is complete: the last
      // update ended with a grid sync)
      const bool lv = lowv && !force64;
      double tr[1] = {lv ? true_residual_rows(a, lane, nullptr, ctx, a.r32)
                         : true_residual_rows(a, lane, a.r, ctx)};
      grid_sum<1>(tr, a.partials, a.gs, epoch, slot, lds, &lds_i);
      ++refresh;
      it_ref = it;
      // A recursive residual far below the true one leaves beta meaningless,
      // and so does a recursive residual that met the tolerance while the
      // true one did not (the s = A p recurrence drifted): restart the
      // search direction from the recomputed residual.
      if (tr[0] > 4.0 * rr || (rr <= stop2 && tr[0] > stop2)) first = true;
      rr = tr[0];
      if (tr[0] <= stop2) break;
      if (it >= a.maxiter) { status = MF_NOT_CONVERGED; break; }
      if (mixed && !force64 && (tr[0] > 0.25 * rr_ref || refresh > a.max_refresh)) {
        // stagnation: FP64 tiles (and FP64 vectors) from here, restart
        force64 = 1;
        first = true;
      }
      if (lowv && force64) {
        // leaving the FP32 vectors (force64 implies a restart, first):
        // the FP64 recurrence starts from an FP64 residual
        lowv = false;
        if (lv) {
          double t2[1] = {true_residual_rows(a, lane, a.r, ctx)};
          grid_sum<1>(t2, a.partials, a.gs, epoch, slot, lds, &lds_i);
          rr = t2[0];
        }
      }
      rr_ref = rr;
      if (!first) {
        // s = A p in FP64 so that s and the new r stay consistent
        if (lowv) {
          for (int i = row0 + lane; i < row1; i += 64) a.p[i] = (double)a.p32[i];
          grid_sync(a.gs, epoch, &lds_i);
          solver_spmv<kAll64, false>(a, a.p, nullptr, nullptr, 0, ctx, lane,
                                   nullptr, [&](const int i, const double sv) { a.s32[i] = (float)sv; });
        } else {
          solver_spmv<kAll64, false>(a, a.p, nullptr, nullptr, 0, ctx, lane,
                                   nullptr, [&](const int i, const double sv) { a.s[i] = sv; });
        }
      }
      grid_sync(a.gs, epoch, &lds_i);
      continue;
    }
    double beta, alpha;
    if (first) {
      beta = 0.0;
      alpha = gamma / delta;
    } else {
      beta = gamma / gamma_old;
      alpha = gamma / (delta - beta * gamma / alpha_old);
    }
    if (!isfinite(alpha) || alpha == 0.0 || (mixed && !force64 && !(alpha > 0.0))) {
      if (mixed && !force64) {
        // The low-precision operator lost definiteness: FP64 tiles,
        // recompute r = b - A x, restart.
        force64 = 1;
        first = true;
        pending = true;
        continue;
      }
      status = MF_ERR_BREAKDOWN;
      break;
    }
    first = false;
    alpha = uniform_double(alpha);
    beta = uniform_double(beta);
    gamma_old = gamma;
    alpha_old = alpha;
    ++it;
    // p = r + beta p, s = w + beta s, x += alpha p, r -= alpha s; the rule
    // policy looks at |r_new - r_old| per segment, the learned one records
    // the segment norms of r_new for its next inference.
    const bool upd = MODE == kPerSeg && (it % a.check_interval) == 0;
    const bool lupd = learned && !force64 && (it % a.check_interval) == 0;
    // Four 64-row groups per step, every load of a step issued before its
    // arithmetic, so a wave keeps 20 loads in flight instead of 5.
    for (int base = row0; base < row1; base += 64 * kUpd) {
      double rv[kUpd], rn[kUpd];
      if (lowv) {
        float r_[kUpd], p_[kUpd], q_[kUpd], s_[kUpd];
        double x_[kUpd];
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base + 64 * u + lane;
          const bool ok = i < row1;
          r_[u] = ok ? a.r32[i] : 0.0f;
          p_[u] = ok ? a.p32[i] : 0.0f;
          if constexpr (RES)
            q_[u] = ok ? (float)ctx.L->acc[res_acc_index(i - row0)] : 0.0f;
          else
            q_[u] = ok ? a.q32[i] : 0.0f;
          s_[u] = ok ? a.s32[i] : 0.0f;
          x_[u] = ok ? a.x[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base + 64 * u + lane;
          rv[u] = r_[u];
          const double pv = rv[u] + beta * (double)p_[u];
          const double sv = (double)q_[u] + beta * (double)s_[u];
          const float pf = (float)pv, sf = (float)sv;
          const float rf = (float)(rv[u] - alpha * (double)sf);
          rn[u] = rf;
          if (i < row1) {
            wt_store(a.p32 + i, pf);
            wt_store(a.s32 + i, sf);
            wt_store(a.x + i, x_[u] + alpha * (double)pf);
            wt_store(a.r32 + i, rf);
          }
        }
      } else {
        double r_[kUpd], p_[kUpd], q_[kUpd], s_[kUpd], x_[kUpd];
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base + 64 * u + lane;
          const bool ok = i < row1;
          r_[u] = ok ? a.r[i] : 0.0;
          p_[u] = ok ? a.p[i] : 0.0;
          if constexpr (RES)
            q_[u] = ok ? ctx.L->acc[res_acc_index(i - row0)] : 0.0;
          else
            q_[u] = ok ? a.q[i] : 0.0;
          s_[u] = ok ? a.s[i] : 0.0;
          x_[u] = ok ? a.x[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base + 64 * u + lane;
          rv[u] = r_[u];
          const double pv = rv[u] + beta * p_[u];
          const double sv = q_[u] + beta * s_[u];
          rn[u] = rv[u] - alpha * sv;
          if (i < row1) {
            wt_store(a.p + i, pv);
            wt_store(a.s + i, sv);
            wt_store(a.x + i, x_[u] + alpha * pv);
            wt_store(a.r + i, rn[u]);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kUpd; ++u) {
        const int i = base + 64 * u + lane;
        double d = i < row1 ? (lupd ? rn[u] * rn[u] : fabs(rn[u] - rv[u])) : 0.0;
        if (upd) {
          d = group16_max(d);
          if ((lane & 15) == 0 && i < row1) a.seg_p[i >> 4] = (uint8_t)rule_level(d);
        }
        if (lupd) {
          d = group16_sum(d);
          if ((lane & 15) == 0 && i < row1) a.seg2[i >> 4] = (float)d;
        }
      }
    }
    infer = lupd;
    restage = upd;
    MF_MARK(3);
    grid_sync(a.gs, epoch, &lds_i);
    MF_MARK(4);
    if (a.fixed_iters > 0 && it >= a.fixed_iters) break;
  }
  if (mixed) flush_hist(cnt, a.hist, lds, RES);
  // Final FP64 residual for the report.
  double fin[1] = {true_residual_rows(a, lane, nullptr, ctx)};
  grid_sum<1>(fin, a.partials, a.gs, epoch, slot, lds, &lds_i);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.stats[kStIters] = it;
    a.stats[kStRefresh] = refresh;
    a.stats[kStFallback] = force64;
    a.stats[kStRecursive] = rr;
    a.stats[kStTrue] = fin[0];
    a.stats[kStBB] = bb;
    a.stats[kStStatus] = status;
  }
}

// BiCGSTAB (van der Vorst) with the reference's scalar formulas
// (src/Mille-feuille-bicg.h:358-460).  A breakdown (v.r^ = 0, omega = 0 or
// rho = 0) or, for the mixed policies, a stagnating FP64 residual restarts
// the recurrence from the FP64 residual with r^ = r (and every tile FP64);
// more than kMaxRestart restarts report a breakdown.
//
// With a mixed policy the work vectors r, r^, p, s, v and t are FP32 (x,
// the scalars, the dot products and every residual recomputation stay FP64;
// FP64 tiles multiply the FP32 vectors in FP64 arithmetic), as in the mixed
// reference (src/main-bicg-mixed.cu:122-723); after a fallback to FP64 tiles
// the recurrence restarts in FP64 vectors.  FP64 vectors: r, rh, p, s, v in
// a.q, t in a.tv; FP32: r32, rh32, p32, s32, v in q32, t in t32.  With
// resident windows t never goes to memory: the x and r update reads it from
// the row accumulators the SpMV t = A s left it in.
constexpr int kMaxRestart = 8;

template <int MODE, bool RES>
__global__ __launch_bounds__(kBlock) void k_bicgstab(SolveArgs a) {
  __shared__ double lds[kMaxGridSum * kWavesPerBlock];
  __shared__ int lds_i;
  __shared__ typename std::conditional<RES, ResLds, WaveLds>::type rl[kWavesPerBlock];
  const int lane = threadIdx.x & 63;
  const int wid = solver_wave_id();
  const int tr0 = __builtin_amdgcn_readfirstlane(a.wave_tr[wid]);
  const int tr1 = __builtin_amdgcn_readfirstlane(a.wave_tr[wid + 1]);
  const int row0 = tr0 * 16, row1 = tr1 * 16;
  constexpr bool mixed = MODE != kAll64;
  const bool learned = MODE == kPerTile && a.mlp != nullptr;
  unsigned epoch = 0;
  int slot = 0;
  PrecCount cnt;
  double *v = a.q, *t = a.tv;
  // (The rule policy multiplies p and s at different segment levels in one
  // iteration: the host runs it with streamed windows.)
  WaveCtx<RES> ctx;
  ctx.L = &rl[threadIdx.x >> 6];
  solver_stage<MODE>(a, ctx, tr0, tr1, a.seg_p, lane);
  bool lowv = mixed;  // FP32 vectors (until a fallback to FP64 tiles)

  double red[2] = {0.0, 0.0};
  solver_spmv<kAll64, false>(a, a.x, nullptr, nullptr, 0, ctx, lane, nullptr,
    [&](const int i, const double ax) {
      const double bv = a.b[i];
      double rv = bv - ax;
      if (mixed) {
        const float rf = (float)rv;
        a.r32[i] = rf;
        a.rh32[i] = rf;
        a.p32[i] = rf;
        rv = rf;
      } else {
        a.r[i] = rv;
        a.rh[i] = rv;
        a.p[i] = rv;
      }
      red[0] += rv * rv;
      red[1] += bv * bv;
    });
  if (learned) {
    // segment norms of p = r0 for the first inference
    for (int base = row0; base < row1; base += 64) {
      const int i = base + lane;
      const double pv = i < row1 ? (double)a.p32[i] : 0.0;
      const double d2 = group16_sum(pv * pv);
      if ((lane & 15) == 0 && i < row1) a.seg2[i >> 4] = (float)d2;
    }
  }
  grid_sum<2>(red, a.partials, a.gs, epoch, slot, lds, &lds_i);
  double rho = red[0], rr = red[0];
  const double bb = red[1];
  const double stop2 = uniform_double(a.tol2 * bb);
  double rr_ref = rr;
  int it = 0, it_ref = 0, refresh = 0, force64 = 0, status = 0, restarts = 0;
  bool restart = false;
  bool done = bb == 0.0 || (a.fixed_iters <= 0 && rr <= stop2);
  bool infer = learned;  // learned policy: new levels in the next v = A p
  double pp_last = rr;   // ||p||^2 at the last inference point

  while (!done) {
    if (restart) {
      if (++restarts > kMaxRestart) { status = MF_ERR_BREAKDOWN; break; }
      // leaving the FP32 vectors: the FP64 recurrence restarts from here
      if (lowv && force64) lowv = false;
      // x of every wave complete (the sums before this point order values only)
      grid_sync(a.gs, epoch, &lds_i);
      double tr[1] = {restart_rows(a, lane, mixed, true, ctx, lowv)};
      grid_sum<1>(tr, a.partials, a.gs, epoch, slot, lds, &lds_i);
      ++refresh;
      rr = rho = tr[0];
      rr_ref = rr;
      it_ref = it;
      restart = false;
      if (rr <= stop2) break;
    }
    ++it;
    // v = A p, v.rh
    double vr[1] = {0.0};
    if (infer && !force64 && solver_infer<RES>(a, infer_ctx(a, pp_last, rr, bb), tr0, tr1, lane))
      solver_stage<MODE>(a, ctx, tr0, tr1, a.seg_p, lane);
    infer = false;
    solver_spmv<MODE, mixed>(a, a.p, a.p32, a.seg_p, lowv ? 2 : force64, ctx, lane, &cnt,
      [&](const int i, const double vv) {
        if (lowv) {
          a.q32[i] = (float)vv;
          vr[0] += vv * (double)a.rh32[i];
        } else {
          v[i] = vv;
          vr[0] += vv * a.rh[i];
        }
      });
    grid_sum_own<RES>(a, vr, epoch, slot, lds, &lds_i);  // next: own rows of s
    if (!(vr[0] != 0.0) || !isfinite(vr[0])) {
      if (mixed) force64 = 1;
      restart = true;
      continue;
    }
    const double alpha = uniform_double(rho / vr[0]);
    // s = r - alpha v (rule policy: |s_new - s_old| per segment)
    const bool upd = MODE == kPerSeg && (it % a.check_interval) == 0;
    for (int base0 = row0; base0 < row1; base0 += 64 * kUpd) {
      double so[kUpd], sn[kUpd];
      if (lowv) {
        float s_[kUpd], r_[kUpd], v_[kUpd];
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base0 + 64 * u + lane;
          const bool ok = i < row1;
          s_[u] = ok && upd ? a.s32[i] : 0.0f;
          r_[u] = ok ? a.r32[i] : 0.0f;
          v_[u] = ok ? a.q32[i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base0 + 64 * u + lane;
          const float sf = (float)((double)r_[u] - alpha * (double)v_[u]);
          if (i < row1) wt_store(a.s32 + i, sf);
          so[u] = s_[u];
          sn[u] = sf;
        }
      } else {
        double s_[kUpd], r_[kUpd], v_[kUpd];
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base0 + 64 * u + lane;
          const bool ok = i < row1;
          s_[u] = ok && upd ? a.s[i] : 0.0;
          r_[u] = ok ? a.r[i] : 0.0;
          v_[u] = ok ? v[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base0 + 64 * u + lane;
          const double sv = r_[u] - alpha * v_[u];
          if (i < row1) wt_store(a.s + i, sv);
          so[u] = s_[u];
          sn[u] = sv;
        }
      }
      if (upd) {
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base0 + 64 * u + lane;
          const double d = group16_max(i < row1 ? fabs(sn[u] - so[u]) : 0.0);
          if ((lane & 15) == 0 && i < row1) a.seg_s[i >> 4] = (uint8_t)rule_level(d);
        }
      }
    }
    grid_sync(a.gs, epoch, &lds_i);
    // t = A s, t.s and t.t
    double tw[2] = {0.0, 0.0};
    solver_spmv<MODE, mixed>(a, a.s, a.s32, a.seg_s, lowv ? 2 : force64, ctx, lane, &cnt,
      [&](const int i, const double tvv) {
        double sv;
        if (lowv) {
          if constexpr (!RES) a.t32[i] = (float)tvv;
          sv = a.s32[i];
        } else {
          if constexpr (!RES) t[i] = tvv;
          sv = a.s[i];
        }
        tw[0] += tvv * sv;
        tw[1] += tvv * tvv;
      });
    grid_sum_own<RES>(a, tw, epoch, slot, lds, &lds_i);  // next: own rows of x and r
    const double omega = uniform_double(tw[1] != 0.0 && isfinite(tw[1]) ? tw[0] / tw[1] : 0.0);
    // x += alpha p + omega s, r = s - omega t, rho' = r.rh, ||r||^2
    double nw[2] = {0.0, 0.0};
    // four 64-row groups per step, loads first (20 in flight per lane)
    for (int base = row0 + lane; base < row1; base += 64 * kUpd) {
      double x_[kUpd], p_[kUpd], s_[kUpd], t_[kUpd], h_[kUpd];
#pragma unroll
      for (int u = 0; u < kUpd; ++u) {
        const int i = base + 64 * u;
        const bool ok = i < row1;
        x_[u] = ok ? a.x[i] : 0.0;
        if (lowv) {
          p_[u] = ok ? (double)a.p32[i] : 0.0;
          s_[u] = ok ? (double)a.s32[i] : 0.0;
          h_[u] = ok ? (double)a.rh32[i] : 0.0;
        } else {
          p_[u] = ok ? a.p[i] : 0.0;
          s_[u] = ok ? a.s[i] : 0.0;
          h_[u] = ok ? a.rh[i] : 0.0;
        }
        if constexpr (RES)
          t_[u] = ok ? ctx.L->acc[res_acc_index(i - row0)] : 0.0;
        else
          t_[u] = ok ? (lowv ? (double)a.t32[i] : t[i]) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kUpd; ++u) {
        const int i = base + 64 * u;
        if (i < row1) {
          wt_store(a.x + i, x_[u] + (alpha * p_[u] + omega * s_[u]));
          double rv = s_[u] - omega * t_[u];
          if (lowv) {
            const float rf = (float)rv;
            wt_store(a.r32 + i, rf);
            rv = rf;
          } else {
            wt_store(a.r + i, rv);
          }
          nw[0] += rv * h_[u];
          nw[1] += rv * rv;
        }
      }
    }
    grid_sum_own<RES>(a, nw, epoch, slot, lds, &lds_i);  // next: own rows of p (or a restart)
    double rho_new = nw[0], rr_new = nw[1];
    if (a.fixed_iters > 0) {
      if (it >= a.fixed_iters) { rr = rr_new; break; }
    } else {
      const bool converged = rr_new <= stop2;
      const bool periodic = mixed && !force64 &&
                            (rr_new <= a.ratio2 * rr_ref || it - it_ref >= a.refresh_every);
      if (converged || periodic || omega == 0.0 || !isfinite(rr_new)) {
        double tr[2] = {0.0, 0.0};
        grid_sync(a.gs, epoch, &lds_i);  // x of every wave complete
        solver_spmv<kAll64, false>(a, a.x, nullptr, nullptr, 0, ctx, lane, nullptr,
          [&](const int i, const double ax) {
            double rv = a.b[i] - ax;
            double hv;
            if (lowv) {
              const float rf = (float)rv;
              a.r32[i] = rf;
              rv = rf;
              hv = a.rh32[i];
            } else {
              a.r[i] = rv;
              hv = a.rh[i];
            }
            tr[0] += rv * rv;
            tr[1] += rv * hv;
          });
        grid_sum<2>(tr, a.partials, a.gs, epoch, slot, lds, &lds_i);
        ++refresh;
        it_ref = it;
        if (tr[0] <= stop2) { rr = tr[0]; break; }
        if (mixed && !force64 && (tr[0] > 0.25 * rr_ref || refresh > a.max_refresh)) {
          force64 = 1;
          restart = true;
        }
        // A recursive residual far below the true one makes beta meaningless.
        if (tr[0] > 4.0 * rr_new) restart = true;
        rr_new = tr[0];
        rho_new = tr[1];
        rr_ref = tr[0];
      }
      if (it >= a.maxiter) { rr = rr_new; status = MF_NOT_CONVERGED; break; }
      if (omega == 0.0 || rho == 0.0 || rho_new == 0.0 || !isfinite(rho_new)) restart = true;
      if (restart) continue;
    }
    const double beta = uniform_double((rho_new / rho) * (alpha / omega));
    rho = rho_new;
    rr = rr_new;
    // p = r + beta (p - omega v)
    const bool updp = MODE == kPerSeg && (it % a.check_interval) == 0;
    const bool lupd = learned && !force64 && (it % a.check_interval) == 0;
    double pp[1] = {0.0};
    for (int base0 = row0; base0 < row1; base0 += 64 * kUpd) {
      double po[kUpd], pn[kUpd];
      if (lowv) {
        float p_[kUpd], r_[kUpd], v_[kUpd];
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base0 + 64 * u + lane;
          const bool ok = i < row1;
          p_[u] = ok ? a.p32[i] : 0.0f;
          r_[u] = ok ? a.r32[i] : 0.0f;
          v_[u] = ok ? a.q32[i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base0 + 64 * u + lane;
          const float pf =
              (float)((double)r_[u] + beta * ((double)p_[u] - omega * (double)v_[u]));
          if (i < row1) wt_store(a.p32 + i, pf);
          po[u] = p_[u];
          pn[u] = pf;
        }
      } else {
        double p_[kUpd], r_[kUpd], v_[kUpd];
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base0 + 64 * u + lane;
          const bool ok = i < row1;
          p_[u] = ok ? a.p[i] : 0.0;
          r_[u] = ok ? a.r[i] : 0.0;
          v_[u] = ok ? v[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kUpd; ++u) {
          const int i = base0 + 64 * u + lane;
          const double pv = r_[u] + beta * (p_[u] - omega * v_[u]);
          if (i < row1) wt_store(a.p + i, pv);
          po[u] = p_[u];
          pn[u] = pv;
        }
      }
#pragma unroll
      for (int u = 0; u < kUpd; ++u) {
        const int i = base0 + 64 * u + lane;
        double d = i < row1 ? (lupd ? pn[u] * pn[u] : fabs(pn[u] - po[u])) : 0.0;
        if (updp) {
          d = group16_max(d);
          if ((lane & 15) == 0 && i < row1) a.seg_p[i >> 4] = (uint8_t)rule_level(d);
        }
        if (lupd) {
          pp[0] += d;
          d = group16_sum(d);
          if ((lane & 15) == 0 && i < row1) a.seg2[i >> 4] = (float)d;
        }
      }
    }
    if (lupd) {
      grid_sum<1>(pp, a.partials, a.gs, epoch, slot, lds, &lds_i);
      pp_last = pp[0];
      infer = true;
    } else {
      grid_sync(a.gs, epoch, &lds_i);
    }
  }
  if (mixed) flush_hist(cnt, a.hist, lds, RES);
  grid_sync(a.gs, epoch, &lds_i);  // x of every wave complete
  double fin[1] = {true_residual_rows(a, lane, nullptr, ctx)};
  grid_sum<1>(fin, a.partials, a.gs, epoch, slot, lds, &lds_i);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.stats[kStIters] = it;
    a.stats[kStRefresh] = refresh;
    a.stats[kStFallback] = force64;
    a.stats[kStRecursive] = rr;
    a.stats[kStTrue] = fin[0];
    a.stats[kStBB] = bb;
    a.stats[kStStatus] = status;
  }
}

// Error-bound policy: every tile takes the cheapest payload whose largest
// quantisation error is within `theta` of the tile's largest magnitude
// (INT8, then FP16, then FP32).  With theta at FP32 rounding (2^-24) a tile
// drops below FP32 only when INT8 or FP16 represent it as well as FP32 does,
// e.g. the constant couplings of a stencil.
__global__ void k_errbound_levels(int T, const double *__restrict__ qerr,
                                  const double *__restrict__ amax, uint8_t *__restrict__ prec,
                                  double theta) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const double lim = theta * amax[t];
  uint8_t l = MF_PREC_FP32;
  if (qerr[2 * T + t] <= lim) l = MF_PREC_INT8;
  else if (qerr[T + t] <= lim) l = MF_PREC_FP16;
  prec[t] = l;
}

// The FP32 / FP16 / INT8 levels of each solver wave's tiles (tile-rows
// wave_tr[w] .. wave_tr[w+1]) set to the most precise of them, so that each
// window streams one level group instead of two or three (on the 3D 7-point
// 128^3 Laplacian the FP32 diagonal / INT8 off-diagonal split of the error
// bound costs 21.3 ms per CG solve, every tile FP32 19.3 ms, on MI355X).
// One wave per solver wave; FP64 and SKIP tiles keep their level.
__global__ void k_wave_uniform_levels(uint8_t *__restrict__ prec, const int *__restrict__ tile_ptr,
                                      const int *__restrict__ wave_tr, int nwaves) {
  const int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);  // alike in the wave
  const int lane = threadIdx.x & 63;
  if (w >= nwaves) return;
  const int t0 = tile_ptr[wave_tr[w]], t1 = tile_ptr[wave_tr[w + 1]];
  int m = MF_PREC_SKIP;
  for (int t = t0 + lane; t < t1; t += kWave) {
    const int l = prec[t];
    if (l >= MF_PREC_FP32 && l <= MF_PREC_INT8) m = min(m, l);
  }
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) m = min(m, __shfl_xor(m, k));
  for (int t = t0 + lane; t < t1; t += kWave) {
    const int l = prec[t];
    if (l >= MF_PREC_FP32 && l <= MF_PREC_INT8 && l != m) prec[t] = (uint8_t)m;
  }
}

mf_status wave_uniform_levels(mf_tiled *t, const int *d_wave_tr, int nwaves, void *stream) {
  if (nwaves > 0 && t->tilenum > 0) {
    k_wave_uniform_levels<<<(nwaves + 3) / 4, 256, 0, (hipStream_t)stream>>>(t->dev.prec, t->dev.tile_ptr,
                                                                             d_wave_tr, nwaves);
    MF_HIP(hipGetLastError());
  }
  return MF_OK;
}

// ------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------

// Error-bound levels of every tile of t into t->dev.prec (stream-ordered).
mf_status apply_errbound_levels(mf_tiled *t, void *stream) {
  if (t->tilenum > 0) {
    k_errbound_levels<<<(t->tilenum + 255) / 256, 256, 0, (hipStream_t)stream>>>(
        t->tilenum, t->dev.qerr, t->dev.amax, t->dev.prec, 0x1p-24);
    MF_HIP(hipGetLastError());
  }
  return MF_OK;
}

namespace {

struct Workspace {
  double *r, *p, *q, *rh, *s, *tv;
  float *p32, *s32, *r32, *q32, *rh32 = nullptr, *t32 = nullptr;
  uint8_t *seg_p, *seg_s;
  float *seg2;
  double *partials;
  unsigned long long *mbox;
  GridSync *gs;
  double *stats;
  unsigned long long *hist;
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Successive work vectors start this many bytes further apart than their
// sizes: with power-of-two lengths, rows of equal index in different vectors
// would otherwise share their address bits below the vector size, and so the
// same HBM channels, in every streaming pass over several vectors.
constexpr size_t kVecSkew = 4096 + 256;

// Carves the handle's device workspace (grown on demand).
mf_status get_workspace(mf_tiled *t, int G, Workspace &w) {
  const size_t n = t->dev.npad, segs = t->tilen > t->tilem ? t->tilen : t->tilem;
  size_t need = 6 * align256(n * 8) + 6 * align256(n * 4) + 12 * kVecSkew + 2 * align256(segs) + align256(segs * 4) +
                align256(2 * kMaxGridSum * (size_t)G * 8) + align256(((size_t)G + 1) * kMboxWords * 8) +
                align256(sizeof(GridSync)) + align256(kStCount * 8) +
                align256(5 * 8);
  if (t->work_bytes < need) {
    if (t->work) MF_HIP(hipFree(t->work));
    t->work = nullptr;
    t->work_bytes = 0;
    MF_HIP(hipMalloc(&t->work, need));
    t->work_bytes = need;
  }
  char *c = (char *)t->work;
  auto take = [&](size_t bytes) { char *p = c; c += align256(bytes); return (void *)p; };
  auto vec = [&](size_t bytes) { void *p = take(bytes); c += kVecSkew; return p; };
  w.r = (double *)vec(n * 8);
  w.p = (double *)vec(n * 8);
  w.q = (double *)vec(n * 8);
  w.rh = (double *)vec(n * 8);
  w.s = (double *)vec(n * 8);
  w.tv = (double *)vec(n * 8);
  w.p32 = (float *)vec(n * 4);
  w.s32 = (float *)vec(n * 4);
  w.r32 = (float *)vec(n * 4);
  w.q32 = (float *)vec(n * 4);
  w.rh32 = (float *)vec(n * 4);
  w.t32 = (float *)vec(n * 4);
  w.seg_p = (uint8_t *)take(segs);
  w.seg_s = (uint8_t *)take(segs);
  w.seg2 = (float *)take(segs * 4);
  w.partials = (double *)take(2 * kMaxGridSum * (size_t)G * 8);
  w.mbox = (unsigned long long *)take(((size_t)G + 1) * kMboxWords * 8);
  w.gs = (GridSync *)take(sizeof(GridSync));
  w.stats = (double *)take(kStCount * 8);
  w.hist = (unsigned long long *)take(5 * 8);
  return MF_OK;
}

template <typename K>
mf_status grid_for(K kernel, int override_blocks, int *G, int block = kBlock, bool one_per_cu = false) {
  int dev = 0, ncu = 0, nb = 0;
  MF_HIP(hipGetDevice(&dev));
  MF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  MF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, block, 0));
  if (one_per_cu && nb > 1) nb = 1;
  if (nb < 1) return fail(MF_ERR_COOP, "persistent kernel does not fit on a CU");
  const int maxg = nb * ncu;
  *G = override_blocks > 0 ? std::min(override_blocks, maxg) : maxg;
  if (*G > kMaxBlocks) *G = kMaxBlocks;
  return MF_OK;
}

// Per-wave tile-row ranges, balanced on each tile-row's SpMV and vector-update
// work, and each wave's SpMV descriptors.  A tile-row larger than one window
// and than half a wave's share of the SpMV work is heavy: it is cut into
// pieces spread over the waves with the least work, so hub rows of power-law
// matrices do not serialise on one wave.  Cached per wave count.
// Host side of a solver schedule: per-wave tile-row bounds, resident flags,
// the waves' SpMV descriptors and the heavy tile-row pieces.
struct HostSchedule {
  std::vector<int> b, res, ww, hn, hbase, htr, wh;
  std::vector<WaveDesc> wd;
  int npieces = 0, all_resident = 1;
};

void build_schedule(const mf_tiled *t, int nwaves, HostSchedule &hsch) {
  const int tm = t->tilem;
  auto slots = [&](int r) { return t->pay_off[t->tile_ptr[r + 1]] - t->pay_off[t->tile_ptr[r]]; };
  auto tiles = [&](int r) { return t->tile_ptr[r + 1] - t->tile_ptr[r]; };
  // a tile-row's work: its payload slots and tiles in the SpMV, and its 16
  // rows in the vector updates (about 128 slots' worth)
  auto cost = [&](int r) { return (double)slots(r) + 8.0 * tiles(r) + 128.0; };
  std::vector<char> heavy(tm, 0);
  std::vector<int> b;
  {
    double total = 0.0;
    for (int r = 0; r < tm; ++r) total += cost(r);
    const double share = total / nwaves;
    for (int r = 0; r < tm; ++r)
      heavy[r] = (tiles(r) > kWinTiles || slots(r) > 4 * kWinChunks) && cost(r) > 0.5 * share;
    // row ownership balanced on the work that stays with the owner
    std::vector<double> pre(tm + 1, 0.0);
    for (int r = 0; r < tm; ++r) pre[r + 1] = pre[r] + (heavy[r] ? 16.0 : cost(r));
    b.assign(nwaves + 1, tm);
    b[0] = 0;
    int r = 0;
    for (int w = 1; w < nwaves; ++w) {
      const double target = pre[tm] * (double)w / (double)nwaves;
      while (r < tm && pre[r] < target) ++r;
      b[w] = std::max(r, b[w - 1]);
    }
    // The vector passes take a wave's rows 64 * kUpd at a time, so a wave
    // one tile-row over a multiple of 4 * kUpd tile-rows pays a whole extra
    // pass (about 2048 slots' worth), which the additive cost above does not
    // see.  An equal split of tile-rows is kept instead when its busiest wave,
    // counted that way, has less work (regular matrices whose boundary
    // tile-rows are lighter than the rest, such as stencils).
    auto busiest = [&](const std::vector<int> &bb) {
      double m = 0.0;
      for (int w = 0; w < nwaves; ++w) {
        double c = 0.0;
        for (int q = bb[w]; q < bb[w + 1]; ++q)
          if (!heavy[q]) c += (double)slots(q) + 8.0 * tiles(q);
        const int passes = (16 * (bb[w + 1] - bb[w]) + 64 * kUpd - 1) / (64 * kUpd);
        m = std::max(m, c + 2048.0 * passes);
      }
      return m;
    };
    std::vector<int> even(nwaves + 1);
    for (int w = 0; w <= nwaves; ++w) even[w] = (int)((long long)tm * w / nwaves);
    if (busiest(even) < busiest(b)) b = even;
  }
  // descriptors of each wave's non-heavy tile-rows, and the wave's load
  std::vector<std::vector<WaveDesc>> per(nwaves);
  std::vector<double> load(nwaves, 0.0);
  for (int w = 0; w < nwaves; ++w) {
    int r = b[w];
    while (r < b[w + 1]) {
      if (heavy[r]) { ++r; continue; }
      int e = r;
      while (e < b[w + 1] && !heavy[e]) load[w] += cost(e++);
      window_partition_range(t, r, e, 4 * kWinChunks, per[w]);
      r = e;
    }
  }
  // heavy tile-rows: one-window pieces, each to the least loaded wave
  // (a piece spans up to 8 windows, so a hub tile-row of 10^5 tiles is
  // about 100 pieces)
  constexpr int kPieceTiles = 8 * kWinTiles, kPieceSlots = 8 * 4 * kWinChunks;
  std::vector<int> hn, hbase;
  int npieces = 0;
  std::priority_queue<std::pair<double, int>, std::vector<std::pair<double, int>>,
                      std::greater<std::pair<double, int>>> pq;
  for (int w = 0; w < nwaves; ++w) pq.push({load[w], w});
  for (int r = 0; r < tm; ++r) {
    if (!heavy[r]) continue;
    const int h = (int)hn.size();
    std::vector<WaveDesc> pcs;
    const int np = cut_pieces(t, r, h, kPieceTiles, kPieceSlots, 0, pcs);
    for (const WaveDesc &d : pcs) {
      auto top = pq.top();
      pq.pop();
      per[top.second].push_back(d);
      top.first += (double)(d.Z1 - d.Z0) + 8.0 * (d.T1 - d.ta);
      pq.push(top);
    }
    hn.push_back(np);
    hbase.push_back(npieces);
    hsch.htr.push_back(r);
    npieces += np;
  }
  // heavy tile-rows each wave owns (they are numbered in tile-row order)
  hsch.wh.assign(nwaves + 1, 0);
  {
    size_t h = 0;
    for (int w = 0; w < nwaves; ++w) {
      hsch.wh[w] = (int)h;
      while (h < hsch.htr.size() && hsch.htr[h] < b[w + 1]) ++h;
    }
    hsch.wh[nwaves] = (int)h;
  }
  std::vector<int> ww(nwaves + 1, 0);
  std::vector<WaveDesc> wd;
  for (int w = 0; w < nwaves; ++w) {
    wd.insert(wd.end(), per[w].begin(), per[w].end());
    ww[w + 1] = (int)wd.size();
  }
  // resident windows: a non-empty tile-row range without heavy tile-rows or
  // pieces of other waves' heavy tile-rows, within the LDS window limits
  std::vector<int> res(nwaves, 0);
  for (int w = 0; w < nwaves; ++w) {
    const int r0 = b[w], r1 = b[w + 1];
    if (r1 <= r0) continue;
    bool ok = true;
    for (int r = r0; r < r1 && ok; ++r) ok = !heavy[r];
    for (const WaveDesc &d : per[w]) ok = ok && !(d.flags & kDescPiece);
    const int tiles = t->tile_ptr[r1] - t->tile_ptr[r0];
    const int chunks = (t->pay_off[t->tile_ptr[r1]] - t->pay_off[t->tile_ptr[r0]]) / 4;
    res[w] = ok && res_fits(r1 - r0, tiles, chunks) ? 1 : 0;
  }
  hsch.all_resident = 1;
  for (int w = 0; w < nwaves; ++w)
    if (!res[w] && b[w + 1] > b[w]) hsch.all_resident = 0;
  if (wd.empty()) wd.push_back(WaveDesc{0, 0, 0, 0, 0, 0, 0, 0});
  hsch.b = std::move(b);
  hsch.res = std::move(res);
  hsch.ww = std::move(ww);
  hsch.wd = std::move(wd);
  hsch.hn = std::move(hn);
  hsch.hbase = std::move(hbase);
  hsch.npieces = npieces;
}

mf_status ensure_schedule(mf_tiled *t, int nwaves) {
  const int key = nwaves;
  if (t->sched.wave_tr && t->sched_waves_for == key) return MF_OK;
  for (void *p : {(void *)t->sched.wave_tr, (void *)t->sched.wave_res, (void *)t->sched.wave_win, (void *)t->sched.win_desc,
                  (void *)t->sched.heavy_part, (void *)t->sched.heavy_cnt, (void *)t->sched.heavy_n,
                  (void *)t->sched.heavy_base, (void *)t->sched.heavy_tr, (void *)t->sched.wave_heavy})
    if (p) MF_HIP(hipFree(p));
  t->sched = Schedule();
  t->sched_waves_for = -1;
  HostSchedule hsch;
  build_schedule(t, nwaves, hsch);
  const std::vector<int> &b = hsch.b, &res = hsch.res, &ww = hsch.ww, &hn = hsch.hn, &hbase = hsch.hbase;
  const std::vector<WaveDesc> &wd = hsch.wd;
  const int npieces = hsch.npieces;
  t->sched.all_resident = hsch.all_resident;
  const int nh = (int)hn.size();
  MF_HIP(hipMalloc(&t->sched.wave_tr, sizeof(int) * (nwaves + 1)));
  MF_HIP(hipMemcpy(t->sched.wave_tr, b.data(), sizeof(int) * (nwaves + 1), hipMemcpyHostToDevice));
  MF_HIP(hipMalloc(&t->sched.wave_res, sizeof(int) * nwaves));
  MF_HIP(hipMemcpy(t->sched.wave_res, res.data(), sizeof(int) * nwaves, hipMemcpyHostToDevice));
  MF_HIP(hipMalloc(&t->sched.wave_win, sizeof(int) * (nwaves + 1)));
  MF_HIP(hipMemcpy(t->sched.wave_win, ww.data(), sizeof(int) * (nwaves + 1), hipMemcpyHostToDevice));
  MF_HIP(hipMalloc(&t->sched.win_desc, sizeof(WaveDesc) * wd.size()));
  MF_HIP(hipMemcpy(t->sched.win_desc, wd.data(), sizeof(WaveDesc) * wd.size(), hipMemcpyHostToDevice));
  MF_HIP(hipMalloc(&t->sched.heavy_part, sizeof(double) * 16 * std::max(npieces, 1)));
  MF_HIP(hipMalloc(&t->sched.heavy_cnt, sizeof(unsigned) * std::max(nh, 1)));
  MF_HIP(hipMalloc(&t->sched.heavy_n, sizeof(int) * std::max(nh, 1)));
  MF_HIP(hipMalloc(&t->sched.heavy_base, sizeof(int) * std::max(nh, 1)));
  MF_HIP(hipMalloc(&t->sched.heavy_tr, sizeof(int) * std::max(nh, 1)));
  MF_HIP(hipMalloc(&t->sched.wave_heavy, sizeof(int) * (nwaves + 1)));
  MF_HIP(hipMemcpy(t->sched.wave_heavy, hsch.wh.data(), sizeof(int) * (nwaves + 1), hipMemcpyHostToDevice));
  if (nh > 0) {
    MF_HIP(hipMemcpy(t->sched.heavy_n, hn.data(), sizeof(int) * nh, hipMemcpyHostToDevice));
    MF_HIP(hipMemcpy(t->sched.heavy_base, hbase.data(), sizeof(int) * nh, hipMemcpyHostToDevice));
    MF_HIP(hipMemcpy(t->sched.heavy_tr, hsch.htr.data(), sizeof(int) * nh, hipMemcpyHostToDevice));
  }
  t->sched.nheavy = nh;
  t->sched.nwaves = nwaves;
  t->sched_waves_for = key;
  return MF_OK;
}

int spmv_mode_for(int policy) {
  switch (policy) {
    case MF_POLICY_FP64: return kAll64;
    case MF_POLICY_RULE: return kPerSeg;
    default: return kPerTile;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Allocation-free solves (mf_cg_solve_noalloc): the workspace and schedule
// live in one file-scope device array of kScratchBytes per device, whose
// address (hipGetSymbolAddress, once per device) replaces the handle's
// allocations.  Calls on one device share it: they must not overlap.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)12 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kScratchDevices] = {};

// The scratch array's device address on device `dev` (current).
mf_status scratch_address(int dev, unsigned char **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0 || dev >= kScratchDevices) return fail(MF_ERR_ARG, "noalloc: device index out of range");
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(MF_ERR_HIP, std::string("noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = (unsigned char *)g_scratch_ptr[dev];
  return MF_OK;
}

// Fill of n 32-bit words (a kernel, not a memset: stays valid under capture).
__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

mf_status fill_u32(void *p, size_t bytes, uint32_t v, hipStream_t st) {
  const size_t n = bytes / 4;
  if (n == 0) return MF_OK;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 1024);
  k_fill_u32<<<blocks, 256, 0, st>>>((uint32_t *)p, n, v);
  MF_HIP(hipGetLastError());
  return MF_OK;
}

// Carves the workspace and the schedule of a solve out of the scratch array;
// fails when they do not fit.  Host schedule arrays are copied in.
mf_status scratch_layout(const mf_tiled *t, int G, const HostSchedule &hs, unsigned char *base,
                         Workspace &w, Schedule &sc, hipStream_t st) {
  const size_t n = t->dev.npad, segs = t->tilen > t->tilem ? t->tilen : t->tilem;
  const int nwaves = (int)hs.b.size() - 1, nh = (int)hs.hn.size();
  size_t off = 0;
  auto take = [&](size_t bytes) { const size_t o = off; off += align256(bytes); return o; };
  const size_t o_vec = take(6 * align256(n * 8) + 4 * align256(n * 4));
  const size_t o_segp = take(segs), o_segs = take(segs), o_seg2 = take(segs * 4);
  const size_t o_part = take(2 * kMaxGridSum * (size_t)G * 8);
  const size_t o_mbox = take(((size_t)G + 1) * kMboxWords * 8);
  const size_t o_gs = take(sizeof(GridSync)), o_stats = take(kStCount * 8), o_hist = take(5 * 8);
  const size_t o_tr = take(sizeof(int) * (nwaves + 1)), o_res = take(sizeof(int) * nwaves);
  const size_t o_win = take(sizeof(int) * (nwaves + 1)), o_desc = take(sizeof(WaveDesc) * hs.wd.size());
  const size_t o_hpart = take(sizeof(double) * 16 * std::max(hs.npieces, 1));
  const size_t o_hcnt = take(sizeof(unsigned) * std::max(nh, 1)), o_hn = take(sizeof(int) * std::max(nh, 1));
  const size_t o_hbase = take(sizeof(int) * std::max(nh, 1));
  const size_t o_htr = take(sizeof(int) * std::max(nh, 1)), o_wh = take(sizeof(int) * (nwaves + 1));
  if (off > kScratchBytes)
    return fail(MF_ERR_ARG, "mf_cg_solve_noalloc: the system needs " + std::to_string(off) +
                                " bytes of workspace, more than the " + std::to_string(kScratchBytes) +
                                " bytes of the allocation-free scratch (about 150,000 rows)");
  unsigned char *v = base + o_vec;
  w.r = (double *)v; v += align256(n * 8);
  w.p = (double *)v; v += align256(n * 8);
  w.q = (double *)v; v += align256(n * 8);
  w.rh = (double *)v; v += align256(n * 8);
  w.s = (double *)v; v += align256(n * 8);
  w.tv = (double *)v; v += align256(n * 8);
  w.p32 = (float *)v; v += align256(n * 4);
  w.s32 = (float *)v; v += align256(n * 4);
  w.r32 = (float *)v; v += align256(n * 4);
  w.q32 = (float *)v;
  w.seg_p = base + o_segp;
  w.seg_s = base + o_segs;
  w.seg2 = (float *)(base + o_seg2);
  w.partials = (double *)(base + o_part);
  w.mbox = (unsigned long long *)(base + o_mbox);
  w.gs = (GridSync *)(base + o_gs);
  w.stats = (double *)(base + o_stats);
  w.hist = (unsigned long long *)(base + o_hist);
  sc = Schedule();
  sc.nwaves = nwaves;
  sc.nheavy = nh;
  sc.all_resident = hs.all_resident;
  sc.wave_tr = (int *)(base + o_tr);
  sc.wave_res = (int *)(base + o_res);
  sc.wave_win = (int *)(base + o_win);
  sc.win_desc = (WaveDesc *)(base + o_desc);
  sc.heavy_part = (double *)(base + o_hpart);
  sc.heavy_cnt = (unsigned *)(base + o_hcnt);
  sc.heavy_n = (int *)(base + o_hn);
  sc.heavy_base = (int *)(base + o_hbase);
  sc.heavy_tr = (int *)(base + o_htr);
  sc.wave_heavy = (int *)(base + o_wh);
  MF_HIP(hipMemcpyAsync(sc.wave_heavy, hs.wh.data(), sizeof(int) * (nwaves + 1), hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(sc.wave_tr, hs.b.data(), sizeof(int) * (nwaves + 1), hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(sc.wave_res, hs.res.data(), sizeof(int) * nwaves, hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(sc.wave_win, hs.ww.data(), sizeof(int) * (nwaves + 1), hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(sc.win_desc, hs.wd.data(), sizeof(WaveDesc) * hs.wd.size(), hipMemcpyHostToDevice, st));
  if (nh > 0) {
    MF_HIP(hipMemcpyAsync(sc.heavy_n, hs.hn.data(), sizeof(int) * nh, hipMemcpyHostToDevice, st));
    MF_HIP(hipMemcpyAsync(sc.heavy_base, hs.hbase.data(), sizeof(int) * nh, hipMemcpyHostToDevice, st));
    MF_HIP(hipMemcpyAsync(sc.heavy_tr, hs.htr.data(), sizeof(int) * nh, hipMemcpyHostToDevice, st));
  }
  // the scratch keeps the last call's contents: clear what starts at zero
  MF_TRY(fill_u32(base + o_vec, o_segp - o_vec, 0u, st));  // vectors (zero padding)
  MF_TRY(fill_u32(base + o_segp, o_part - o_segp, 0u, st));  // segment levels (FP64 = 0), norms
  MF_TRY(fill_u32(base + o_mbox, o_tr - o_mbox, 0u, st));  // mailboxes, barrier, stats, histogram
  MF_TRY(fill_u32(base + o_hcnt, align256(sizeof(unsigned) * std::max(nh, 1)), 0u, st));
  return MF_OK;
}

mf_status run_persistent(mf_tiled *t, const double *d_b, double *d_x, const mf_solve_opts *o,
                         mf_report *rep, bool bicg, bool noalloc = false) {
  if (!t || !d_b || !d_x) return fail(MF_ERR_ARG, "solve: null argument");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, "solve: host-only handle");
  mf_solve_opts opt;
  mf_solve_opts_default(&opt);
  if (o) opt = *o;
  if (opt.maxiter <= 0) opt.maxiter = 10000;
  if (opt.check_interval <= 0) opt.check_interval = 8;
  if (!(opt.refresh_ratio > 0.0) || opt.refresh_ratio >= 1.0) opt.refresh_ratio = 0.1;
  if (opt.max_refresh <= 0) opt.max_refresh = 64;
  if (opt.refresh_every <= 0) opt.refresh_every = 64;
  if (!(opt.tol > 0.0)) opt.tol = 1e-10;
  if (opt.policy < MF_POLICY_FP64 || opt.policy > MF_POLICY_ERRBOUND)
    return fail(MF_ERR_ARG, "solve: unknown policy");
  // Without an installed MLP the learned policy runs the reference's rule.
  if (opt.policy == MF_POLICY_LEARNED && !t->dev.has_mlp) opt.policy = MF_POLICY_RULE;
  MF_HIP(hipSetDevice(t->dev.device));
  hipStream_t st = (hipStream_t)opt.stream;
  const int mode = spmv_mode_for(opt.policy);

  const void *kern;
  int G = 0, block = kBlock;
  mf_status sst = MF_OK;
  bool resident_windows = false;
  HostSchedule nsched;  // allocation-free solves: the schedule built on the host
  if (opt.form < MF_FORM_AUTO || opt.form > MF_FORM_RESIDENT) return fail(MF_ERR_ARG, "solve: unknown form");
  {
    // The grid is at most the occupancy of the streamed kernel (one block of
    // 16 waves per CU for both forms).  Unless the memory form is asked for,
    // the resident-window kernel runs when every wave's share fits a
    // resident window (not for the rule policy's BiCGSTAB, which multiplies
    // p and s at different segment levels).
    const void *kstream = bicg ? (mode == kAll64 ? (const void *)k_bicgstab<kAll64, false>
                                  : mode == kPerSeg ? (const void *)k_bicgstab<kPerSeg, false>
                                                    : (const void *)k_bicgstab<kPerTile, false>)
                               : (mode == kAll64 ? (const void *)k_cg<kAll64, false>
                                  : mode == kPerSeg ? (const void *)k_cg<kPerSeg, false>
                                                    : (const void *)k_cg<kPerTile, false>);
    // Without an override, a small system gets about 1024 rows per block (at
    // least 4 blocks; whole multiples of 8 from 8 on, for the XCD-contiguous
    // block order), so the grid-wide sums and syncs span fewer blocks: on
    // MI355X an FP64 CG iteration of a 4096-row system takes 10.1 us on 4
    // blocks against 13.5 us on the full grid, and of a 32768-row one 11.1 us
    // on 32 blocks against 14.4 us.  From about 256k rows on, the full grid.
    int want = opt.blocks;
    if (want <= 0) {
      const int rows = t->tilem * 16;
      want = std::max(4, (rows + 1023) / 1024);
      if (want >= 8) want = (want + 7) / 8 * 8;
    }
    sst = grid_for(kstream, want, &G);
    if (sst != MF_OK) return sst;
    if (noalloc) {
      build_schedule(t, G * kWavesPerBlock, nsched);
    } else {
      sst = ensure_schedule(t, G * kWavesPerBlock);
      if (sst != MF_OK) return sst;
    }
    kern = kstream;
    const int all_res = noalloc ? nsched.all_resident : t->sched.all_resident;
    if (all_res && opt.form != MF_FORM_MEMORY && !(bicg && mode == kPerSeg)) {
      int Gr = 0;
      const void *kres = bicg ? (mode == kAll64 ? (const void *)k_bicgstab<kAll64, true>
                                                : (const void *)k_bicgstab<kPerTile, true>)
                              : (mode == kAll64 ? (const void *)k_cg<kAll64, true>
                                 : mode == kPerSeg ? (const void *)k_cg<kPerSeg, true>
                                                   : (const void *)k_cg<kPerTile, true>);
      sst = grid_for(kres, want, &Gr);
      if (sst != MF_OK) return sst;
      if (Gr == G) {
        kern = kres;
        resident_windows = true;
      }
    }
  }
  Workspace w;
  Schedule scr_sched;
  const Schedule *sch = &t->sched;
  const size_t n = t->dev.npad, segs = t->tilen > t->tilem ? t->tilen : t->tilem;
  if (noalloc) {
    unsigned char *base = nullptr;
    MF_TRY(scratch_address(t->dev.device, &base));
    MF_TRY(scratch_layout(t, G, nsched, base, w, scr_sched, st));
    sch = &scr_sched;
  } else {
    sst = get_workspace(t, G, w);
    if (sst != MF_OK) return sst;
    MF_HIP(hipMemsetAsync(w.gs, 0, sizeof(GridSync), st));
    MF_HIP(hipMemsetAsync(w.mbox, 0, ((size_t)G + 1) * kMboxWords * 8, st));
    if (t->sched.nheavy > 0)
      MF_HIP(hipMemsetAsync(t->sched.heavy_cnt, 0, sizeof(unsigned) * t->sched.nheavy, st));
    MF_HIP(hipMemsetAsync(w.hist, 0, 5 * 8, st));
    MF_HIP(hipMemsetAsync(w.stats, 0, kStCount * 8, st));
    MF_HIP(hipMemsetAsync(w.seg_p, MF_PREC_FP64, segs, st));
    MF_HIP(hipMemsetAsync(w.seg_s, MF_PREC_FP64, segs, st));
    // Padding entries of the work vectors must be zero: tiles whose columns
    // lie beyond the rows gather them.
    MF_HIP(hipMemsetAsync(w.r, 0, (size_t)((char *)w.tv - (char *)w.r) + n * 8, st));
    MF_HIP(hipMemsetAsync(w.p32, 0, (size_t)((char *)w.t32 - (char *)w.p32) + n * 4, st));
  }
  if (opt.policy == MF_POLICY_FIXED)
    MF_HIP(hipMemcpyAsync(t->dev.prec, t->dev.prec_user, t->tilenum, hipMemcpyDeviceToDevice, st));
  if (opt.policy == MF_POLICY_ERRBOUND && t->tilenum > 0) {
    k_errbound_levels<<<(t->tilenum + 255) / 256, 256, 0, st>>>(t->tilenum, t->dev.qerr, t->dev.amax,
                                                               t->dev.prec, 0x1p-24);
    MF_HIP(hipGetLastError());
    MF_TRY(wave_uniform_levels(t, sch->wave_tr, G * kWavesPerBlock, st));
  }

  SolveArgs a;
  a.A = view_of(t);
  a.b = d_b;
  a.x = d_x;
  a.r = w.r; a.p = w.p; a.q = w.q; a.rh = w.rh; a.s = w.s; a.tv = w.tv;
  a.p32 = w.p32; a.s32 = w.s32; a.r32 = w.r32; a.q32 = w.q32; a.rh32 = w.rh32; a.t32 = w.t32;
  a.seg_p = w.seg_p; a.seg_s = w.seg_s;
  a.wave_tr = sch->wave_tr;
  a.wave_win = sch->wave_win;
  a.win_desc = sch->win_desc;
  a.heavy_part = sch->heavy_part;
  a.heavy_cnt = sch->heavy_cnt;
  a.heavy_n = sch->heavy_n;
  a.heavy_base = sch->heavy_base;
  a.heavy_tr = sch->heavy_tr;
  a.wave_heavy = sch->wave_heavy;
  a.partials = w.partials;
  a.mbox = w.mbox;
  a.gs = w.gs;
  a.stats = w.stats;
  a.hist = w.hist;
  a.tol2 = opt.tol * opt.tol;
  a.maxiter = opt.maxiter;
  a.fixed_iters = opt.fixed_iters;
  a.policy = opt.policy;
  a.check_interval = opt.check_interval;
  a.ratio2 = opt.refresh_ratio * opt.refresh_ratio;
  a.max_refresh = opt.max_refresh;
  a.refresh_every = opt.refresh_every;
  const bool learned = opt.policy == MF_POLICY_LEARNED;
  a.mlp = learned ? policy_mlp(t) : nullptr;
  a.feat16 = t->dev.feat16;
  a.seg2 = w.seg2;
  a.prec_rw = t->dev.prec;
  a.nseg = (int)segs;
  if (learned && !noalloc) MF_HIP(hipMemsetAsync(w.seg2, 0, segs * 4, st));

#ifdef MF_PHASE_TRACE
  const char *trace_file = std::getenv("MF_PHASE_TRACE_FILE");
  const size_t trace_n = (size_t)kTraceIters * kTracePhases * G;
  a.trace = nullptr;
  if (trace_file && !bicg) {
    MF_HIP(hipMalloc(&a.trace, trace_n * 8));
    MF_HIP(hipMemsetAsync(a.trace, 0, trace_n * 8, st));
  }
#endif
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!noalloc) {
    MF_HIP(hipEventCreate(&e0));
    MF_HIP(hipEventCreate(&e1));
    MF_HIP(hipEventRecord(e0, st));
  }
  void *args[] = {&a};
  // The grid never exceeds what the occupancy query says is resident at once.
  // A cooperative launch also guarantees it; when the runtime refuses one
  // (e.g. another process holds the device's cooperative queue) the same grid
  // is launched normally: the barrier needs only atomics, and its bounded
  // spin turns a block that is not resident into a reported abort.
  hipError_t le = hipLaunchCooperativeKernel(kern, dim3(G), dim3(block), args, 0, st);
  if (le == hipErrorIllegalState || le == hipErrorNotSupported) {
    (void)hipGetLastError();
    le = hipLaunchKernel(kern, dim3(G), dim3(block), args, 0, st);
  }
  if (le != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return fail(MF_ERR_COOP, std::string("hipLaunchCooperativeKernel: ") + hipGetErrorString(le));
  }
  float ms = 0.0f;
  if (noalloc) {
    MF_HIP(hipStreamSynchronize(st));  // (no events: the report's time is 0)
  } else {
    MF_HIP(hipEventRecord(e1, st));
    MF_HIP(hipEventSynchronize(e1));
    MF_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }

#ifdef MF_PHASE_TRACE
  if (a.trace) {
    std::vector<unsigned long long> tr(trace_n);
    MF_HIP(hipMemcpy(tr.data(), a.trace, trace_n * 8, hipMemcpyDeviceToHost));
    MF_HIP(hipFree(a.trace));
    if (FILE *f = std::fopen(trace_file, "ab")) {
      const long long hdr[4] = {G, kTraceIters, kTracePhases, opt.policy};
      std::fwrite(hdr, 8, 4, f);
      std::fwrite(tr.data(), 8, trace_n, f);
      std::fclose(f);
    }
  }
#endif
  double stats[kStCount];
  unsigned long long hist[5];
  GridSync gsh;
  MF_HIP(hipMemcpy(&gsh, w.gs, sizeof(gsh), hipMemcpyDeviceToHost));
  if (gsh.abort) return fail(MF_ERR_COOP, "solve: grid barrier timed out (grid not co-resident?)");
  MF_HIP(hipMemcpy(stats, w.stats, sizeof(stats), hipMemcpyDeviceToHost));
  MF_HIP(hipMemcpy(hist, w.hist, sizeof(hist), hipMemcpyDeviceToHost));
  const double bb = stats[kStBB];
  const int status = (int)stats[kStStatus];
  mf_report r;
  std::memset(&r, 0, sizeof(r));
  r.iterations = (int)stats[kStIters];
  r.refreshes = (int)stats[kStRefresh];
  r.fell_back = (int)stats[kStFallback];
  r.rel_residual = bb > 0 ? std::sqrt(stats[kStTrue] / bb) : std::sqrt(stats[kStTrue]);
  r.rel_residual_recursive = bb > 0 ? std::sqrt(stats[kStRecursive] / bb) : 0.0;
  r.time_ms = ms;
  for (int i = 0; i < 5; ++i) r.prec_hist[i] = (long long)hist[i];
  r.grid_blocks = G;
  r.form_used = resident_windows ? MF_FORM_RESIDENT : MF_FORM_MEMORY;
  r.policy_used = opt.policy;
  int ret = status;
  if (ret == MF_OK && opt.fixed_iters <= 0 && !(r.rel_residual <= opt.tol)) ret = MF_NOT_CONVERGED;
  r.status = ret;
  if (rep) *rep = r;
  if (ret == MF_ERR_BREAKDOWN) set_error("solve: breakdown (zero or non-finite denominator)");
  if (ret == MF_NOT_CONVERGED) set_error("solve: tolerance not reached");
  return ret;
}

}  // namespace mf

using namespace mf;

extern "C" void mf_solve_opts_default(mf_solve_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->tol = 1e-10;
  o->maxiter = 10000;
  o->policy = MF_POLICY_FP64;
  o->check_interval = 8;
  o->refresh_ratio = 0.1;
  o->max_refresh = 64;
  o->refresh_every = 64;
}

extern "C" mf_status mf_cg_solve(mf_tiled *t, const double *d_b, double *d_x,
                                 const mf_solve_opts *opts, mf_report *rep) {
  return run_persistent(t, d_b, d_x, opts, rep, false);
}

extern "C" mf_status mf_bicgstab_solve(mf_tiled *t, const double *d_b, double *d_x,
                                       const mf_solve_opts *opts, mf_report *rep) {
  return run_persistent(t, d_b, d_x, opts, rep, true);
}

extern "C" mf_status mf_cg_solve_noalloc(mf_tiled *t, const double *d_b, double *d_x,
                                        const mf_solve_opts *opts, mf_report *rep) {
  return run_persistent(t, d_b, d_x, opts, rep, false, true);
}
