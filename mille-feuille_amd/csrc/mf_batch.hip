// mf_batch.hip -- many independent systems, each with several right-hand
// sides, solved by CG in one persistent launch captured in a hipGraph.
//
// A work item is one (system, right-hand side) pair.  The systems are split
// over 8 queues, one per XCD (blocks are dispatched round-robin, block b on
// XCD b % 8), by the longest-processing-time rule on their work (entries
// times right-hand sides); inside a queue they run largest first, with a
// system's right-hand sides next to each other.  Every block takes the next
// item of its XCD's queue, solves it with its waves and takes another; when
// that queue is empty it takes items from the other queues in turn.  So the
// blocks solving one system's right-hand sides run at the same time on one
// XCD and read the system's tiles through that XCD's L2.
//
// Inside a block the solve is the single-system memory-form CG: the 8 waves
// own nnz-balanced tile-row ranges of the system, the SpMV is the streamed
// wave SpMV, dot products are deterministic block sums and every phase ends
// at a block barrier; there is no grid barrier, so blocks need not be
// co-resident and the launch is an ordinary one.  Precision follows the
// handle's current tile levels (FP64, or the FIXED / ERRBOUND levels set
// before the launch) with FP64 scalars, FP64 x, FP32 work vectors under mixed
// levels, periodic FP64 true-residual recomputation and FP64 fallback on
// stagnation.
//
// Timing: each item records the realtime clock (100 MHz) at start and end;
// a system's time is its last end minus its first start.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "mf_device.h"

namespace mf {

// Waves per (system, right-hand side) item: 4, 8 (default) or 16, picked by
// opts->batch_waves or by mf_tune_batch.
constexpr int kBatchWavesDefault = 8;

struct BatchSys {
  TileView A;
  const int *wave_tr;  // waves per item + 1 tile-row bounds
  const double *b;     // npad * nrhs, right-hand side j at j * npad
  double *x;           // npad * nrhs
  int npad;
  int mixed;           // 1: per-tile levels (A.prec), 0: every tile FP64
};

struct BatchArgs {
  const BatchSys *sys;
  const int2 *items;  // (system, right-hand side), queue by queue
  const int *qoff;    // kQueues + 1: items of queue q at [qoff[q], qoff[q+1])
  int nitems;
  int *head;          // kQueues work-queue heads
  double *work;       // per block: r, p, q (npad_max doubles each), p32
  long long stride;   // doubles per block
  int *iters;
  double *relres;
  unsigned long long *tstamp;  // per item: start, end
  double tol2;
  int maxiter;
  double ratio2;
  int max_refresh;
  int refresh_every;
};

constexpr int kQueues = 8;  // one per XCD

__device__ __forceinline__ unsigned long long realtime() { return __builtin_amdgcn_s_memrealtime(); }

__global__ void k_batch_reset(int *head, int nitems, int *iters, double *relres) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kQueues) head[i] = 0;
  if (i < nitems) {
    iters[i] = 0;
    relres[i] = -1.0;
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch_cg(BatchArgs a) {
  __shared__ double lds[BLOCK / 64];
  __shared__ int item_s;
  __shared__ WaveLds wl[BLOCK / 64];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  WaveLds &L = wl[wv];
  double *r = a.work + (long long)blockIdx.x * a.stride;
  PrecCount cnt;
  int q = blockIdx.x % kQueues, tried = 0;  // this XCD's queue first
  for (;;) {
    if (threadIdx.x == 0) {
      // next item of queue q; an empty queue passes to the next one, and
      // after all kQueues are empty the block ends (every head only grows)
      int it = -1;
      while (tried < kQueues) {
        const int k = atomicAdd(&a.head[q], 1);
        if (a.qoff[q] + k < a.qoff[q + 1]) {
          it = a.qoff[q] + k;
          break;
        }
        q = (q + 1) % kQueues;
        ++tried;
      }
      item_s = it;
    }
    __syncthreads();
    // item, system and wave range are wave-uniform: scalar registers
    const int item = __builtin_amdgcn_readfirstlane(item_s);
    __syncthreads();
    if (item < 0) break;
    const unsigned long long t_start = realtime();
    const int sidx = __builtin_amdgcn_readfirstlane(a.items[item].x);
    const int rhs = __builtin_amdgcn_readfirstlane(a.items[item].y);
    const BatchSys &S = a.sys[sidx];
    const int npad = S.npad;
    const double *b = S.b + (long long)rhs * npad;
    double *x = S.x + (long long)rhs * npad;
    double *p = r + npad, *q = p + npad;
    float *p32 = (float *)(q + npad);
    const TileView &A = S.A;
    const int tr0 = __builtin_amdgcn_readfirstlane(S.wave_tr[wv]);
    const int tr1 = __builtin_amdgcn_readfirstlane(S.wave_tr[wv + 1]);
    const int row0 = tr0 * 16, row1 = tr1 * 16;
    // gathered padding of p beyond the tile-rows must be zero
    for (int i = 16 * A.tilem + threadIdx.x; i < npad; i += BLOCK) {
      p[i] = 0.0;
      p32[i] = 0.0f;
    }
    // Mixed items keep r, p and q in FP32 (r32 and q32 over the first halves
    // of r and q) while their tiles are mixed; x, the scalars and the residual
    // recomputations stay FP64, and FP64 tiles multiply p32 in FP64
    // arithmetic.  A fallback to FP64 tiles restarts in FP64 vectors.
    const bool mixed = S.mixed != 0;
    bool lowv = mixed;
    float *r32 = (float *)r, *q32 = (float *)q;
    // r = b - A x, p = r
    double rr0 = 0.0, bb0 = 0.0;
    wave_spmv<kAll64, false>(A, tr0, tr1, x, nullptr, nullptr, 0, L, lane, nullptr,
      [&](const int i, const double ax) {
        const double bv = b[i];
        double rv = bv - ax;
        if (lowv) {
          const float rf = (float)rv;
          r32[i] = rf;
          p32[i] = rf;
          rv = rf;
        } else {
          r[i] = rv;
          p[i] = rv;
        }
        rr0 += rv * rv;
        bb0 += bv * bv;
      });
    double rr = block_sum(rr0, lds);
    const double bb = block_sum(bb0, lds);
    const double stop2 = a.tol2 * bb;
    double rr_ref = rr;
    int it = 0, it_ref = 0, refresh = 0;
    int force64 = S.mixed ? 0 : 1;
    bool restart = false;
    bool done = bb == 0.0 || rr <= stop2;
    __syncthreads();
    while (!done) {
      if (restart) {
        if (lowv && force64) lowv = false;
        double tr0v = 0.0;
        wave_spmv<kAll64, false>(A, tr0, tr1, x, nullptr, nullptr, 0, L, lane, nullptr,
          [&](const int i, const double ax) {
            double rv = b[i] - ax;
            if (lowv) {
              const float rf = (float)rv;
              r32[i] = rf;
              p32[i] = rf;
              rv = rf;
            } else {
              r[i] = rv;
              p[i] = rv;
            }
            tr0v += rv * rv;
          });
        rr = block_sum(tr0v, lds);
        rr_ref = rr;
        it_ref = it;
        ++refresh;
        restart = false;
        if (rr <= stop2) break;
        __syncthreads();
      }
      ++it;
      double pq0 = 0.0;
      wave_spmv<kPerTile, false, 3, false, true>(A, wave_desc(A, tr0, tr1), p, p32, nullptr,
                                                 lowv ? 2 : force64, L, lane, &cnt,
                                                 [&](const int i, const double qv) {
          if (lowv) {
            q32[i] = (float)qv;
            pq0 += (double)p32[i] * qv;
          } else {
            q[i] = qv;
            pq0 += p[i] * qv;
          }
        });
      const double pq = block_sum(pq0, lds);
      if (mixed && !force64 && !(pq > 0.0 && isfinite(pq))) {
        force64 = 1;
        restart = true;
        continue;
      }
      if (!(pq != 0.0) || !isfinite(pq)) break;
      const double alpha = rr / pq;
      double rn0 = 0.0;
      if (lowv) {
        for (int i = row0 + lane; i < row1; i += 64) {
          x[i] += alpha * (double)p32[i];
          const float rf = (float)((double)r32[i] - alpha * (double)q32[i]);
          r32[i] = rf;
          rn0 += (double)rf * rf;
        }
      } else {
        for (int i = row0 + lane; i < row1; i += 64) {
          x[i] += alpha * p[i];
          const double rv = r[i] - alpha * q[i];
          r[i] = rv;
          rn0 += rv * rv;
        }
      }
      double rr_new = block_sum(rn0, lds);
      const bool converged = rr_new <= stop2;
      const bool periodic = mixed && !force64 &&
                            (rr_new <= a.ratio2 * rr_ref || it - it_ref >= a.refresh_every);
      if (converged || periodic) {
        double tv = 0.0;
        wave_spmv<kAll64, false>(A, tr0, tr1, x, nullptr, nullptr, 0, L, lane, nullptr,
          [&](const int i, const double ax) {
            double rv = b[i] - ax;
            if (lowv) {
              const float rf = (float)rv;
              r32[i] = rf;
              rv = rf;
            } else {
              r[i] = rv;
            }
            tv += rv * rv;
          });
        const double tr = block_sum(tv, lds);
        ++refresh;
        it_ref = it;
        if (tr <= stop2) { rr = tr; break; }
        if (mixed && !force64 && (tr > 0.25 * rr_ref || refresh > a.max_refresh)) {
          force64 = 1;
          restart = true;
        }
        if (tr > 4.0 * rr_new) restart = true;
        rr_new = tr;
        rr_ref = tr;
      }
      if (it >= a.maxiter) { rr = rr_new; break; }
      if (restart) {
        __syncthreads();
        continue;
      }
      const double beta = rr_new / rr;
      rr = rr_new;
      if (lowv) {
        for (int i = row0 + lane; i < row1; i += 64)
          p32[i] = (float)((double)r32[i] + beta * (double)p32[i]);
      } else {
        for (int i = row0 + lane; i < row1; i += 64) p[i] = r[i] + beta * p[i];
      }
      __syncthreads();
    }
    __syncthreads();
    // FP64 relative residual of the result
    double fv = 0.0;
    wave_spmv<kAll64, false>(A, tr0, tr1, x, nullptr, nullptr, 0, L, lane, nullptr,
      [&](const int i, const double ax) {
        const double d = b[i] - ax;
        fv += d * d;
      });
    const double fin = block_sum(fv, lds);
    if (threadIdx.x == 0) {
      a.iters[item] = it;
      a.relres[item] = bb > 0.0 ? sqrt(fin / bb) : sqrt(fin);
      a.tstamp[2 * item] = t_start;
      a.tstamp[2 * item + 1] = realtime();
    }
    __syncthreads();
  }
}

}  // namespace mf

using namespace mf;

extern "C" mf_status mf_batch_cg_solve(mf_tiled **systems, int nsys, int nrhs, const double **d_b,
                                       double **d_x, const mf_solve_opts *o, int *iterations,
                                       double *rel_residual, double *per_system_ms,
                                       double *total_ms) {
  if (!systems || nsys <= 0 || nrhs <= 0 || !d_b || !d_x)
    return fail(MF_ERR_ARG, "mf_batch_cg_solve: bad argument");
  mf_solve_opts opt;
  mf_solve_opts_default(&opt);
  if (o) opt = *o;
  if (opt.maxiter <= 0) opt.maxiter = 10000;
  if (!(opt.refresh_ratio > 0.0) || opt.refresh_ratio >= 1.0) opt.refresh_ratio = 0.1;
  if (opt.max_refresh <= 0) opt.max_refresh = 64;
  if (opt.refresh_every <= 0) opt.refresh_every = 64;
  if (!(opt.tol > 0.0)) opt.tol = 1e-10;
  if (opt.policy != MF_POLICY_FP64 && opt.policy != MF_POLICY_FIXED && opt.policy != MF_POLICY_ERRBOUND)
    return fail(MF_ERR_UNSUPPORTED, "mf_batch_cg_solve: policies FP64, FIXED and ERRBOUND only");
  const int dev = systems[0] ? systems[0]->dev.device : -1;
  for (int i = 0; i < nsys; ++i) {
    if (!systems[i] || systems[i]->dev.device != dev || dev < 0 || !d_b[i] || !d_x[i])
      return fail(MF_ERR_ARG, "mf_batch_cg_solve: systems must be device handles on one device");
    if (systems[i]->tilem != systems[i]->tilen)
      return fail(MF_ERR_ARG, "mf_batch_cg_solve: square systems only");
  }
  MF_HIP(hipSetDevice(dev));
  hipStream_t st = (hipStream_t)opt.stream;
  hipStream_t own = nullptr;
  if (!st) {
    MF_HIP(hipStreamCreate(&own));  // capture needs a non-null stream
    st = own;
  }
  struct Owned {
    std::vector<void *> p;
    hipStream_t s = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Owned() {
      if (ge) (void)hipGraphExecDestroy(ge);
      if (g) (void)hipGraphDestroy(g);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      for (void *q : p)
        if (q) (void)hipFree(q);
      if (s) (void)hipStreamDestroy(s);
    }
  } own_res;
  own_res.s = own;
  auto dalloc = [&](void **q, size_t bytes) -> mf_status {
    MF_HIP(hipMalloc(q, bytes));
    own_res.p.push_back(*q);
    return MF_OK;
  };

  // per-system levels, schedules and descriptors
  std::vector<BatchSys> hs(nsys);
  const int nwv = opt.batch_waves == 0 ? kBatchWavesDefault : opt.batch_waves;
  if (nwv != 4 && nwv != 8 && nwv != 16)
    return fail(MF_ERR_ARG, "mf_batch_cg_solve: batch_waves must be 0, 4, 8 or 16");
  const int block = 64 * nwv;
  std::vector<int> htr((size_t)nsys * (nwv + 1));
  int npad_max = 16;
  for (int i = 0; i < nsys; ++i) {
    mf_tiled *t = systems[i];
    if (opt.policy == MF_POLICY_FIXED)
      MF_HIP(hipMemcpyAsync(t->dev.prec, t->dev.prec_user, t->tilenum, hipMemcpyDeviceToDevice, st));
    if (opt.policy == MF_POLICY_ERRBOUND) MF_TRY(apply_errbound_levels(t, st));
    const std::vector<int> w = balance_tilerows(t, nwv);
    std::copy(w.begin(), w.end(), htr.begin() + (size_t)i * (nwv + 1));
    npad_max = std::max(npad_max, t->dev.npad);
  }
  int *d_tr = nullptr;
  MF_TRY(dalloc((void **)&d_tr, sizeof(int) * htr.size()));
  MF_HIP(hipMemcpyAsync(d_tr, htr.data(), sizeof(int) * htr.size(), hipMemcpyHostToDevice, st));
  if (opt.policy == MF_POLICY_ERRBOUND)  // one level group per window
    for (int i = 0; i < nsys; ++i) MF_TRY(wave_uniform_levels(systems[i], d_tr + (size_t)i * (nwv + 1), nwv, st));
  for (int i = 0; i < nsys; ++i) {
    hs[i].A = view_of(systems[i]);
    hs[i].wave_tr = d_tr + (size_t)i * (nwv + 1);
    hs[i].b = d_b[i];
    hs[i].x = d_x[i];
    hs[i].npad = systems[i]->dev.npad;
    hs[i].mixed = opt.policy == MF_POLICY_FP64 ? 0 : 1;
  }
  // systems over the queues: largest first, each to the least loaded queue
  // (work = entries; every system has nrhs items); items queue by queue
  std::vector<int> order(nsys);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int u, int v) { return systems[u]->csrsize > systems[v]->csrsize; });
  std::vector<std::vector<int>> qsys(kQueues);
  std::vector<double> qload(kQueues, 0.0);
  for (int s : order) {
    const int q = (int)(std::min_element(qload.begin(), qload.end()) - qload.begin());
    qsys[q].push_back(s);
    qload[q] += (double)systems[s]->csrsize + 1.0;
  }
  const int nitems = nsys * nrhs;
  std::vector<int2> hitems;
  std::vector<int> hqoff(kQueues + 1, 0);
  hitems.reserve(nitems);
  for (int q = 0; q < kQueues; ++q) {
    for (int s : qsys[q])
      for (int j = 0; j < nrhs; ++j) hitems.push_back(make_int2(s, j));
    hqoff[q + 1] = (int)hitems.size();
  }

  int ncu = 0, nb = 0;
  MF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const void *kern = nwv == 4 ? (const void *)k_batch_cg<256>
                   : nwv == 8 ? (const void *)k_batch_cg<512> : (const void *)k_batch_cg<1024>;
  MF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, block, 0));
  if (nb < 1) nb = 1;
  int G = std::min(nitems, ncu * nb);
  if (opt.blocks > 0) G = std::min(G, opt.blocks);
  // p32 needs npad_max floats: round the per-block stride up to whole doubles
  const long long stride = 3LL * npad_max + (npad_max + 1) / 2 + 32;

  BatchSys *d_sys = nullptr;
  int2 *d_items = nullptr;
  int *d_head = nullptr, *d_iters = nullptr, *d_qoff = nullptr;
  double *d_rel = nullptr, *d_work = nullptr;
  unsigned long long *d_ts = nullptr;
  MF_TRY(dalloc((void **)&d_sys, sizeof(BatchSys) * nsys));
  MF_TRY(dalloc((void **)&d_items, sizeof(int2) * nitems));
  MF_TRY(dalloc((void **)&d_head, sizeof(int) * kQueues));
  MF_TRY(dalloc((void **)&d_qoff, sizeof(int) * (kQueues + 1)));
  MF_TRY(dalloc((void **)&d_iters, sizeof(int) * nitems));
  MF_TRY(dalloc((void **)&d_rel, sizeof(double) * nitems));
  MF_TRY(dalloc((void **)&d_ts, sizeof(unsigned long long) * 2 * nitems));
  MF_TRY(dalloc((void **)&d_work, sizeof(double) * stride * G));
  MF_HIP(hipMemcpyAsync(d_sys, hs.data(), sizeof(BatchSys) * nsys, hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(d_items, hitems.data(), sizeof(int2) * nitems, hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(d_qoff, hqoff.data(), sizeof(int) * (kQueues + 1), hipMemcpyHostToDevice, st));
  MF_HIP(hipStreamSynchronize(st));

  BatchArgs args;
  args.sys = d_sys;
  args.items = d_items;
  args.qoff = d_qoff;
  args.nitems = nitems;
  args.head = d_head;
  args.work = d_work;
  args.stride = stride;
  args.iters = d_iters;
  args.relres = d_rel;
  args.tstamp = d_ts;
  args.tol2 = opt.tol * opt.tol;
  args.maxiter = opt.maxiter;
  args.ratio2 = opt.refresh_ratio * opt.refresh_ratio;
  args.max_refresh = opt.max_refresh;
  args.refresh_every = opt.refresh_every;

  // capture: queue reset (a kernel, not a memset node) + the persistent solve
  MF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  k_batch_reset<<<(std::max(nitems, kQueues) + 255) / 256, 256, 0, st>>>(d_head, nitems, d_iters, d_rel);
  if (nwv == 4) k_batch_cg<256><<<G, 256, 0, st>>>(args);
  else if (nwv == 8) k_batch_cg<512><<<G, 512, 0, st>>>(args);
  else k_batch_cg<1024><<<G, 1024, 0, st>>>(args);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(st, &g);
  own_res.g = g;
  if (ce != hipSuccess) return fail(MF_ERR_HIP, std::string("mf_batch_cg_solve capture: ") + hipGetErrorString(ce));
  MF_HIP(hipGraphInstantiate(&own_res.ge, g, nullptr, nullptr, 0));
  MF_HIP(hipEventCreate(&own_res.e0));
  MF_HIP(hipEventCreate(&own_res.e1));
  MF_HIP(hipEventRecord(own_res.e0, st));
  MF_HIP(hipGraphLaunch(own_res.ge, st));
  MF_HIP(hipEventRecord(own_res.e1, st));
  MF_HIP(hipEventSynchronize(own_res.e1));
  float ms = 0.0f;
  MF_HIP(hipEventElapsedTime(&ms, own_res.e0, own_res.e1));
  if (total_ms) *total_ms = ms;

  std::vector<int> hi(nitems);
  std::vector<double> hr(nitems);
  std::vector<unsigned long long> ht(2 * (size_t)nitems);
  MF_HIP(hipMemcpy(hi.data(), d_iters, sizeof(int) * nitems, hipMemcpyDeviceToHost));
  MF_HIP(hipMemcpy(hr.data(), d_rel, sizeof(double) * nitems, hipMemcpyDeviceToHost));
  MF_HIP(hipMemcpy(ht.data(), d_ts, sizeof(unsigned long long) * 2 * nitems, hipMemcpyDeviceToHost));
  bool all = true;
  for (int k = 0; k < nitems; ++k) {
    const int s = hitems[k].x, j = hitems[k].y;
    if (iterations) iterations[(size_t)s * nrhs + j] = hi[k];
    if (rel_residual) rel_residual[(size_t)s * nrhs + j] = hr[k];
    if (!(hr[k] >= 0.0 && hr[k] <= opt.tol)) all = false;
  }
  if (per_system_ms) {
    for (int s = 0; s < nsys; ++s) {
      unsigned long long t0 = ~0ull, t1 = 0;
      for (int k = 0; k < nitems; ++k)
        if (hitems[k].x == s) {
          t0 = std::min(t0, ht[2 * (size_t)k]);
          t1 = std::max(t1, ht[2 * (size_t)k + 1]);
        }
      per_system_ms[s] = t1 > t0 ? (double)(t1 - t0) * 1e-5 : 0.0;  // 100 MHz ticks
    }
  }
  if (!all) return fail(MF_NOT_CONVERGED, "mf_batch_cg_solve: some systems did not reach the tolerance");
  return MF_OK;
}
