// mf_tune.hip -- autotuner: candidate schedules of the stand-alone SpMV and
// of the batched CG are timed on the device and the fastest is kept.
//
// SpMV: the streamed kernel's entries per wave set the number and length of
// the per-wave windows (more waves hide more latency, longer waves pay their
// setup less often); the row-group kernel is a candidate where tile-rows are
// regular.  Batched CG: the waves per (system, right-hand side) item trade
// per-item parallelism against the number of items in flight.
#include <algorithm>
#include <vector>

#include "mf_device.h"

namespace mf {

namespace {

// Median of n timed launches of f (HIP events on the null stream) after w
// warm-up launches, in microseconds.
template <typename F>
mf_status time_median(F f, int w, int n, float *us) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  MF_HIP(hipEventCreate(&e0));
  hipError_t he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(MF_ERR_HIP, std::string("hipEventCreate: ") + hipGetErrorString(he));
  }
  std::vector<float> ts;
  mf_status st = MF_OK;
  for (int i = 0; i < w + n && st == MF_OK; ++i) {
    if (hipEventRecord(e0, nullptr) != hipSuccess) st = fail(MF_ERR_HIP, "hipEventRecord");
    if (st == MF_OK) st = f();
    if (st == MF_OK && hipEventRecord(e1, nullptr) != hipSuccess) st = fail(MF_ERR_HIP, "hipEventRecord");
    if (st == MF_OK && hipEventSynchronize(e1) != hipSuccess) st = fail(MF_ERR_HIP, "hipEventSynchronize");
    float ms = 0.0f;
    if (st == MF_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = fail(MF_ERR_HIP, "hipEventElapsedTime");
    if (st == MF_OK && i >= w) ts.push_back(ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != MF_OK) return st;
  std::sort(ts.begin(), ts.end());
  *us = 1e3f * ts[ts.size() / 2];
  return MF_OK;
}

struct DevBuf {
  std::vector<void *> p;
  ~DevBuf() {
    for (void *q : p)
      if (q) (void)hipFree(q);
  }
  mf_status alloc(void **q, size_t bytes) {
    MF_HIP(hipMalloc(q, bytes));
    p.push_back(*q);
    return MF_OK;
  }
};

}  // namespace

}  // namespace mf

using namespace mf;

extern "C" mf_status mf_tune_spmv(mf_tiled *t, int mode, mf_tune_report *rep) {
  if (!t || !rep) return fail(MF_ERR_ARG, "mf_tune_spmv: null argument");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, "mf_tune_spmv: host-only handle");
  if (mode != kAll64 && mode != kPerTile) return fail(MF_ERR_ARG, "mf_tune_spmv: bad mode");
  MF_HIP(hipSetDevice(t->dev.device));
  const int np = t->dev.npad;
  DevBuf buf;
  double *x = nullptr, *y = nullptr;
  float *x32 = nullptr;
  MF_TRY(buf.alloc((void **)&x, sizeof(double) * np));
  MF_TRY(buf.alloc((void **)&y, sizeof(double) * np));
  MF_TRY(buf.alloc((void **)&x32, sizeof(float) * np));
  MF_HIP(hipMemset(x, 0, sizeof(double) * np));
  MF_HIP(hipMemset(x32, 0, sizeof(float) * np));
  const int old_nnz = t->dev.spmv_wave_nnz, old_kernel = t->dev.spmv_kernel;
  std::vector<int> cand = {384, 512, 768, 1024, 1536, 2048};
  if (regular_tilerows(t)) cand.push_back(0);
  *rep = mf_tune_report{};
  rep->ncand = (int)cand.size();
  float best = 0.0f;
  int bi = -1;
  for (size_t i = 0; i < cand.size(); ++i) {
    if (cand[i] > 0) {
      MF_TRY(mf_tile_set_schedule(t, cand[i]));
      t->dev.spmv_kernel = MF_SPMV_STREAM;
    } else {
      t->dev.spmv_kernel = MF_SPMV_ROWGROUP;
    }
    float us = 0.0f;
    const mf_status st = time_median([&] { return mf_spmv(t, x, x32, y, mode, nullptr); }, 5, 25, &us);
    if (st != MF_OK) {
      t->dev.spmv_kernel = old_kernel;
      (void)mf_tile_set_schedule(t, old_nnz);
      return st;
    }
    rep->cand[i] = cand[i];
    rep->us[i] = us;
    if (bi < 0 || us < best) {
      best = us;
      bi = (int)i;
    }
  }
  rep->best = cand[bi];
  if (cand[bi] > 0) {
    MF_TRY(mf_tile_set_schedule(t, cand[bi]));
    t->dev.spmv_kernel = MF_SPMV_STREAM;
  } else {
    MF_TRY(mf_tile_set_schedule(t, old_nnz));
    t->dev.spmv_kernel = MF_SPMV_ROWGROUP;
  }
  return MF_OK;
}

extern "C" mf_status mf_tune_batch(mf_tiled **systems, int nsys, int nrhs, const double **d_b,
                                   double **d_x, const mf_solve_opts *opts, int iters,
                                   mf_tune_report *rep) {
  if (!systems || nsys <= 0 || nrhs <= 0 || !d_b || !d_x || !rep)
    return fail(MF_ERR_ARG, "mf_tune_batch: bad argument");
  if (iters <= 0) return fail(MF_ERR_ARG, "mf_tune_batch: iters must be positive");
  for (int i = 0; i < nsys; ++i)
    if (!systems[i] || systems[i]->dev.device < 0 || !d_x[i])
      return fail(MF_ERR_ARG, "mf_tune_batch: systems must be device handles");
  MF_HIP(hipSetDevice(systems[0]->dev.device));
  // copies of the initial guesses: the timed runs overwrite them
  DevBuf buf;
  std::vector<double *> xs(nsys), x0(nsys);
  for (int i = 0; i < nsys; ++i) {
    const size_t bytes = sizeof(double) * (size_t)systems[i]->dev.npad * nrhs;
    MF_TRY(buf.alloc((void **)&x0[i], bytes));
    MF_TRY(buf.alloc((void **)&xs[i], bytes));
    MF_HIP(hipMemcpy(x0[i], d_x[i], bytes, hipMemcpyDeviceToDevice));
  }
  mf_solve_opts o;
  mf_solve_opts_default(&o);
  if (opts) o = *opts;
  o.maxiter = iters;
  const int cand[3] = {4, 8, 16};
  *rep = mf_tune_report{};
  rep->ncand = 3;
  float best = 0.0f;
  int bi = -1;
  for (int c = 0; c < 3; ++c) {
    o.batch_waves = cand[c];
    std::vector<float> ts;
    for (int r = 0; r < 3; ++r) {
      for (int i = 0; i < nsys; ++i)
        MF_HIP(hipMemcpy(xs[i], x0[i], sizeof(double) * (size_t)systems[i]->dev.npad * nrhs,
                         hipMemcpyDeviceToDevice));
      double ms = 0.0;
      const mf_status st = mf_batch_cg_solve(systems, nsys, nrhs, d_b, xs.data(), &o, nullptr,
                                             nullptr, nullptr, &ms);
      if (st != MF_OK && st != MF_NOT_CONVERGED) return st;
      ts.push_back((float)(ms * 1e3));
    }
    std::sort(ts.begin(), ts.end());
    rep->cand[c] = cand[c];
    rep->us[c] = ts[1];
    if (bi < 0 || ts[1] < best) {
      best = ts[1];
      bi = c;
    }
  }
  rep->best = cand[bi];
  set_error("");
  return MF_OK;
}
