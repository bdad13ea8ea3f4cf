// mf_multikernel.hip -- multi-kernel CG and BiCGSTAB: one launch per SpMV,
// dot product and vector update, with the scalars formed on the host from
// values read back every iteration.  This is the launch-bound baseline the
// single-kernel solvers are compared with (the reference's cg_solve_reduce,
// src/main-cg.cu:245-279, and multi-kernel bicgstab, src/main-bicg.cu:1619-1656,
// which run a fixed 10000 iterations; these stop at the FP64 tolerance).
#include <cmath>
#include <cstring>
#include <vector>

#include "mf_device.h"

namespace mf {

mf_status launch_spmv(const mf_tiled *t, const double *d_x, const float *d_x32, double *d_y,
                      int mode, hipStream_t s);

namespace {

constexpr int kDotBlocks = 1024;

// Deterministic two-stage dot products: block partials in a fixed order,
// then one block adds the partials in block order.
__global__ __launch_bounds__(256) void k_dot_partial(const double *__restrict__ a,
                                                     const double *__restrict__ b, int n,
                                                     double *__restrict__ part) {
  __shared__ double lds[4];
  double s = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += a[i] * b[i];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_dot_final(const double *__restrict__ part, int nb,
                                                   double *__restrict__ out) {
  __shared__ double lds[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) *out = s;
}

// y = a*x + b*y (elementwise), z optional: z = y - c*w
__global__ void k_axpby(int n, double a, const double *__restrict__ x, double b,
                        double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

// z = x + a*y
__global__ void k_xpay_to(int n, const double *__restrict__ x, double a,
                          const double *__restrict__ y, double *__restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = x[i] + a * y[i];
}

// p = r + beta (p - omega v)
__global__ void k_bicg_p(int n, const double *__restrict__ r, double beta, double omega,
                         const double *__restrict__ v, double *__restrict__ p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

// x += alpha p + omega s
__global__ void k_bicg_x(int n, double alpha, const double *__restrict__ p, double omega,
                         const double *__restrict__ s, double *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += alpha * p[i] + omega * s[i];
}

struct MK {
  int n;
  hipStream_t s;
  double *part, *scal;
  mf_status dot(const double *a, const double *b, double *out_host) {
    k_dot_partial<<<kDotBlocks, 256, 0, s>>>(a, b, n, part);
    k_dot_final<<<1, 256, 0, s>>>(part, kDotBlocks, scal);
    MF_HIP(hipGetLastError());
    MF_HIP(hipMemcpyAsync(out_host, scal, sizeof(double), hipMemcpyDeviceToHost, s));
    MF_HIP(hipStreamSynchronize(s));
    return MF_OK;
  }
  int grid() const { return (n + 255) / 256; }
};

struct Buffers {
  std::vector<void *> ptrs;
  ~Buffers() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  mf_status get(T **p, size_t n) {
    MF_HIP(hipMalloc((void **)p, n * sizeof(T)));
    ptrs.push_back(*p);
    MF_HIP(hipMemset(*p, 0, n * sizeof(T)));
    return MF_OK;
  }
};


void fill_report(mf_report *rep, int it, double rr, double tr, double bb, double ms, int status) {
  if (!rep) return;
  std::memset(rep, 0, sizeof(*rep));
  rep->iterations = it;
  rep->rel_residual_recursive = bb > 0 ? std::sqrt(rr / bb) : 0.0;
  rep->rel_residual = bb > 0 ? std::sqrt(tr / bb) : std::sqrt(tr);
  rep->time_ms = ms;
  rep->status = status;
}

mf_status true_res(const mf_tiled *t, MK &k, const double *b, const double *x, double *tmp,
                   double *tr) {
  MF_TRY(launch_spmv(t, x, nullptr, tmp, kAll64, k.s));
  k_axpby<<<k.grid(), 256, 0, k.s>>>(k.n, 1.0, b, -1.0, tmp);  // tmp = b - A x
  MF_HIP(hipGetLastError());
  return k.dot(tmp, tmp, tr);
}

}  // namespace
}  // namespace mf

using namespace mf;

extern "C" mf_status mf_cg_solve_multikernel(mf_tiled *t, const double *d_b, double *d_x,
                                             const mf_solve_opts *o, mf_report *rep) {
  if (!t || !d_b || !d_x) return fail(MF_ERR_ARG, "mf_cg_solve_multikernel: null argument");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, "solve: host-only handle");
  mf_solve_opts opt;
  mf_solve_opts_default(&opt);
  if (o) opt = *o;
  if (opt.maxiter <= 0) opt.maxiter = 10000;
  MF_HIP(hipSetDevice(t->dev.device));
  MK k{t->dev.npad, (hipStream_t)opt.stream, nullptr, nullptr};
  Buffers B;
  double *r, *p, *q;
  MF_TRY(B.get(&r, k.n));
  MF_TRY(B.get(&p, k.n));
  MF_TRY(B.get(&q, k.n));
  MF_TRY(B.get(&k.part, kDotBlocks));
  MF_TRY(B.get(&k.scal, 1));
  hipEvent_t e0, e1;
  MF_HIP(hipEventCreate(&e0));
  MF_HIP(hipEventCreate(&e1));
  MF_HIP(hipEventRecord(e0, k.s));
  // r = b - A x, p = r
  MF_TRY(launch_spmv(t, d_x, nullptr, q, kAll64, k.s));
  k_xpay_to<<<k.grid(), 256, 0, k.s>>>(k.n, d_b, -1.0, q, r);
  MF_HIP(hipMemcpyAsync(p, r, sizeof(double) * k.n, hipMemcpyDeviceToDevice, k.s));
  double rr = 0, bb = 0, pq = 0, tr = 0;
  MF_TRY(k.dot(r, r, &rr));
  MF_TRY(k.dot(d_b, d_b, &bb));
  const double stop2 = opt.tol * opt.tol * bb;
  int it = 0, status = MF_OK;
  bool done = bb == 0.0 || (opt.fixed_iters <= 0 && rr <= stop2);
  while (!done) {
    ++it;
    MF_TRY(launch_spmv(t, p, nullptr, q, kAll64, k.s));
    MF_TRY(k.dot(p, q, &pq));
    if (!(pq != 0.0) || !std::isfinite(pq)) { status = MF_ERR_BREAKDOWN; break; }
    const double alpha = rr / pq;
    k_axpby<<<k.grid(), 256, 0, k.s>>>(k.n, alpha, p, 1.0, d_x);
    k_axpby<<<k.grid(), 256, 0, k.s>>>(k.n, -alpha, q, 1.0, r);
    MF_HIP(hipGetLastError());
    double rn = 0;
    MF_TRY(k.dot(r, r, &rn));
    if (opt.fixed_iters > 0) {
      if (it >= opt.fixed_iters) { rr = rn; break; }
    } else {
      if (rn <= stop2) {
        MF_TRY(true_res(t, k, d_b, d_x, q, &tr));
        if (tr <= stop2) { rr = rn; break; }
        MF_HIP(hipMemcpyAsync(r, q, sizeof(double) * k.n, hipMemcpyDeviceToDevice, k.s));
        rn = tr;
      }
      if (it >= opt.maxiter) { rr = rn; status = MF_NOT_CONVERGED; break; }
    }
    const double beta = rn / rr;
    rr = rn;
    k_axpby<<<k.grid(), 256, 0, k.s>>>(k.n, 1.0, r, beta, p);  // p = r + beta p
    MF_HIP(hipGetLastError());
  }
  MF_HIP(hipEventRecord(e1, k.s));
  MF_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MF_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  MF_TRY(true_res(t, k, d_b, d_x, q, &tr));
  int ret = status;
  if (ret == MF_OK && opt.fixed_iters <= 0 && !(bb == 0.0 || std::sqrt(tr / bb) <= opt.tol)) ret = MF_NOT_CONVERGED;
  fill_report(rep, it, rr, tr, bb, ms, ret);
  if (ret != MF_OK) set_error("mf_cg_solve_multikernel: not converged or breakdown");
  return ret;
}

extern "C" mf_status mf_bicgstab_solve_multikernel(mf_tiled *t, const double *d_b, double *d_x,
                                                   const mf_solve_opts *o, mf_report *rep) {
  if (!t || !d_b || !d_x) return fail(MF_ERR_ARG, "mf_bicgstab_solve_multikernel: null argument");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, "solve: host-only handle");
  mf_solve_opts opt;
  mf_solve_opts_default(&opt);
  if (o) opt = *o;
  if (opt.maxiter <= 0) opt.maxiter = 10000;
  MF_HIP(hipSetDevice(t->dev.device));
  MK k{t->dev.npad, (hipStream_t)opt.stream, nullptr, nullptr};
  Buffers B;
  double *r, *rh, *p, *v, *s, *tt;
  MF_TRY(B.get(&r, k.n));
  MF_TRY(B.get(&rh, k.n));
  MF_TRY(B.get(&p, k.n));
  MF_TRY(B.get(&v, k.n));
  MF_TRY(B.get(&s, k.n));
  MF_TRY(B.get(&tt, k.n));
  MF_TRY(B.get(&k.part, kDotBlocks));
  MF_TRY(B.get(&k.scal, 1));
  hipEvent_t e0, e1;
  MF_HIP(hipEventCreate(&e0));
  MF_HIP(hipEventCreate(&e1));
  MF_HIP(hipEventRecord(e0, k.s));
  MF_TRY(launch_spmv(t, d_x, nullptr, v, kAll64, k.s));
  k_xpay_to<<<k.grid(), 256, 0, k.s>>>(k.n, d_b, -1.0, v, r);
  MF_HIP(hipMemcpyAsync(rh, r, sizeof(double) * k.n, hipMemcpyDeviceToDevice, k.s));
  MF_HIP(hipMemcpyAsync(p, r, sizeof(double) * k.n, hipMemcpyDeviceToDevice, k.s));
  double rho = 0, bb = 0, rr = 0, tr = 0;
  MF_TRY(k.dot(r, rh, &rho));
  MF_TRY(k.dot(d_b, d_b, &bb));
  rr = rho;
  const double stop2 = opt.tol * opt.tol * bb;
  int it = 0, status = MF_OK;
  bool done = bb == 0.0 || (opt.fixed_iters <= 0 && rr <= stop2);
  while (!done) {
    ++it;
    MF_TRY(launch_spmv(t, p, nullptr, v, kAll64, k.s));
    double vr = 0;
    MF_TRY(k.dot(v, rh, &vr));
    if (!(vr != 0.0) || !std::isfinite(vr)) { status = MF_ERR_BREAKDOWN; break; }
    const double alpha = rho / vr;
    k_xpay_to<<<k.grid(), 256, 0, k.s>>>(k.n, r, -alpha, v, s);
    MF_TRY(launch_spmv(t, s, nullptr, tt, kAll64, k.s));
    double ts = 0, t2 = 0;
    MF_TRY(k.dot(tt, s, &ts));
    MF_TRY(k.dot(tt, tt, &t2));
    const double omega = t2 != 0.0 ? ts / t2 : 0.0;
    k_bicg_x<<<k.grid(), 256, 0, k.s>>>(k.n, alpha, p, omega, s, d_x);
    k_xpay_to<<<k.grid(), 256, 0, k.s>>>(k.n, s, -omega, tt, r);
    MF_HIP(hipGetLastError());
    double rho_new = 0, rn = 0;
    MF_TRY(k.dot(r, rh, &rho_new));
    MF_TRY(k.dot(r, r, &rn));
    if (opt.fixed_iters > 0) {
      if (it >= opt.fixed_iters) { rr = rn; break; }
    } else {
      if (rn <= stop2 || omega == 0.0) {
        MF_TRY(true_res(t, k, d_b, d_x, v, &tr));
        if (tr <= stop2) { rr = rn; break; }
        MF_HIP(hipMemcpyAsync(r, v, sizeof(double) * k.n, hipMemcpyDeviceToDevice, k.s));
        rn = tr;
        MF_TRY(k.dot(r, rh, &rho_new));
      }
      if (omega == 0.0 || rho == 0.0) { rr = rn; status = MF_ERR_BREAKDOWN; break; }
      if (it >= opt.maxiter) { rr = rn; status = MF_NOT_CONVERGED; break; }
    }
    const double beta = (rho_new / rho) * (alpha / omega);
    rho = rho_new;
    rr = rn;
    k_bicg_p<<<k.grid(), 256, 0, k.s>>>(k.n, r, beta, omega, v, p);
    MF_HIP(hipGetLastError());
  }
  MF_HIP(hipEventRecord(e1, k.s));
  MF_HIP(hipEventSynchronize(e1));
  float ms = 0;
  MF_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  MF_TRY(true_res(t, k, d_b, d_x, v, &tr));
  int ret = status;
  if (ret == MF_OK && opt.fixed_iters <= 0 && !(bb == 0.0 || std::sqrt(tr / bb) <= opt.tol)) ret = MF_NOT_CONVERGED;
  fill_report(rep, it, rr, tr, bb, ms, ret);
  if (ret != MF_OK) set_error("mf_bicgstab_solve_multikernel: not converged or breakdown");
  return ret;
}
