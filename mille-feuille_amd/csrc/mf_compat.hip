// mf_compat.hip -- the reference's solver entry points with their exact
// argument lists (src/main-cg.cu:6, 379, 1071; src/main-cg-mixed.cu:2238;
// src/main-bicg.cu:13, 711, 1336; src/main-bicg-mixed.cu:1355), implemented on
// the single-kernel solvers.
//
// Behaviour kept from the reference: rows truncated to a multiple of 16 with
// colA = ori (src/main-cg.cu:9-11), x set to zero before solving
// (src/main-cg.cu:29), the FP64 CPU check ||A x - b|| / ||b|| over the tile
// arrays (src/main-cg.cu:978-1011), the "time_cg=... ms" line and one CSV row
// appended to data/cg_performance.csv or data/bicg_performance.csv
// (src/main-cg.cu:1012-1022, src/main-bicg.cu:676-688; a missing data/ folder
// prints "open error!" and returns, as there).  For BiCGSTAB, rhs is
// overwritten with the row sums of A (src/main-bicg.cu:480-487).
// CSV fields 4 and 5 are the reference's: for CG the FP64 check and the
// absolute recursive residual norm sqrt(r.r) (src/main-cg.cu:1013); for
// BiCGSTAB the recursive ||r|| / ||rhs|| and ||r||^2 / ||r0||
// (src/main-bicg.cu:675-678, where k_residual accumulates r.r and err_rel is
// ||r0||).
//
// The same names mean FP64 solvers in main-cg / main-bicg and mixed ones in
// main-cg-mixed / main-bicg-mixed (src/main-cg-mixed.cu:2238,
// src/main-bicg-mixed.cu:1355).  This file is therefore built twice: into
// libmillefeuille_hip.so, whose entry points default to FP64 (cg_solve_reduce
// and bicgstab to the multi-kernel solvers, as src/main-cg.cu:6 and
// src/main-bicg.cu:1336), and with MF_COMPAT_MIXED into
// libmillefeuille_mixed.so, whose entry points all default to the
// single-kernel solvers under the reference's rule policy.
// Differences: the solve stops at the FP64 relative residual 1e-10 instead of
// running a fixed 100 (CG) / 50 (BiCGSTAB) iterations; MF_COMPAT_FIXED_ITERS=k
// in the environment restores a fixed count.  MF_COMPAT_POLICY selects the
// policy of every entry point (fp64, rule, errbound, fixed, learned); a policy
// other than fp64 always runs the single-kernel solver.  The *_mixed entry
// points default to the rule policy in both builds.  The learned policy loads
// its MLP from MF_POLICY_FILE, else from policy_mlp.f32 next to this library.
// One line per solve on stdout names the solver that ran: "mf: cg
// single-kernel policy=rule iterations=N levels=a/b/c/d/e", the last field
// being the tile-SpMV counts at FP64/FP32/FP16/INT8/skipped.
#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mf_device.h"

namespace {

#ifdef MF_COMPAT_MIXED
constexpr int kDefaultPolicy = MF_POLICY_RULE;
#else
constexpr int kDefaultPolicy = MF_POLICY_FP64;
#endif

int env_int(const char *name, int dflt) {
  const char *v = std::getenv(name);
  return v && *v ? std::atoi(v) : dflt;
}

int env_policy(int dflt) {
  const char *v = std::getenv("MF_COMPAT_POLICY");
  if (!v || !*v) return dflt;
  std::string s(v);
  if (s == "fp64") return MF_POLICY_FP64;
  if (s == "rule") return MF_POLICY_RULE;
  if (s == "fixed") return MF_POLICY_FIXED;
  if (s == "errbound") return MF_POLICY_ERRBOUND;
  if (s == "learned") return MF_POLICY_LEARNED;
  return dflt;
}

const char *policy_name(int p) {
  switch (p) {
    case MF_POLICY_FP64: return "fp64";
    case MF_POLICY_RULE: return "rule";
    case MF_POLICY_FIXED: return "fixed";
    case MF_POLICY_ERRBOUND: return "errbound";
    case MF_POLICY_LEARNED: return "learned";
    default: return "?";
  }
}

void print_solver(const char *method, bool multikernel, const mf_report &rep) {
  std::printf("mf: %s %s policy=%s iterations=%d levels=%lld/%lld/%lld/%lld/%lld\n", method,
              multikernel ? "multi-kernel" : "single-kernel", policy_name(rep.policy_used),
              rep.iterations, rep.prec_hist[0], rep.prec_hist[1], rep.prec_hist[2], rep.prec_hist[3],
              rep.prec_hist[4]);
}

double norm2(const double *v, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += v[i] * v[i];
  return std::sqrt(s);
}

// Installs the shipped MLP (1476 little-endian float32) on the handle;
// without it the learned policy runs the rule policy.
void install_shipped_policy(mf_tiled *t) {
  std::string path;
  if (const char *e = std::getenv("MF_POLICY_FILE")) path = e;
  if (path.empty()) {
    Dl_info info;
    if (dladdr((void *)&install_shipped_policy, &info) && info.dli_fname) {
      path = info.dli_fname;
      const size_t slash = path.rfind('/');
      path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/policy_mlp.f32";
    }
  }
  std::vector<float> p(MF_MLP_NPARAM);
  FILE *f = path.empty() ? nullptr : std::fopen(path.c_str(), "rb");
  const bool ok = f && std::fread(p.data(), sizeof(float), p.size(), f) == p.size();
  if (f) std::fclose(f);
  if (!ok) {
    std::printf("mf: no learned policy at %s; using the rule policy\n", path.c_str());
    return;
  }
  float *d = nullptr;
  if (hipMalloc(&d, sizeof(float) * p.size()) != hipSuccess) return;
  if (hipMemcpy(d, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice) == hipSuccess &&
      mf_policy_set_mlp(t, d) != MF_OK)
    std::printf("mf: %s\n", mf_last_error());
  (void)hipFree(d);
}

// FP64 check over the tile arrays, as the reference's post-solve loop.
double tile_check(const mf_tiled *t, const double *x, const double *b, int n) {
  mf_tile_arrays a;
  mf_tile_arrays_get(t, &a);
  std::vector<double> bx(n, 0.0), xp((size_t)t->tilen * 16, 0.0);
  for (int i = 0; i < n && i < (int)xp.size(); ++i) xp[i] = x[i];
  for (int bi = 0; bi < t->tilem; ++bi) {
    const int h = bi == t->tilem - 1 ? t->rowA - (t->tilem - 1) * 16 : 16;
    for (int tt = a.tile_ptr[bi]; tt < a.tile_ptr[bi + 1]; ++tt) {
      const int xo = a.tile_columnidx[tt] * 16, off = a.csr_offset[tt];
      const unsigned char *rp = a.Blockcsr_Ptr + a.csrptr_offset[tt];
      const int nzt = a.csr_offset[tt + 1] - off;
      for (int r = 0; r < h; ++r) {
        const int stop = r == h - 1 ? nzt : rp[r + 1];
        double s = 0.0;
        for (int k = rp[r]; k < stop; ++k) s += xp[xo + a.Tile_csr_Col[off + k]] * a.Blockcsr_Val[off + k];
        if (bi * 16 + r < n) bx[bi * 16 + r] += s;
      }
    }
  }
  double num = 0.0, den = 0.0;
  for (int i = 0; i < n; ++i) {
    num += (bx[i] - b[i]) * (bx[i] - b[i]);
    den += b[i] * b[i];
  }
  return den > 0 ? std::sqrt(num) / std::sqrt(den) : std::sqrt(num);
}

void append_csv(const char *path, const char *filename, const char *row) {
  FILE *f = std::fopen(path, "a");
  if (!f) {
    std::printf("open error!\n");
    return;
  }
  std::fprintf(f, "%s,%s", filename ? filename : "", row);
  std::fclose(f);
}

struct DevVec {
  double *p = nullptr;
  ~DevVec() {
    if (p) (void)hipFree(p);
  }
};

// Shared CG driver: rows [0, rowA), solve on device 0, x (n entries) out.
void compat_cg(int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x, double *b,
               int n, char *filename, int nnzR, int ori, int policy, bool multikernel) {
  multikernel = multikernel && policy == MF_POLICY_FP64;
  const int rowA = (n / 16) * 16;
  std::memset(x, 0, sizeof(double) * n);
  mf_tiled *t = nullptr;
  if (mf_tile_create(rowA, ori, RowPtr, ColIdx, Val, Val_Low, 0, &t) != MF_OK) {
    std::printf("mf: %s\n", mf_last_error());
    return;
  }
  if (policy == MF_POLICY_LEARNED) install_shipped_policy(t);
  int npad = 0;
  mf_tile_info(t, nullptr, nullptr, nullptr, nullptr, nullptr, &npad);
  DevVec db, dx;
  std::vector<double> hb(npad, 0.0), hx(npad, 0.0);
  for (int i = 0; i < rowA; ++i) hb[i] = b[i];
  if (hipMalloc(&db.p, sizeof(double) * npad) != hipSuccess ||
      hipMalloc(&dx.p, sizeof(double) * npad) != hipSuccess) {
    std::printf("mf: device allocation failed\n");
    mf_tile_free(t);
    return;
  }
  (void)hipMemcpy(db.p, hb.data(), sizeof(double) * npad, hipMemcpyHostToDevice);
  (void)hipMemcpy(dx.p, hx.data(), sizeof(double) * npad, hipMemcpyHostToDevice);
  mf_solve_opts o;
  mf_solve_opts_default(&o);
  o.policy = policy;
  o.fixed_iters = env_int("MF_COMPAT_FIXED_ITERS", 0);
  mf_report rep;
  std::memset(&rep, 0, sizeof(rep));
  mf_status st = multikernel ? mf_cg_solve_multikernel(t, db.p, dx.p, &o, &rep)
                             : mf_cg_solve(t, db.p, dx.p, &o, &rep);
  if (st < 0) std::printf("mf: %s\n", mf_last_error());
  (void)hipMemcpy(hx.data(), dx.p, sizeof(double) * npad, hipMemcpyDeviceToHost);
  for (int i = 0; i < n; ++i) x[i] = i < rowA ? hx[i] : 0.0;
  print_solver("cg", multikernel, rep);
  std::printf("time_cg=%lf ms\n", rep.time_ms);
  const double l2 = tile_check(t, x, b, rowA);
  char row[256];
  std::snprintf(row, sizeof row, "%d,%.3f,%d,%e,%e\n", rep.iterations, rep.time_ms, nnzR, l2,
                rep.rel_residual_recursive * norm2(b, rowA));
  append_csv("data/cg_performance.csv", filename, row);
  mf_tile_free(t);
}

int compat_bicg(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR,
                char *filename, int n, int policy, bool multikernel) {
  multikernel = multikernel && policy == MF_POLICY_FP64;
  const int rowA = (n / 16) * 16;
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int k = RowPtr[i]; k < RowPtr[i + 1]; ++k) s += Val[k];
    rhs[i] = s;
  }
  std::memset(x, 0, sizeof(double) * n);
  mf_tiled *t = nullptr;
  if (mf_tile_create(rowA, n, RowPtr, ColIdx, Val, nullptr, 0, &t) != MF_OK) {
    std::printf("mf: %s\n", mf_last_error());
    return -1;
  }
  if (policy == MF_POLICY_LEARNED) install_shipped_policy(t);
  int npad = 0;
  mf_tile_info(t, nullptr, nullptr, nullptr, nullptr, nullptr, &npad);
  DevVec db, dx;
  std::vector<double> hb(npad, 0.0), hx(npad, 0.0);
  for (int i = 0; i < rowA; ++i) hb[i] = rhs[i];
  if (hipMalloc(&db.p, sizeof(double) * npad) != hipSuccess ||
      hipMalloc(&dx.p, sizeof(double) * npad) != hipSuccess) {
    mf_tile_free(t);
    return -1;
  }
  (void)hipMemcpy(db.p, hb.data(), sizeof(double) * npad, hipMemcpyHostToDevice);
  (void)hipMemcpy(dx.p, hx.data(), sizeof(double) * npad, hipMemcpyHostToDevice);
  mf_solve_opts o;
  mf_solve_opts_default(&o);
  o.policy = policy;
  o.fixed_iters = env_int("MF_COMPAT_FIXED_ITERS", 0);
  mf_report rep;
  std::memset(&rep, 0, sizeof(rep));
  mf_status st = multikernel ? mf_bicgstab_solve_multikernel(t, db.p, dx.p, &o, &rep)
                             : mf_bicgstab_solve(t, db.p, dx.p, &o, &rep);
  if (st < 0) std::printf("mf: %s\n", mf_last_error());
  (void)hipMemcpy(hx.data(), dx.p, sizeof(double) * npad, hipMemcpyDeviceToHost);
  for (int i = 0; i < n; ++i) x[i] = i < rowA ? hx[i] : 0.0;
  print_solver("bicgstab", multikernel, rep);
  std::printf("time_bicg=%lf ms\n", rep.time_ms);
  // ||r|| / ||rhs|| and ||r||^2 / ||r0|| with r0 = rhs (x starts at zero)
  const double rel = rep.rel_residual_recursive, bnorm = norm2(rhs, rowA);
  char row[256];
  std::snprintf(row, sizeof row, "%d,%.3f,%d,%e,%e\n", rep.iterations, rep.time_ms, nnzR, rel,
                rel * rel * bnorm);
  append_csv("data/bicg_performance.csv", filename, row);
  mf_tile_free(t);
  return st == MF_OK ? 0 : 1;
}

}  // namespace

extern "C" void cg_solve_reduce(int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x,
                                double *b, int n, int *iter, int maxiter, double threshold,
                                char *filename, int nnzR, int ori) {
  (void)iter; (void)maxiter; (void)threshold;
  compat_cg(RowPtr, ColIdx, Val, Val_Low, x, b, n, filename, nnzR, ori, env_policy(kDefaultPolicy), true);
}

extern "C" void cg_solve_sync(int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x,
                              double *b, int n, int *iter, int maxiter, double threshold,
                              char *filename, int nnzR, int ori) {
  (void)iter; (void)maxiter; (void)threshold;
  compat_cg(RowPtr, ColIdx, Val, Val_Low, x, b, n, filename, nnzR, ori, env_policy(kDefaultPolicy), false);
}

extern "C" void cg_solve_inc(int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x,
                             double *b, int n, int *iter, int maxiter, double threshold,
                             char *filename, int nnzR, int ori) {
  (void)iter; (void)maxiter; (void)threshold;
  compat_cg(RowPtr, ColIdx, Val, Val_Low, x, b, n, filename, nnzR, ori, env_policy(kDefaultPolicy), false);
}

extern "C" void cg_solve_mixed(int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x,
                               double *b, int n, int *iter, int maxiter, double threshold,
                               char *filename, int nnzR, int ori) {
  (void)iter; (void)maxiter; (void)threshold;
  compat_cg(RowPtr, ColIdx, Val, Val_Low, x, b, n, filename, nnzR, ori, env_policy(MF_POLICY_RULE), false);
}

extern "C" int bicgstab(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR,
                        char *filename, int n) {
  return compat_bicg(RowPtr, ColIdx, Val, rhs, x, nnzR, filename, n, env_policy(kDefaultPolicy), true);
}

extern "C" int bicgstab_sync(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x,
                             int nnzR, char *filename, int n) {
  return compat_bicg(RowPtr, ColIdx, Val, rhs, x, nnzR, filename, n, env_policy(kDefaultPolicy), false);
}

extern "C" int bicgstab_inc(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x,
                            int nnzR, char *filename, int n) {
  return compat_bicg(RowPtr, ColIdx, Val, rhs, x, nnzR, filename, n, env_policy(kDefaultPolicy), false);
}

extern "C" int bicgstab_mixed(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x,
                              int nnzR, char *filename, int n) {
  return compat_bicg(RowPtr, ColIdx, Val, rhs, x, nnzR, filename, n, env_policy(MF_POLICY_RULE), false);
}
