// mf_learned.hip -- the learned tile-precision policy: per-tile features,
// labels from the solver's own iterates, and the MLP's forward, backward and
// Adam kernels.
//
// Features of tile t (8 floats, all O(1)):
//   0  nnz_t / 256
//   1  value range: log2(max|a| / min|a|) / 24 over the tile's nonzeros
//   2  FP16 payload error: log2(qerr16 / max|a| + 2^-40) / 40
//   3  INT8 payload error: log2(qerr8 / max|a| + 2^-40) / 40
//   4  weight of the tile in its tile-row: sum|a_t| / sum|a_tilerow|
//   5  diagonal dominance of the tile-row: log2(min_i |a_ii| / sum_j!=i |a_ij|) / 8,
//      clamped to [-1, 1]
//   6  local residual norm: log2(||v_seg|| sqrt(nseg) / ||v|| + 2^-40) / 40 for
//      the 16-entry segment of the multiplied vector v the tile reads
//   7  iteration progress: log10(||r|| / ||b||) / 16
// Features 0-5 depend on the matrix only and are built once per handle.
//
// Label of tile t for a vector v: the cheapest level L in INT8, FP16, FP32
// whose product error on v, e_t(L) = ||(deq_L(A_t) fp32(v) - A_t v)_rows||_2
// computed exactly as the SpMV computes it, is at most
// theta * ||A v||_2 / sqrt(tilenum); FP64 otherwise.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mf_device.h"
#include "mf_mlp.h"

using namespace mf;

namespace mf {

constexpr int kGS = 128;                 // samples per gradient block (= threads)
constexpr int kGSP = kGS + 1;            // padded LDS row (no bank conflicts)
constexpr int kPartial = kMlpParams + 1; // gradient partials + loss per block

// ---- parameter packing into the MFMA lane layout ---------------------------
__device__ __forceinline__ _Float16 h16(float v) { return (_Float16)v; }

__global__ void k_mlp_pack(const float *__restrict__ p, MlpDevice *__restrict__ w) {
  const int l = threadIdx.x;  // 64 threads
  const int i = l & 15, k0 = 4 * (l >> 4);
  for (int m = 0; m < 2; ++m) {
    half4_t a;
    for (int v = 0; v < 4; ++v) {
      const int k = k0 + v;
      a[v] = k < kMlpIn ? h16(p[kOffW1 + (16 * m + i) * kMlpIn + k]) : (_Float16)0;
    }
    w->a[m][l] = a;
  }
  for (int m2 = 0; m2 < 2; ++m2)
    for (int kb = 0; kb < 2; ++kb) {
      half4_t a;
      for (int v = 0; v < 4; ++v) a[v] = h16(p[kOffW2 + (16 * m2 + i) * kMlpH1 + 16 * kb + k0 + v]);
      w->a[2 + 2 * m2 + kb][l] = a;
    }
  for (int kb = 0; kb < 2; ++kb) {
    half4_t a;
    for (int v = 0; v < 4; ++v)
      a[v] = i < kMlpOut ? h16(p[kOffW3 + i * kMlpH2 + 16 * kb + k0 + v]) : (_Float16)0;
    w->a[6 + kb][l] = a;
  }
  for (int m = 0; m < 2; ++m) {
    float4_t c1, c2;
    for (int v = 0; v < 4; ++v) {
      c1[v] = (float)h16(p[kOffB1 + 16 * m + k0 + v]);
      c2[v] = (float)h16(p[kOffB2 + 16 * m + k0 + v]);
    }
    w->c1[m][l] = c1;
    w->c2[m][l] = c2;
  }
  float4_t c3;
  for (int v = 0; v < 4; ++v) c3[v] = k0 + v < kMlpOut ? (float)h16(p[kOffB3 + k0 + v]) : 0.0f;
  w->c3[l] = c3;
}

// ---- forward (the same MFMA path as in-kernel inference) -------------------
__global__ __launch_bounds__(256) void k_mlp_forward(const MlpDevice *__restrict__ w,
                                                     const float *__restrict__ feat, int n,
                                                     float *__restrict__ logits) {
  const int lane = threadIdx.x & 63;
  const int gw = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int base = gw * 16; base < n; base += nw * 16) {
    const int row = base + (lane & 15);
    half4_t f = {0, 0, 0, 0};
    if (lane < 32 && row < n) {
      const float4 v = *(const float4 *)(feat + (size_t)row * 8 + (lane < 16 ? 0 : 4));
      f.x = (_Float16)v.x;
      f.y = (_Float16)v.y;
      f.z = (_Float16)v.z;
      f.w = (_Float16)v.w;
    }
    const float4_t z = mlp_eval16(w, lane, f);
    if (lane < 16 && row < n) {
      float4 o;
      o.x = z.x;
      o.y = z.y;
      o.z = z.z;
      o.w = z.w;
      *(float4 *)(logits + (size_t)row * 4) = o;
    }
  }
}

// ---- gradient of the mean cross-entropy ------------------------------------
// Block b takes batch positions [b*kGS, b*kGS + kGS); position q maps to
// sample (pa*(start+q) + pc) mod n (pa = 0: start+q).  Phase 1: one thread per
// sample runs forward and backward in FP32 with the FP16-rounded weights and
// leaves activations and deltas in LDS; phase 2: one thread per parameter sums
// its gradient over the block's samples in a fixed order.
__global__ __launch_bounds__(kGS) void k_mlp_grad_blocks(const float *__restrict__ params,
                                                         const float *__restrict__ feat,
                                                         const int *__restrict__ label, int n,
                                                         long long start, int count,
                                                         unsigned long long pa,
                                                         unsigned long long pc, float inv_count,
                                                         float *__restrict__ partial) {
  __shared__ float W[kMlpParams];
  __shared__ float F[kMlpIn][kGSP];
  __shared__ float A1[kMlpH1][kGSP];
  __shared__ float A2[kMlpH2][kGSP];
  __shared__ float D1[kMlpH1][kGSP];
  __shared__ float D2[kMlpH2][kGSP];
  __shared__ float DZ[kMlpOut][kGSP];
  __shared__ float Ls[kGS];
  const int tid = threadIdx.x;
  for (int q = tid; q < kMlpParams; q += kGS) W[q] = (float)(_Float16)params[q];
  __syncthreads();
  const int pos = blockIdx.x * kGS + tid;
  float loss = 0.0f;
  if (pos < count) {
    // Forward and backward of one sample, activations kept in this thread's
    // LDS column (conflict-free: rows are padded to kGSP).
    const unsigned long long g = (unsigned long long)(start + pos);
    const long long s = pa == 0 ? (long long)g : (long long)((pa * g + pc) % (unsigned long long)n);
    for (int c = 0; c < kMlpIn; ++c) F[c][tid] = feat[s * kMlpIn + c];
    for (int k = 0; k < kMlpH1; ++k) {
      float acc = W[kOffB1 + k];
#pragma unroll
      for (int c = 0; c < kMlpIn; ++c) acc += W[kOffW1 + k * kMlpIn + c] * F[c][tid];
      A1[k][tid] = fmaxf(acc, 0.0f);
    }
    for (int i = 0; i < kMlpH2; ++i) {
      float acc = W[kOffB2 + i];
#pragma unroll 8
      for (int k = 0; k < kMlpH1; ++k) acc += W[kOffW2 + i * kMlpH1 + k] * A1[k][tid];
      A2[i][tid] = fmaxf(acc, 0.0f);
    }
    float z[kMlpOut];
    float mx = -INFINITY;
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) {
      float acc = W[kOffB3 + o];
      for (int i = 0; i < kMlpH2; ++i) acc += W[kOffW3 + o * kMlpH2 + i] * A2[i][tid];
      z[o] = acc;
      mx = fmaxf(mx, acc);
    }
    float se = 0.0f;
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) se += __expf(z[o] - mx);
    const int lab = label[s];
    float zl = z[0];
#pragma unroll
    for (int o = 1; o < kMlpOut; ++o)
      if (o == lab) zl = z[o];
    loss = __logf(se) + mx - zl;
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o)
      DZ[o][tid] = (__expf(z[o] - mx) / se - (o == lab ? 1.0f : 0.0f)) * inv_count;
    for (int i = 0; i < kMlpH2; ++i) {
      float acc = 0.0f;
#pragma unroll
      for (int o = 0; o < kMlpOut; ++o) acc += W[kOffW3 + o * kMlpH2 + i] * DZ[o][tid];
      D2[i][tid] = A2[i][tid] > 0.0f ? acc : 0.0f;
    }
    for (int k = 0; k < kMlpH1; ++k) {
      float acc = 0.0f;
#pragma unroll 8
      for (int i = 0; i < kMlpH2; ++i) acc += W[kOffW2 + i * kMlpH1 + k] * D2[i][tid];
      D1[k][tid] = A1[k][tid] > 0.0f ? acc : 0.0f;
    }
  } else {
    for (int c = 0; c < kMlpIn; ++c) F[c][tid] = 0.0f;
    for (int k = 0; k < kMlpH1; ++k) A1[k][tid] = D1[k][tid] = 0.0f;
    for (int i = 0; i < kMlpH2; ++i) A2[i][tid] = D2[i][tid] = 0.0f;
    for (int o = 0; o < kMlpOut; ++o) DZ[o][tid] = 0.0f;
  }
  Ls[tid] = loss;
  __syncthreads();
  float *out = partial + (size_t)blockIdx.x * kPartial;
  for (int q = tid; q < kMlpParams; q += kGS) {
    const float *dsrc;
    const float *asrc = nullptr;
    if (q < kOffB1) {
      dsrc = D1[(q - kOffW1) / kMlpIn];
      asrc = F[(q - kOffW1) % kMlpIn];
    } else if (q < kOffW2) {
      dsrc = D1[q - kOffB1];
    } else if (q < kOffB2) {
      dsrc = D2[(q - kOffW2) / kMlpH1];
      asrc = A1[(q - kOffW2) % kMlpH1];
    } else if (q < kOffW3) {
      dsrc = D2[q - kOffB2];
    } else if (q < kOffB3) {
      dsrc = DZ[(q - kOffW3) / kMlpH2];
      asrc = A2[(q - kOffW3) % kMlpH2];
    } else {
      dsrc = DZ[q - kOffB3];
    }
    float g = 0.0f;
    if (asrc)
      for (int s = 0; s < kGS; ++s) g += dsrc[s] * asrc[s];
    else
      for (int s = 0; s < kGS; ++s) g += dsrc[s];
    out[q] = g;
  }
  if (tid == 0) {
    float L = 0.0f;
    for (int s = 0; s < kGS; ++s) L += Ls[s];
    out[kMlpParams] = L;
  }
}

// Sum of the block partials in block order: gradient and mean loss.
__global__ void k_mlp_reduce(const float *__restrict__ partial, int nblk, float inv_count,
                             float *__restrict__ grad, float *__restrict__ loss) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > kMlpParams) return;
  float g = 0.0f;
  for (int b = 0; b < nblk; ++b) g += partial[(size_t)b * kPartial + q];
  if (q < kMlpParams) grad[q] = g;
  else if (loss) *loss = g * inv_count;
}

// Adam on the FP32 master parameters (beta1 0.9, beta2 0.999, eps 1e-8).
__global__ void k_mlp_adam(const float *__restrict__ partial, int nblk, float *__restrict__ params,
                           float *__restrict__ m, float *__restrict__ v, float lr, float bc1,
                           float bc2, float *__restrict__ loss_acc, float inv_count) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > kMlpParams) return;
  float g = 0.0f;
  for (int b = 0; b < nblk; ++b) g += partial[(size_t)b * kPartial + q];
  if (q == kMlpParams) {
    *loss_acc += g;
    return;
  }
  const float mn = 0.9f * m[q] + 0.1f * g;
  const float vn = 0.999f * v[q] + 0.001f * g * g;
  m[q] = mn;
  v[q] = vn;
  params[q] -= lr * (mn / bc1) / (sqrtf(vn / bc2) + 1e-8f);
}

// ---- features and labels ---------------------------------------------------
// Squared norm of every 16-entry segment of v (rows < n count).
__global__ void k_seg_norm2(const double *__restrict__ v, int n, int nseg, float *__restrict__ seg2,
                            double *__restrict__ seg2d) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nseg) return;
  double s = 0.0;
  for (int i = 16 * c; i < 16 * c + 16 && i < n; ++i) s += v[i] * v[i];
  seg2[c] = (float)s;
  seg2d[c] = s;
}

// Deterministic single-block sum.
__global__ __launch_bounds__(1024) void k_sum_one_block(const double *__restrict__ v, int n,
                                                        double *__restrict__ out) {
  __shared__ double lds[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += v[i];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) *out = s;
}

__global__ void k_features(int T, const uint16_t *__restrict__ feat16,
                           const int *__restrict__ tile_col, const float *__restrict__ seg2,
                           float seg_scale, float progress, float *__restrict__ feat) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  for (int c = 0; c < 6; ++c)
    feat[(size_t)t * 8 + c] = (float)__builtin_bit_cast(_Float16, feat16[(size_t)t * 8 + c]);
  feat[(size_t)t * 8 + 6] = (float)(_Float16)feature_local_norm(seg2[tile_col[t]], seg_scale);
  feat[(size_t)t * 8 + 7] = (float)(_Float16)progress;
}

// y = A v in FP64 per tile-row (one thread per row), for ||A v||.
__global__ void k_rows_av2(TileView A, const double *__restrict__ v, int nrows,
                           double *__restrict__ out2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const int R = i >> 4, r = i & 15;
  double acc = 0.0;
  for (int t = A.tile_ptr[R]; t < A.tile_ptr[R + 1]; ++t) {
    const int xo = A.tile_col[t] << 4;
    const int k0 = A.tile_off[t], k1 = k0 + A.tile_nz[t] + 1;
    for (int k = k0; k < k1; ++k) {
      const int rc = A.rc[k];
      if ((rc >> 4) == r) acc += A.v64[k] * v[xo + (rc & 15)];
    }
  }
  out2[i] = acc * acc;
}

__global__ void k_labels(TileView A, const double *__restrict__ v, double thr,
                         int *__restrict__ label) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.tilenum) return;
  const int xo = A.tile_col[t] << 4;
  const float sc16 = A.s16[t], sc8 = A.s8[t];
  double e32 = 0.0, e16 = 0.0, e8 = 0.0;  // sums of squared row errors
  double r32 = 0.0, r16 = 0.0, r8 = 0.0;  // running row errors
  int row = -1;
  const int k0 = A.tile_off[t], k1 = k0 + A.tile_nz[t] + 1;
  for (int k = k0; k < k1; ++k) {
    const int rc = A.rc[k];
    if ((rc >> 4) != row) {
      e32 += r32 * r32;
      e16 += r16 * r16;
      e8 += r8 * r8;
      r32 = r16 = r8 = 0.0;
      row = rc >> 4;
    }
    const double xv = v[xo + (rc & 15)];
    const float x32 = (float)xv;
    const double exact = A.v64[k] * xv;
    r32 += (double)(A.v32[k] * x32) - exact;
    r16 += (double)(half_bits_to_float(A.v16[k]) * x32 * sc16) - exact;
    r8 += (double)((float)A.v8[k] * x32 * sc8) - exact;
  }
  e32 += r32 * r32;
  e16 += r16 * r16;
  e8 += r8 * r8;
  const double t2 = thr * thr;
  int lv = MF_PREC_FP64;
  if (e32 <= t2) lv = MF_PREC_FP32;
  if (e16 <= t2) lv = MF_PREC_FP16;
  if (e8 <= t2) lv = MF_PREC_INT8;
  label[t] = lv;
}

// ---- host helpers ----------------------------------------------------------
namespace {

uint16_t float_to_half_bits(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }

// Static features 0-5 of every tile, from the host tile arrays and the
// device quantisation errors.
mf_status build_static_features(mf_tiled *t) {
  if (t->dev.feat16) return MF_OK;
  const int T = t->tilenum, TM = t->tilem;
  std::vector<double> qerr(3 * (size_t)std::max(T, 1));
  if (T > 0)
    MF_HIP(hipMemcpy(qerr.data(), t->dev.qerr, sizeof(double) * 3 * (size_t)T, hipMemcpyDeviceToHost));
  std::vector<double> tsum(T, 0.0), amax(T, 0.0), amin(T, 0.0);
  std::vector<double> rowsum(TM, 0.0), dom(TM, INFINITY);
  for (int R = 0; R < TM; ++R) {
    double diag[16] = {0}, off[16] = {0};
    bool has[16] = {false};
    for (int ti = t->tile_ptr[R]; ti < t->tile_ptr[R + 1]; ++ti) {
      const int c = t->tile_columnidx[ti];
      const int z0 = t->csr_offset[ti], z1 = t->csr_offset[ti + 1];
      const int p0 = t->csrptr_offset[ti];
      double mx = 0.0, mn = INFINITY, s = 0.0;
      for (int r = 0; r < 16; ++r) {
        const int a = z0 + t->Blockcsr_Ptr[p0 + r];
        const int b = r < 15 ? z0 + t->Blockcsr_Ptr[p0 + r + 1] : z1;
        for (int k = a; k < b && k < z1; ++k) {
          const double v = std::fabs(t->Blockcsr_Val[k]);
          s += v;
          mx = std::max(mx, v);
          if (v > 0.0) mn = std::min(mn, v);
          has[r] = true;
          if (16 * c + t->Tile_csr_Col[k] == 16 * R + r) diag[r] += v;
          else off[r] += v;
        }
      }
      tsum[ti] = s;
      amax[ti] = mx;
      amin[ti] = mn == INFINITY ? 0.0 : mn;
      rowsum[R] += s;
    }
    for (int r = 0; r < 16; ++r) {
      if (!has[r]) continue;
      const double d = off[r] > 0.0 ? diag[r] / off[r] : INFINITY;
      dom[R] = std::min(dom[R], d);
    }
  }
  std::vector<uint16_t> f16((size_t)std::max(T, 1) * 8, 0);
  for (int R = 0; R < TM; ++R)
    for (int ti = t->tile_ptr[R]; ti < t->tile_ptr[R + 1]; ++ti) {
      const int nz = t->csr_offset[ti + 1] - t->csr_offset[ti];
      float f[6];
      f[0] = nz / 256.0f;
      f[1] = amin[ti] > 0.0 ? (float)(std::log2(amax[ti] / amin[ti]) / 24.0) : 0.0f;
      const double am = amax[ti];
      f[2] = am > 0.0 ? (float)(std::log2(qerr[T + ti] / am + 0x1p-40) / 40.0) : -1.0f;
      f[3] = am > 0.0 ? (float)(std::log2(qerr[2 * (size_t)T + ti] / am + 0x1p-40) / 40.0) : -1.0f;
      f[4] = rowsum[R] > 0.0 ? (float)(tsum[ti] / rowsum[R]) : 0.0f;
      double d = dom[R];
      double fd = d == INFINITY ? 1.0 : (d > 0.0 ? std::log2(d) / 8.0 : -1.0);
      f[5] = (float)std::max(-1.0, std::min(1.0, fd));
      for (int c = 0; c < 6; ++c) f16[(size_t)ti * 8 + c] = float_to_half_bits(f[c]);
    }
  uint16_t *d = nullptr;
  MF_HIP(hipMalloc(&d, sizeof(uint16_t) * f16.size()));
  const hipError_t e = hipMemcpy(d, f16.data(), sizeof(uint16_t) * f16.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return fail(MF_ERR_HIP, std::string("features upload: ") + hipGetErrorString(e));
  }
  t->dev.feat16 = d;
  return MF_OK;
}

// Segment norms of v and ||v||^2 into freshly allocated buffers.
struct SegNorms {
  float *seg2 = nullptr;
  double *seg2d = nullptr;
  double *sum = nullptr;
  ~SegNorms() {
    if (seg2) (void)hipFree(seg2);
    if (seg2d) (void)hipFree(seg2d);
    if (sum) (void)hipFree(sum);
  }
};

mf_status seg_norms(const mf_tiled *t, const double *d_v, hipStream_t st, SegNorms &s,
                    double *norm2) {
  const int nseg = std::max(t->tilem, t->tilen);
  MF_HIP(hipMalloc(&s.seg2, sizeof(float) * nseg));
  MF_HIP(hipMalloc(&s.seg2d, sizeof(double) * nseg));
  MF_HIP(hipMalloc(&s.sum, sizeof(double)));
  k_seg_norm2<<<(nseg + 255) / 256, 256, 0, st>>>(d_v, t->dev.npad, nseg, s.seg2, s.seg2d);
  MF_HIP(hipGetLastError());
  k_sum_one_block<<<1, 1024, 0, st>>>(s.seg2d, nseg, s.sum);
  MF_HIP(hipGetLastError());
  MF_HIP(hipMemcpyAsync(norm2, s.sum, sizeof(double), hipMemcpyDeviceToHost, st));
  MF_HIP(hipStreamSynchronize(st));
  return MF_OK;
}

mf_status check_handle(const mf_tiled *t, const char *fn) {
  if (!t) return fail(MF_ERR_ARG, std::string(fn) + ": null handle");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, std::string(fn) + ": host-only handle");
  return MF_OK;
}

}  // namespace

// Used by the solver: packed MLP of a handle, or null.
const MlpDevice *policy_mlp(const mf_tiled *t) {
  return t->dev.has_mlp ? (const MlpDevice *)t->dev.mlp : nullptr;
}

}  // namespace mf

extern "C" mf_status mf_mlp_forward(const float *d_params, const float *d_feat, int nrows,
                                    float *d_logits, void *stream) {
  if (!d_params || nrows < 0 || (nrows > 0 && (!d_feat || !d_logits)))
    return fail(MF_ERR_ARG, "mf_mlp_forward: bad argument");
  if (nrows == 0) return MF_OK;
  hipStream_t st = (hipStream_t)stream;
  MlpDevice *w = nullptr;
  MF_HIP(hipMalloc(&w, sizeof(MlpDevice)));
  k_mlp_pack<<<1, 64, 0, st>>>(d_params, w);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    const int batches = (nrows + 15) / 16;
    const int blocks = std::min((batches + 3) / 4, 4096);
    k_mlp_forward<<<blocks, 256, 0, st>>>(w, d_feat, nrows, d_logits);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(w);
  if (e != hipSuccess) return fail(MF_ERR_HIP, std::string("mf_mlp_forward: ") + hipGetErrorString(e));
  return MF_OK;
}

extern "C" mf_status mf_mlp_grad(const float *d_params, const float *d_feat, const int *d_label,
                                 int nrows, float *d_grad, float *d_loss, void *stream) {
  if (!d_params || !d_grad || nrows <= 0 || !d_feat || !d_label)
    return fail(MF_ERR_ARG, "mf_mlp_grad: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const int nblk = (nrows + kGS - 1) / kGS;
  float *part = nullptr;
  MF_HIP(hipMalloc(&part, sizeof(float) * kPartial * (size_t)nblk));
  const float inv = 1.0f / (float)nrows;
  k_mlp_grad_blocks<<<nblk, kGS, 0, st>>>(d_params, d_feat, d_label, nrows, 0, nrows, 0, 0, inv, part);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    k_mlp_reduce<<<(kPartial + 255) / 256, 256, 0, st>>>(part, nblk, inv, d_grad, d_loss);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(part);
  if (e != hipSuccess) return fail(MF_ERR_HIP, std::string("mf_mlp_grad: ") + hipGetErrorString(e));
  return MF_OK;
}

extern "C" mf_status mf_mlp_train(float *d_params, const float *d_feat, const int *d_label,
                                  int nrows, int epochs, int batch, float lr, unsigned seed,
                                  float *loss_out, void *stream) {
  if (!d_params || !d_feat || !d_label || nrows <= 0 || epochs < 0 || batch <= 0 || !(lr > 0.0f))
    return fail(MF_ERR_ARG, "mf_mlp_train: bad argument");
  hipStream_t st = (hipStream_t)stream;
  const int bsz = std::min(batch, nrows);
  const int nblk = (bsz + kGS - 1) / kGS;
  float *buf = nullptr;
  const size_t nb = (size_t)kPartial * nblk + 2 * (size_t)kMlpParams + 2;
  MF_HIP(hipMalloc(&buf, sizeof(float) * nb));
  float *part = buf, *m = part + (size_t)kPartial * nblk, *v = m + kMlpParams;
  float *lacc = v + kMlpParams;
  hipError_t e = hipMemsetAsync(m, 0, sizeof(float) * (2 * (size_t)kMlpParams + 2), st);
  unsigned long long rng = 0x9E3779B97F4A7C15ull ^ seed;
  auto next = [&rng]() {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return rng;
  };
  long long step = 0;
  float last = 0.0f;
  for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
    // epoch order: sample (pa*q + pc) mod n with pa coprime to n
    unsigned long long pa = 1, pc = 0;
    if (nrows > 1) {
      pa = (next() % (unsigned long long)nrows) | 1ull;
      auto gcd = [](unsigned long long a, unsigned long long b) {
        while (b) { const unsigned long long r = a % b; a = b; b = r; }
        return a;
      };
      while (gcd(pa, (unsigned long long)nrows) != 1) pa += 2;
      pc = next() % (unsigned long long)nrows;
    }
    e = hipMemsetAsync(lacc, 0, sizeof(float), st);
    for (long long s0 = 0; s0 < nrows && e == hipSuccess; s0 += bsz) {
      const int cnt = (int)std::min<long long>(bsz, nrows - s0);
      const int nbk = (cnt + kGS - 1) / kGS;
      const float inv = 1.0f / (float)cnt;
      k_mlp_grad_blocks<<<nbk, kGS, 0, st>>>(d_params, d_feat, d_label, nrows, s0, cnt, pa, pc, inv, part);
      ++step;
      const float bc1 = 1.0f - std::pow(0.9f, (float)step);
      const float bc2 = 1.0f - std::pow(0.999f, (float)step);
      k_mlp_adam<<<(kPartial + 255) / 256, 256, 0, st>>>(part, nbk, d_params, m, v, lr, bc1, bc2, lacc, inv);
      e = hipGetLastError();
    }
    if (e == hipSuccess) {
      float l = 0.0f;
      e = hipMemcpyAsync(&l, lacc, sizeof(float), hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      last = l / (float)nrows;
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(MF_ERR_HIP, std::string("mf_mlp_train: ") + hipGetErrorString(e));
  if (loss_out) *loss_out = last;
  return MF_OK;
}

extern "C" mf_status mf_policy_set_mlp(mf_tiled *t, const float *d_params) {
  MF_TRY(check_handle(t, "mf_policy_set_mlp"));
  MF_HIP(hipSetDevice(t->dev.device));
  if (!d_params) {
    t->dev.has_mlp = 0;
    return MF_OK;
  }
  MF_TRY(build_static_features(t));
  if (!t->dev.mlp) MF_HIP(hipMalloc(&t->dev.mlp, sizeof(MlpDevice)));
  k_mlp_pack<<<1, 64>>>(d_params, (MlpDevice *)t->dev.mlp);
  MF_HIP(hipGetLastError());
  MF_HIP(hipDeviceSynchronize());
  t->dev.has_mlp = 1;
  return MF_OK;
}

extern "C" mf_status mf_policy_features(mf_tiled *t, const double *d_p, double res_ratio,
                                        float *d_feat, void *stream) {
  MF_TRY(check_handle(t, "mf_policy_features"));
  if (!d_p || !d_feat) return fail(MF_ERR_ARG, "mf_policy_features: null vector");
  MF_HIP(hipSetDevice(t->dev.device));
  MF_TRY(build_static_features(t));
  hipStream_t st = (hipStream_t)stream;
  SegNorms s;
  double n2 = 0.0;
  MF_TRY(seg_norms(t, d_p, st, s, &n2));
  const int nseg = std::max(t->tilem, t->tilen);
  const float seg_scale = n2 > 0.0 ? (float)(std::sqrt((double)nseg) / std::sqrt(n2)) : 0.0f;
  const float prog = (float)(std::log10(std::max(res_ratio, 1e-16)) / 16.0);
  if (t->tilenum > 0) {
    k_features<<<(t->tilenum + 255) / 256, 256, 0, st>>>(t->tilenum, t->dev.feat16, t->dev.tile_col,
                                                          s.seg2, seg_scale, prog, d_feat);
    MF_HIP(hipGetLastError());
  }
  MF_HIP(hipStreamSynchronize(st));
  return MF_OK;
}

extern "C" mf_status mf_policy_labels(mf_tiled *t, const double *d_p, double theta, int *d_label,
                                      void *stream) {
  MF_TRY(check_handle(t, "mf_policy_labels"));
  if (!d_p || !d_label || !(theta >= 0.0)) return fail(MF_ERR_ARG, "mf_policy_labels: bad argument");
  MF_HIP(hipSetDevice(t->dev.device));
  hipStream_t st = (hipStream_t)stream;
  const int n = 16 * t->tilem;
  double *buf = nullptr;
  MF_HIP(hipMalloc(&buf, sizeof(double) * ((size_t)n + 1)));
  const TileView A = view_of(t);
  hipError_t e = hipSuccess;
  if (n > 0) {
    k_rows_av2<<<(n + 255) / 256, 256, 0, st>>>(A, d_p, n, buf);
    e = hipGetLastError();
    if (e == hipSuccess) {
      k_sum_one_block<<<1, 1024, 0, st>>>(buf, n, buf + n);
      e = hipGetLastError();
    }
  }
  double w2 = 0.0;
  if (e == hipSuccess && n > 0) e = hipMemcpyAsync(&w2, buf + n, sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess && t->tilenum > 0) {
    const double thr = theta * std::sqrt(w2) / std::sqrt((double)t->tilenum);
    k_labels<<<(t->tilenum + 255) / 256, 256, 0, st>>>(A, d_p, thr, d_label);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(MF_ERR_HIP, std::string("mf_policy_labels: ") + hipGetErrorString(e));
  return MF_OK;
}
