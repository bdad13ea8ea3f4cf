// mf_spmv.hip -- stand-alone tiled SpMV y = A x (wave64, no global atomics),
// in two forms: streamed (wave_spmv: nnz-balanced tile-row ranges, LDS row
// accumulators; any matrix) and row-group (rowgroup_row: one row per lane,
// register sums; tile-rows of similar tile counts).
//
// Launch shape: 256-thread blocks; streamed: one nnz-balanced range of
// tile-rows per wave (about 512 entries); row-group: four tile-rows per wave.  Blocks are dispatched round-robin over the 8
// XCDs (block b runs on XCD b % 8), so the block index is remapped such that
// each XCD walks one contiguous eighth of the matrix: neighbouring tile-rows
// share x segments, and keeping them on one XCD keeps those segments in that
// XCD's L2.
#include "mf_device.h"

namespace mf {

// Occupancy: the latency of a wave's window (metadata -> payload -> x
// gathers) is hidden by waves in flight, so the register budget is capped at
// 64 VGPRs (8 waves per SIMD; 32 waves x 4.75 KiB of LDS per CU) for FP64
// payloads and 80 (6 waves) for per-tile levels, the most that compile
// without spilling.
// PIECES: the schedule has heavy tile-rows cut into pieces (wave_spmv_piece);
// without them the kernel keeps the piece bookkeeping out of its registers.
template <int MODE, bool PIECES>
__global__ __launch_bounds__(256, MODE == kAll64 ? 8 : 5) void k_spmv(TileView A, const WaveDesc *__restrict__ desc,
                                              int nwaves, HeavyView hv, const double *__restrict__ x,
                                              const float *__restrict__ x32,
                                              double *__restrict__ y) {
  __shared__ WaveLds wl[4];
  const int per_xcd = gridDim.x >> 3;  // gridDim.x is a multiple of 8
  const int lb = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  const int w = __builtin_amdgcn_readfirstlane(lb * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (w >= nwaves) return;
  auto store = [&](const int i, const double v) { y[i] = v; };
  if (PIECES)
    wave_spmv_piece<MODE, false, 3, true, true>(A, desc[w], hv, x, x32, nullptr, 0,
                                                wl[threadIdx.x >> 6], lane, nullptr, store);
  else
    wave_spmv<MODE, false>(A, desc[w], x, x32, nullptr, 0, wl[threadIdx.x >> 6], lane, nullptr, store);
}

// Row-group kernel: four consecutive tile-rows per wave, one row per lane,
// row sums in registers (rowgroup_row).  Same XCD-contiguous block order.
template <int MODE>
__global__ __launch_bounds__(256) void k_spmv_rg(TileView A, const double *__restrict__ x,
                                                 const float *__restrict__ x32,
                                                 double *__restrict__ y) {
  const int per_xcd = gridDim.x >> 3;
  const int lb = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  const int lane = threadIdx.x & 63;
  const int R0 = (lb * 4 + (threadIdx.x >> 6)) * 4;
  if (R0 >= A.tilem) return;  // wave-uniform
  const int R = R0 + (lane >> 4);
  const double v = rowgroup_row<MODE, false>(A, R, R < A.tilem, lane & 15, x, x32, nullptr, 0, nullptr);
  if (R < A.tilem) y[16 * R + (lane & 15)] = v;
}

// Zero the padding of y beyond the tile-rows (npad may exceed 16*tilem when
// the matrix has more tile columns than tile-rows).
__global__ void k_zero_tail(double *y, int from, int to) {
  const int i = from + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < to) y[i] = 0.0;
}

mf_status launch_spmv(const mf_tiled *t, const double *d_x, const float *d_x32, double *d_y,
                      int mode, hipStream_t s) {
  const TileView A = view_of(t);
  if (t->dev.spmv_kernel == MF_SPMV_ROWGROUP) {
    const int nwv = (A.tilem + 3) / 4;
    const int blocks = ((nwv + 3) / 4 + 7) / 8 * 8;
    if (A.tilem > 0) {
      if (mode == kAll64)
        k_spmv_rg<kAll64><<<blocks, 256, 0, s>>>(A, d_x, d_x32, d_y);
      else
        k_spmv_rg<kPerTile><<<blocks, 256, 0, s>>>(A, d_x, d_x32, d_y);
      MF_HIP(hipGetLastError());
    }
  } else {
    const int nw = t->dev.spmv_nwaves;
    const HeavyView hv{t->dev.spmv_part, t->dev.spmv_cnt, t->dev.spmv_hn, t->dev.spmv_hbase};
    const int blocks = ((nw + 3) / 4 + 7) / 8 * 8;
    if (A.tilem > 0 && nw > 0) {
      const bool pieces = t->dev.spmv_nheavy > 0;
      const WaveDesc *desc = t->dev.spmv_desc;
      if (mode == kAll64 && pieces)
        k_spmv<kAll64, true><<<blocks, 256, 0, s>>>(A, desc, nw, hv, d_x, d_x32, d_y);
      else if (mode == kAll64)
        k_spmv<kAll64, false><<<blocks, 256, 0, s>>>(A, desc, nw, hv, d_x, d_x32, d_y);
      else if (pieces)
        k_spmv<kPerTile, true><<<blocks, 256, 0, s>>>(A, desc, nw, hv, d_x, d_x32, d_y);
      else
        k_spmv<kPerTile, false><<<blocks, 256, 0, s>>>(A, desc, nw, hv, d_x, d_x32, d_y);
      MF_HIP(hipGetLastError());
    }
  }
  const int from = 16 * A.tilem, to = A.npad;
  if (to > from) {
    k_zero_tail<<<(to - from + 255) / 256, 256, 0, s>>>(d_y, from, to);
    MF_HIP(hipGetLastError());
  }
  return MF_OK;
}

__global__ __launch_bounds__(256) void k_stream_copy(const double4 *__restrict__ src,
                                                     double4 *__restrict__ dst, int64_t n4) {
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n4; i += (int64_t)gridDim.x * 256)
    dst[i] = src[i];
}

__global__ void k_copy_tail(const double *__restrict__ src, double *__restrict__ dst, int64_t from,
                            int64_t n) {
  const int64_t i = from + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace mf

using namespace mf;

extern "C" mf_status mf_stream_copy(const double *d_src, double *d_dst, int64_t n, void *stream) {
  if (!d_src || !d_dst || n < 0) return fail(MF_ERR_ARG, "mf_stream_copy: bad argument");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n4 = n / 4;
  if (((uintptr_t)d_src | (uintptr_t)d_dst) & 31) return fail(MF_ERR_ARG, "mf_stream_copy: 32-byte alignment");
  if (n4 > 0) {
    k_stream_copy<<<2048, 256, 0, s>>>((const double4 *)d_src, (double4 *)d_dst, n4);
    MF_HIP(hipGetLastError());
  }
  if (n > 4 * n4) {
    k_copy_tail<<<1, 64, 0, s>>>(d_src, d_dst, 4 * n4, n);
    MF_HIP(hipGetLastError());
  }
  return MF_OK;
}

extern "C" mf_status mf_tile_set_spmv_kernel(mf_tiled *t, int kind) {
  if (!t) return fail(MF_ERR_ARG, "mf_tile_set_spmv_kernel: null handle");
  if (kind == MF_SPMV_AUTO) kind = default_spmv_kernel(t);
  if (kind != MF_SPMV_STREAM && kind != MF_SPMV_ROWGROUP)
    return fail(MF_ERR_ARG, "mf_tile_set_spmv_kernel: unknown kernel");
  t->dev.spmv_kernel = kind;
  return MF_OK;
}

extern "C" mf_status mf_spmv(mf_tiled *t, const double *d_x, const float *d_x32, double *d_y,
                             int mode, void *stream) {
  if (!t || !d_x || !d_y) return fail(MF_ERR_ARG, "mf_spmv: null argument");
  if (mode != kAll64 && mode != kPerTile) return fail(MF_ERR_ARG, "mf_spmv: bad mode");
  if (mode == kPerTile && !d_x32) return fail(MF_ERR_ARG, "mf_spmv: mode 1 needs the FP32 copy of x");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, "mf_spmv: host-only handle");
  MF_HIP(hipSetDevice(t->dev.device));
  return launch_spmv(t, d_x, d_x32, d_y, mode, (hipStream_t)stream);
}
