// mf_spmv.hip -- stand-alone tiled SpMV y = A x (wave64, one wave per
// tile-row, no atomics).
//
// Launch shape: 256-thread blocks (4 waves = 4 tile-rows).  Blocks are
// dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8), so the
// block index is remapped such that each XCD walks one contiguous eighth of
// the tile-rows: neighbouring tile-rows share x segments, and keeping them on
// one XCD keeps those segments in that XCD's L2.
#include "mf_device.h"

namespace mf {

template <int MODE>
__global__ __launch_bounds__(256) void k_spmv(TileView A, const double *__restrict__ x,
                                              const float *__restrict__ x32,
                                              double *__restrict__ y) {
  const int per_xcd = gridDim.x >> 3;  // gridDim.x is a multiple of 8
  const int lb = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  const int bi = lb * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (bi >= A.tilem) return;
  const double v = tile_row_sum<MODE>(A, bi, lane, x, x32);
  if (lane < 16) y[bi * 16 + lane] = v;
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
  const int blocks = ((A.tilem + 3) / 4 + 7) / 8 * 8;
  if (A.tilem > 0) {
    if (mode == kAll64)
      k_spmv<kAll64><<<blocks, 256, 0, s>>>(A, d_x, d_x32, d_y);
    else
      k_spmv<kPerTile><<<blocks, 256, 0, s>>>(A, d_x, d_x32, d_y);
    MF_HIP(hipGetLastError());
  }
  const int from = 16 * A.tilem, to = A.npad;
  if (to > from) {
    k_zero_tail<<<(to - from + 255) / 256, 256, 0, s>>>(d_y, from, to);
    MF_HIP(hipGetLastError());
  }
  return MF_OK;
}

}  // namespace mf

using namespace mf;

extern "C" mf_status mf_spmv(mf_tiled *t, const double *d_x, const float *d_x32, double *d_y,
                             int mode, void *stream) {
  if (!t || !d_x || !d_y) return fail(MF_ERR_ARG, "mf_spmv: null argument");
  if (mode != kAll64 && mode != kPerTile) return fail(MF_ERR_ARG, "mf_spmv: bad mode");
  if (mode == kPerTile && !d_x32) return fail(MF_ERR_ARG, "mf_spmv: mode 1 needs the FP32 copy of x");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, "mf_spmv: host-only handle");
  MF_HIP(hipSetDevice(t->dev.device));
  return launch_spmv(t, d_x, d_x32, d_y, mode, (hipStream_t)stream);
}
