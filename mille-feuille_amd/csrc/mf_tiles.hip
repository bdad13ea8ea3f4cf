// mf_tiles.hip -- device layout of the tiled matrix: upload, payload packing
// (FP32 / FP16 / INT8 copies with per-tile scales, computed on the GPU) and
// per-tile precision control.
//
// HBM layout per handle (T = tilenum, Z = nnz, ZP = payload slots: each
// tile's entries padded to a multiple of 4, so that a 4-entry chunk never
// straddles two tiles; Z <= ZP < Z + 3T):
//   tile_ptr  (tilem+1) i32    tile-row -> first tile
//   tile_col  T i32            tile column
//   tile_off  (T+1) i32        payload offset of each tile, a multiple of 4
//                              (ZP at the end)
//   tile_row  T i32            tile-row of each tile
//   tile_nz   T u8             nnz of the tile - 1
//   rc        ZP u8            (row << 4) | col of the entry inside its tile
//                              (padding repeats the tile's last byte)
//   v64/v32/v16/v8             ZP entries each (padding 0): FP64, FP32, FP16 (scaled by
//                              a power of two per tile), INT8 (scaled by a
//                              power of two per tile)
//   s16, s8   T f32            per-tile scales
//   qerr      3T f64           per-tile max |a - dequant(a)| for FP32/16/INT8
//   amax      T f64            per-tile max |a|
//   prec      T u8             current precision level of each tile
//   spmv_desc                  per-wave work (tile-rows, tiles, entries)
//                              of the stand-alone SpMV
// Entries keep the reference's order (tile by tile, row by row inside a
// tile, src/csr2block.h:239-252); the per-entry row nibble replaces the
// reference's 16 row pointers per tile on the device, so one SpMV reads
// value + 1 byte per entry and 12-16 bytes per tile.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "mf_device.h"

namespace mf {

// Payload arrays carry kPad extra zero entries: the streamed SpMV reads
// entries in aligned groups of four and may touch up to three past the end.
constexpr int kPad = 64;
constexpr int kSpmvWaveNnz = 1024;

__global__ void k_pack_payloads(int T, const int *__restrict__ off, const double *__restrict__ v64,
                                float *__restrict__ v32, uint16_t *__restrict__ v16,
                                int8_t *__restrict__ v8, float *__restrict__ s16,
                                float *__restrict__ s8, double *__restrict__ qerr,
                                double *__restrict__ amax) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int a = off[t], b = off[t + 1];
  double m = 0.0;
  for (int k = a; k < b; ++k) m = fmax(m, fabs(v64[k]));
  amax[t] = m;
  // FP16: power-of-two scale 2^e with e = floor(log2 max|a|), so the scaled
  // magnitudes lie in [.., 2); scaling by a power of two is exact in FP32, so
  // for in-range tiles the stored value equals (half)(float)a.
  int e = m > 0.0 ? ilogb(m) : 0;
  e = e < -100 ? -100 : (e > 100 ? 100 : e);
  const float sc16 = ldexpf(1.0f, e);
  const float inv16 = ldexpf(1.0f, -e);
  // INT8: smallest power-of-two scale with max|a|/scale <= 127 (exact
  // division; tiles of multiples of the scale, such as the -1 couplings of a
  // Laplacian, are stored exactly), round to nearest even.
  float sc8 = 1.0f;
  if (m > 0.0) {
    int e8 = ilogb(m / 127.0);
    if (ldexp(127.0, e8) < m) ++e8;
    e8 = e8 < -120 ? -120 : (e8 > 120 ? 120 : e8);
    sc8 = ldexpf(1.0f, e8);
  }
  s16[t] = sc16;
  s8[t] = sc8;
  double e32 = 0.0, e16 = 0.0, e8 = 0.0;
  for (int k = a; k < b; ++k) {
    const double v = v64[k];
    const float f = (float)v;
    v32[k] = f;
    e32 = fmax(e32, fabs(v - (double)f));
    const _Float16 h = (_Float16)(f * inv16);
    v16[k] = __builtin_bit_cast(uint16_t, h);
    e16 = fmax(e16, fabs(v - (double)((float)h * sc16)));
    double q = rint(v / (double)sc8);
    q = q > 127.0 ? 127.0 : (q < -127.0 ? -127.0 : q);
    v8[k] = (int8_t)q;
    e8 = fmax(e8, fabs(v - q * (double)sc8));
  }
  qerr[t] = e32;
  qerr[T + t] = e16;
  qerr[2 * T + t] = e8;
}

template <typename T>
static mf_status dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(MF_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  return MF_OK;
}


mf_status upload_spmv_schedule(mf_tiled *t, int wave_nnz);

mf_status upload_tiles(mf_tiled *t, int device) {
  DeviceLayout &d = t->dev;
  MF_HIP(hipSetDevice(device));
  d.device = device;
  const int T = t->tilenum, TM = t->tilem;
  d.npad = kTile * std::max(t->tilem, t->tilen);
  if (d.npad == 0) d.npad = kTile;

  // Payload layout: each tile's entries start at a multiple of 4 (padded per
  // tile to a multiple of 4 with zero values that repeat the tile's last
  // (row,col) byte), so every 4-entry chunk of the streamed SpMV lies in one
  // tile.  Per entry: (row << 4) | col inside the tile, from the tile's row
  // starts.  Per tile: the 16 row starts (rows past the last row of a short
  // final tile-row start at the tile's nnz, which is then <= 15*16) and
  // nnz - 1.
  std::vector<int> &poff = t->pay_off;
  poff.assign(T + 1, 0);
  for (int tt = 0; tt < T; ++tt) {
    const int nzt = t->csr_offset[tt + 1] - t->csr_offset[tt];
    poff[tt + 1] = poff[tt] + ((nzt + 3) & ~3);
  }
  const int ZP = poff[T];
  d.npay = ZP;
  std::vector<int> tile_row(T);
  std::vector<uint8_t> rc((size_t)ZP + kPad, 0);
  std::vector<uint8_t> tnz(T > 0 ? T : 1, 0);
  std::vector<double> v64((size_t)ZP + kPad, 0.0);
  std::vector<float> v32((size_t)ZP + kPad, 0.0f);
  std::vector<uint8_t> rp16((size_t)T * kTile + kPad, 0);
  std::vector<int> diag(TM, -1);
  int max_tiles = 0;
  for (int bi = 0; bi < TM; ++bi) {
    const int h = bi == TM - 1 ? t->rowA - (TM - 1) * kTile : kTile;
    max_tiles = std::max(max_tiles, t->tile_ptr[bi + 1] - t->tile_ptr[bi]);
    for (int tt = t->tile_ptr[bi]; tt < t->tile_ptr[bi + 1]; ++tt) {
      tile_row[tt] = bi;
      if (t->tile_columnidx[tt] == bi) diag[bi] = tt;
      const int off = t->csr_offset[tt], nzt = t->csr_offset[tt + 1] - off;
      const int po = poff[tt];
      tnz[tt] = (uint8_t)(nzt - 1);
      const unsigned char *rp = &t->Blockcsr_Ptr[t->csrptr_offset[tt]];
      for (int r = 0; r < kTile; ++r) rp16[(size_t)tt * kTile + r] = (uint8_t)(r < h ? rp[r] : nzt);
      for (int r = 0; r < h; ++r) {
        const int e = r == h - 1 ? nzt : rp[r + 1];
        for (int k = rp[r]; k < e; ++k)
          rc[po + k] = (uint8_t)((r << 4) | t->Tile_csr_Col[off + k]);
      }
      for (int k = 0; k < nzt; ++k) {
        v64[po + k] = t->Blockcsr_Val[off + k];
        v32[po + k] = t->Blockcsr_Val_Low[off + k];
      }
      for (int k = nzt; k < poff[tt + 1] - po; ++k) rc[po + k] = rc[po + nzt - 1];
    }
  }
  d.max_tiles_per_row = max_tiles;
  d.spmv_kernel = default_spmv_kernel(t);
  MF_TRY(dalloc(&d.tile_ptr, TM + 1));
  MF_TRY(dalloc(&d.tile_col, T));
  MF_TRY(dalloc(&d.tile_off, T + 1));
  MF_TRY(dalloc(&d.tile_row, T));
  MF_TRY(dalloc(&d.tile_nz, T));
  MF_TRY(dalloc(&d.rc, (size_t)ZP + kPad));
  MF_TRY(dalloc(&d.rp16, (size_t)T * kTile + kPad));
  MF_TRY(dalloc(&d.v64, (size_t)ZP + kPad));
  MF_TRY(dalloc(&d.v32, (size_t)ZP + kPad));
  MF_TRY(dalloc(&d.v16, (size_t)ZP + kPad));
  MF_TRY(dalloc(&d.v8, (size_t)ZP + kPad));
  MF_TRY(dalloc(&d.s16, T));
  MF_TRY(dalloc(&d.s8, T));
  MF_TRY(dalloc(&d.qerr, 3 * (size_t)T));
  MF_TRY(dalloc(&d.amax, T));
  MF_TRY(dalloc(&d.prec, T));
  MF_TRY(dalloc(&d.prec_user, T));
  MF_TRY(dalloc(&d.diag_tile, TM));
  MF_HIP(hipMemcpy(d.tile_ptr, t->tile_ptr.data(), sizeof(int) * (TM + 1), hipMemcpyHostToDevice));
  MF_HIP(hipMemcpy(d.tile_col, t->tile_columnidx.data(), sizeof(int) * T, hipMemcpyHostToDevice));
  MF_HIP(hipMemcpy(d.tile_off, poff.data(), sizeof(int) * (T + 1), hipMemcpyHostToDevice));
  if (T > 0) MF_HIP(hipMemcpy(d.tile_nz, tnz.data(), T, hipMemcpyHostToDevice));
  MF_HIP(hipMemcpy(d.tile_row, tile_row.data(), sizeof(int) * T, hipMemcpyHostToDevice));
  MF_HIP(hipMemcpy(d.rc, rc.data(), (size_t)ZP + kPad, hipMemcpyHostToDevice));
  MF_HIP(hipMemcpy(d.rp16, rp16.data(), rp16.size(), hipMemcpyHostToDevice));
  MF_HIP(hipMemset(d.v16, 0, sizeof(uint16_t) * ((size_t)ZP + kPad)));
  MF_HIP(hipMemset(d.v8, 0, (size_t)ZP + kPad));
  MF_HIP(hipMemcpy(d.v64, v64.data(), sizeof(double) * ((size_t)ZP + kPad), hipMemcpyHostToDevice));
  MF_HIP(hipMemcpy(d.diag_tile, diag.data(), sizeof(int) * TM, hipMemcpyHostToDevice));
  // per-wave work of the stand-alone SpMV, about kSpmvWaveNnz entries per
  // wave (retuned by mf_tile_set_schedule)
  MF_TRY(upload_spmv_schedule(t, kSpmvWaveNnz));
  MF_HIP(hipMemset(d.prec, MF_PREC_FP64, T));
  MF_HIP(hipMemset(d.prec_user, MF_PREC_FP64, T));
  if (T > 0) {
    k_pack_payloads<<<(T + 255) / 256, 256>>>(T, d.tile_off, d.v64, d.v32, d.v16, d.v8, d.s16,
                                              d.s8, d.qerr, d.amax);
    MF_HIP(hipGetLastError());
  }
  // The FP32 payload keeps the caller's FP32 copy when one was given
  // (the reference passes Val_Low, src/csr2block.h:248).
  MF_HIP(hipMemcpy(d.v32, v32.data(), sizeof(float) * ((size_t)ZP + kPad), hipMemcpyHostToDevice));
  MF_HIP(hipDeviceSynchronize());
  return MF_OK;
}

bool regular_tilerows(const mf_tiled *t) {
  const int TM = t->tilem;
  if (TM == 0) return false;
  const double mean = (double)t->tilenum / TM;
  return t->dev.max_tiles_per_row <= 4.0 * mean + 8.0;
}

int default_spmv_kernel(const mf_tiled *) {
  // The streamed kernel handles any sparsity and was the faster of the two
  // on the 3D 7-point 128^3 Laplacian on MI355X (mf_tune_spmv also times
  // the row-group kernel on matrices with regular tile-rows).
  return MF_SPMV_STREAM;
}

// Stand-alone SpMV schedule of about wave_nnz payload slots per wave, heavy
// tile-rows cut into one-window pieces, uploaded in place of the previous one.
mf_status upload_spmv_schedule(mf_tiled *t, int wave_nnz) {
  DeviceLayout &d = t->dev;
  std::vector<int> hn, hbase;
  const std::vector<WaveDesc> wd = window_partition(t, wave_nnz, &hn, &hbase);
  const int nw = (int)wd.size(), nh = (int)hn.size();
  const int npieces = nh > 0 ? hbase.back() + hn.back() : 0;
  for (void *p : {(void *)d.spmv_desc, (void *)d.spmv_part, (void *)d.spmv_cnt, (void *)d.spmv_hn,
                  (void *)d.spmv_hbase})
    if (p) MF_HIP(hipFree(p));
  d.spmv_desc = nullptr;
  d.spmv_part = nullptr;
  d.spmv_cnt = nullptr;
  d.spmv_hn = d.spmv_hbase = nullptr;
  d.spmv_nwaves = 0;
  d.spmv_nheavy = 0;
  if (nw > 0) {
    MF_TRY(dalloc(&d.spmv_desc, nw));
    MF_HIP(hipMemcpy(d.spmv_desc, wd.data(), sizeof(WaveDesc) * nw, hipMemcpyHostToDevice));
  }
  MF_TRY(dalloc(&d.spmv_part, 16 * (size_t)std::max(npieces, 1)));
  MF_TRY(dalloc(&d.spmv_cnt, std::max(nh, 1)));
  MF_TRY(dalloc(&d.spmv_hn, std::max(nh, 1)));
  MF_TRY(dalloc(&d.spmv_hbase, std::max(nh, 1)));
  MF_HIP(hipMemset(d.spmv_cnt, 0, sizeof(unsigned) * std::max(nh, 1)));
  if (nh > 0) {
    MF_HIP(hipMemcpy(d.spmv_hn, hn.data(), sizeof(int) * nh, hipMemcpyHostToDevice));
    MF_HIP(hipMemcpy(d.spmv_hbase, hbase.data(), sizeof(int) * nh, hipMemcpyHostToDevice));
  }
  d.spmv_nwaves = nw;
  d.spmv_nheavy = nh;
  d.spmv_wave_nnz = wave_nnz;
  return MF_OK;
}

void free_device(mf_tiled *t) {
  DeviceLayout &d = t->dev;
  if (d.device >= 0) (void)hipSetDevice(d.device);
  void *ptrs[] = {d.tile_ptr, d.tile_col, d.tile_off, d.tile_row, d.rc, d.rp16, d.spmv_desc, d.tile_nz,
                  d.spmv_part, d.spmv_cnt, d.spmv_hn, d.spmv_hbase,
                  d.v64,      d.v32,      d.v16,      d.v8,       d.s16,    d.s8,
                  d.qerr,     d.amax,     d.prec,     d.prec_user, d.diag_tile, d.mlp, d.feat16,
                  t->sched.wave_tr, t->sched.wave_res, t->sched.wave_win, t->sched.win_desc, t->sched.heavy_part,
                  t->sched.heavy_cnt, t->sched.heavy_n, t->sched.heavy_base, t->sched.heavy_tr,
                  t->sched.wave_heavy, t->work};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  d = DeviceLayout();
  t->sched = Schedule();
  t->work = nullptr;
  t->work_bytes = 0;
}

}  // namespace mf

using namespace mf;

extern "C" mf_status mf_tile_create(int rowA, int colA, const int *rowptr, const int *colidx,
                                    const double *val, const float *val_low, int device,
                                    mf_tiled **out) {
  if (!out) return fail(MF_ERR_ARG, "mf_tile_create: null output");
  *out = nullptr;
  if (device != -1) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      return fail(MF_ERR_HIP, "mf_tile_create: no HIP device");
    if (device < 0 || device >= ndev) return fail(MF_ERR_ARG, "mf_tile_create: bad device");
  }
  mf_tiled *t = new mf_tiled();
  mf_status st = build_tiles_host(t, rowA, colA, rowptr, colidx, val, val_low);
  if (st == MF_OK && device != -1) st = upload_tiles(t, device);
  if (device == -1) t->dev.npad = kTile * std::max(std::max(t->tilem, t->tilen), 1);
  if (st != MF_OK) {
    std::string msg = mf_last_error();
    free_device(t);
    delete t;
    set_error(msg);
    return st;
  }
  *out = t;
  return MF_OK;
}

extern "C" void mf_tile_free(mf_tiled *t) {
  if (!t) return;
  free_device(t);
  delete t;
}

extern "C" mf_status mf_tile_set_precision(mf_tiled *t, const unsigned char *levels, int uniform) {
  if (!t) return fail(MF_ERR_ARG, "mf_tile_set_precision: null handle");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, "mf_tile_set_precision: host-only handle");
  MF_HIP(hipSetDevice(t->dev.device));
  const int T = t->tilenum;
  if (levels) {
    for (int i = 0; i < T; ++i)
      if (levels[i] > MF_PREC_SKIP) return fail(MF_ERR_ARG, "mf_tile_set_precision: bad level");
    MF_HIP(hipMemcpy(t->dev.prec, levels, T, hipMemcpyHostToDevice));
    MF_HIP(hipMemcpy(t->dev.prec_user, levels, T, hipMemcpyHostToDevice));
  } else {
    if (uniform < 0 || uniform > MF_PREC_SKIP) return fail(MF_ERR_ARG, "mf_tile_set_precision: bad level");
    MF_HIP(hipMemset(t->dev.prec, uniform, T));
    MF_HIP(hipMemset(t->dev.prec_user, uniform, T));
  }
  return MF_OK;
}

extern "C" mf_status mf_tile_get_precision(const mf_tiled *t, unsigned char *levels) {
  if (!t || !levels) return fail(MF_ERR_ARG, "mf_tile_get_precision: null argument");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, "mf_tile_get_precision: host-only handle");
  MF_HIP(hipSetDevice(t->dev.device));
  MF_HIP(hipMemcpy(levels, t->dev.prec, t->tilenum, hipMemcpyDeviceToHost));
  return MF_OK;
}

extern "C" mf_status mf_tile_get_scales(const mf_tiled *t, float *scale16, float *scale8) {
  if (!t) return fail(MF_ERR_ARG, "mf_tile_get_scales: null handle");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, "mf_tile_get_scales: host-only handle");
  MF_HIP(hipSetDevice(t->dev.device));
  if (scale16) MF_HIP(hipMemcpy(scale16, t->dev.s16, sizeof(float) * t->tilenum, hipMemcpyDeviceToHost));
  if (scale8) MF_HIP(hipMemcpy(scale8, t->dev.s8, sizeof(float) * t->tilenum, hipMemcpyDeviceToHost));
  return MF_OK;
}

extern "C" mf_status mf_tile_get_quant_error(const mf_tiled *t, double *err) {
  if (!t || !err) return fail(MF_ERR_ARG, "mf_tile_get_quant_error: null argument");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, "mf_tile_get_quant_error: host-only handle");
  MF_HIP(hipSetDevice(t->dev.device));
  MF_HIP(hipMemcpy(err, t->dev.qerr, sizeof(double) * 3 * (size_t)t->tilenum, hipMemcpyDeviceToHost));
  return MF_OK;
}

extern "C" mf_status mf_tile_set_schedule(mf_tiled *t, int spmv_wave_nnz) {
  if (!t) return fail(MF_ERR_ARG, "mf_tile_set_schedule: null handle");
  if (t->dev.device < 0) return fail(MF_ERR_ARG, "mf_tile_set_schedule: host-only handle");
  if (spmv_wave_nnz == 0) return MF_OK;
  if (spmv_wave_nnz < 64 || spmv_wave_nnz > (1 << 20))
    return fail(MF_ERR_ARG, "mf_tile_set_schedule: spmv_wave_nnz out of [64, 2^20]");
  MF_HIP(hipSetDevice(t->dev.device));
  MF_HIP(hipDeviceSynchronize());
  MF_TRY(upload_spmv_schedule(t, spmv_wave_nnz));
  return MF_OK;
}
