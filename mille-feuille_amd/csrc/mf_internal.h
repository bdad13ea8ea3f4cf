// mf_internal.h -- shared internal declarations of libmillefeuille_hip.
//
// Host-only (no HIP types) so that the format builder compiles with g++.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "millefeuille.h"

namespace mf {

constexpr int kTile = 16;  // tile edge (reference BLOCK_SIZE, src/common.h:47-50)

// Window of the streamed SpMV: at most kWinTiles tiles spanning at most
// kWinRows tile-rows and kWinChunks 4-entry chunks of payload are staged in
// a wave's LDS at a time.
constexpr int kWinTiles = 128;
constexpr int kWinRows = 16;
constexpr int kWinChunks = 1024;

// Work of one wave of the streamed SpMV: tile-rows [tr0, tr1), their tiles
// [ta, T1) and payload entries [Z0, Z1), precomputed so that a wave starts
// loading payload without first walking the tile pointers.  flags bit 0
// (kDescOneWindow): the tiles fit one window (at most kWinRows tile-rows,
// kWinTiles tiles, 4 * kWinChunks slots).
struct WaveDesc {
  int tr0, tr1, ta, T1, Z0, Z1, flags, aux;
};
constexpr int kDescOneWindow = 1;
// flags bit 1 (kDescPiece): tiles [ta, T1) are piece flags >> 8 of a heavy
// tile-row tr0 split over several waves; aux is the heavy tile-row's index
// (the solvers add the pieces' row sums).
constexpr int kDescPiece = 2;


// Thread-local error message (mf_last_error).
void set_error(const std::string &msg);
mf_status fail(mf_status code, const std::string &msg);

// Reference-identical Tile_create output plus the handle's device layout.
struct DeviceLayout {
  int device = -1;
  int npad = 0;                // padded vector length = 16 * max(tilem, tilen)
  int *tile_ptr = nullptr;     // tilem+1
  int *tile_col = nullptr;     // tilenum
  int *tile_off = nullptr;     // tilenum+1 payload offset of each tile (pay_off)
  int *tile_row = nullptr;     // tilenum (tile-row of each tile)
  uint8_t *tile_nz = nullptr;  // tilenum: entries of the tile - 1
  int npay = 0;                // payload slots: entries padded per tile to 4
  uint8_t *rc = nullptr;       // npay (+pad): (row << 4) | col inside the tile
  uint8_t *rp16 = nullptr;     // 16*tilenum: row starts inside each tile
  int max_tiles_per_row = 0;   // largest tile count of a tile-row
  int spmv_kernel = 0;         // MF_SPMV_STREAM or MF_SPMV_ROWGROUP
  double *v64 = nullptr;       // npay FP64 payload
  float *v32 = nullptr;        // npay FP32 payload
  uint16_t *v16 = nullptr;     // npay FP16 payload (bits), scaled by s16
  int8_t *v8 = nullptr;        // npay INT8 payload, scaled by s8
  float *s16 = nullptr;        // tilenum power-of-two scales
  float *s8 = nullptr;         // tilenum max/127 scales
  double *qerr = nullptr;      // 3*tilenum max |a - dequant_p(a)|, p=FP32,FP16,INT8
  double *amax = nullptr;      // tilenum max |a|
  uint8_t *prec = nullptr;     // tilenum current precision levels
  uint8_t *prec_user = nullptr;// tilenum levels set by mf_tile_set_precision
  int *diag_tile = nullptr;    // tilem: index of the diagonal tile or -1
  // learned policy: MLP in MFMA lane layout (MlpDevice) and per-tile static
  // features (8 FP16 per tile, slots 6-7 filled per inference)
  void *mlp = nullptr;
  uint16_t *feat16 = nullptr;
  int has_mlp = 0;
  // stand-alone SpMV schedule: one descriptor per wave (window_partition),
  // heavy tile-rows cut into one-window pieces whose row sums the last piece
  // to finish adds (16 partial sums per piece; per heavy tile-row a counter,
  // the piece count and the first piece's slot)
  WaveDesc *spmv_desc = nullptr;
  int spmv_nwaves = 0;
  int spmv_wave_nnz = 0;
  double *spmv_part = nullptr;
  unsigned *spmv_cnt = nullptr;
  int *spmv_hn = nullptr;
  int *spmv_hbase = nullptr;
  int spmv_nheavy = 0;
};

struct Schedule {
  // Contiguous tile-row ranges, one per persistent wave, balanced on nnz
  // (the vector phases' row ownership).
  int nwaves = 0;
  int *wave_tr = nullptr;  // device, nwaves+1 tile-row bounds
  // 1 for a wave whose SpMV share (its tile-rows, no heavy one, no pieces)
  // fits one resident LDS window of the persistent solvers
  int *wave_res = nullptr;  // device, nwaves
  int all_resident = 0;     // every non-empty wave's share is resident
  // SpMV work of each wave: win_desc[wave_win[w] .. wave_win[w+1]), the
  // window descriptors of its tile-rows except heavy ones, then the pieces of
  // heavy tile-rows assigned to it.
  int *wave_win = nullptr;       // device, nwaves+1
  WaveDesc *win_desc = nullptr;  // device
  // Heavy tile-rows: 16 FP64 partial row sums per piece, and per heavy
  // tile-row an arrival counter (zero between SpMVs; the last piece resets
  // it), the piece count and the first piece's slot.
  int nheavy = 0;
  double *heavy_part = nullptr;   // device, 16 * pieces
  unsigned *heavy_cnt = nullptr;  // device, nheavy
  int *heavy_n = nullptr;         // device, nheavy
  int *heavy_base = nullptr;      // device, nheavy
  int *heavy_tr = nullptr;        // device, nheavy: tile-row of each heavy tile-row
  int *wave_heavy = nullptr;      // device, nwaves+1: heavy tile-rows each wave owns
};

}  // namespace mf

struct mf_tiled {
  int rowA = 0, colA = 0, tilem = 0, tilen = 0, tilenum = 0;
  int csrsize = 0, csrptrlen = 0;
  std::vector<int> tile_ptr, tile_columnidx, tile_nnz, blknnz, csr_offset,
      csrptr_offset;
  std::vector<unsigned char> blknnznnz, Tile_csr_Col, Blockcsr_Ptr;
  std::vector<double> Blockcsr_Val;
  std::vector<float> Blockcsr_Val_Low;
  // device payload offsets: each tile's entries start at a multiple of 4
  // (entries padded per tile to a multiple of 4), tilenum+1
  std::vector<int> pay_off;
  mf::DeviceLayout dev;
  mf::Schedule sched;
  int sched_waves_for = 0;  // wave count the cached schedule was built for
  // device scratch owned by the handle (solver workspace), grown on demand
  void *work = nullptr;
  size_t work_bytes = 0;
};

namespace mf {
// Host format build (mf_format.cpp).
mf_status build_tiles_host(mf_tiled *t, int rowA, int colA, const int *rowptr,
                           const int *colidx, const double *val,
                           const float *val_low);
// Device upload + payload packing (mf_device.hip).
mf_status upload_tiles(mf_tiled *t, int device);
void free_device(mf_tiled *t);
// nnz-balanced contiguous tile-row partition for `nwaves` waves.
std::vector<int> balance_tilerows(const mf_tiled *t, int nwaves);
// Greedy split into waves of about target_nnz payload slots that each fit
// one SpMV window (<= kWinRows tile-rows, <= kWinTiles tiles, <= kWinChunks
// chunks) unless a single tile-row is larger than a window (such a tile-row
// gets a wave of its own).  Needs pay_off.
// Stand-alone SpMV schedule: descriptors of about target_nnz payload slots,
// except that a tile-row of more than 4 * max(target_nnz, 4 * kWinChunks)
// slots is heavy and cut into one-window pieces (kDescPiece); hn and hbase
// receive each heavy tile-row's piece count and first piece slot.
std::vector<WaveDesc> window_partition(const mf_tiled *t, int target_nnz,
                                       std::vector<int> *hn = nullptr,
                                       std::vector<int> *hbase = nullptr);
// Pieces of tile-row r (heavy tile-row index h) of at most max_tiles tiles and
// max_slots payload slots, appended to out with kDescPiece | extra_flags;
// returns their number.
int cut_pieces(const mf_tiled *t, int r, int h, int max_tiles, int max_slots, int extra_flags,
               std::vector<WaveDesc> &out);
// Tile-rows [r0, r1) cut into descriptors of at most target_nnz payload
// slots, each within one window's limits unless a single tile-row exceeds
// them (such a descriptor walks several windows).
void window_partition_range(const mf_tiled *t, int r0, int r1, int target_nnz,
                            std::vector<WaveDesc> &out);
// Default stand-alone SpMV kernel (MF_SPMV_STREAM or MF_SPMV_ROWGROUP).
int default_spmv_kernel(const mf_tiled *t);
// No tile-row holds more than 4x the mean tile count + 8 (the row-group
// kernel, one tile-row per 16 lanes, wastes little on such matrices).
bool regular_tilerows(const mf_tiled *t);
// Error-bound precision levels into the handle's current levels
// (stream-ordered; stream is a hipStream_t).
mf_status apply_errbound_levels(mf_tiled *t, void *stream);
// The FP32 / FP16 / INT8 levels of each of nwaves waves' tiles (tile-rows
// d_wave_tr[w] .. d_wave_tr[w+1], device) set to the most precise of them.
mf_status wave_uniform_levels(mf_tiled *t, const int *d_wave_tr, int nwaves, void *stream);
}  // namespace mf
