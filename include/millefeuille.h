/*
 * millefeuille.h -- C ABI of libmillefeuille_hip.so, the MI355X (gfx950)
 * tiled SpMV and single-kernel CG / BiCGSTAB library.
 *
 * Conventions
 *   - Every entry point returns an mf_status: 0 = OK, negative = error,
 *     positive = ran but did not converge.  mf_last_error() returns a
 *     thread-local message for the last failure on the calling thread.
 *   - Plain pointers and sizes only.  Pointers named d_* are device pointers
 *     (HBM, allocated by the caller on the handle's device); everything else
 *     is host memory.
 *   - Streams are passed as void* (a hipStream_t; NULL = the null stream).
 *   - One handle may be used by one host thread at a time.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the reference repository root).
 */
#ifndef MILLEFEUILLE_H
#define MILLEFEUILLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int mf_status;

#define MF_OK 0
#define MF_NOT_CONVERGED 1
#define MF_ERR_ARG (-1)
#define MF_ERR_IO (-2)
#define MF_ERR_FORMAT (-3)
#define MF_ERR_HIP (-4)
#define MF_ERR_NOMEM (-5)
#define MF_ERR_UNSUPPORTED (-6)
#define MF_ERR_BREAKDOWN (-7)
#define MF_ERR_COOP (-8)

/* Per-tile precision levels.  Reference flags (src/main-cg-mixed.cu:236-322,
 * d_vis): 0 = FP64, 4 = FP32, 3 = FP16, 2 = INT8, 1 = bypass. */
#define MF_PREC_FP64 0
#define MF_PREC_FP32 1
#define MF_PREC_FP16 2
#define MF_PREC_INT8 3
#define MF_PREC_SKIP 4

/* Host CSR matrix.  Arrays owned by the library when filled by mf_read_mtx
 * (release with mf_csr_free). */
typedef struct mf_csr {
  int m, n;
  int64_t nnz;
  int *rowptr; /* m+1 */
  int *colidx; /* nnz */
  double *val; /* nnz */
  int is_sym;  /* banner said symmetric/hermitian (entries already mirrored) */
} mf_csr;

/* Reads a Matrix Market coordinate file into CSR in file order, mirroring
 * symmetric/hermitian entries.
 * Replaces read_Dmatrix_32 (src/biio2.0/src/biio.h:153-194) with
 * mmio_info/mmio_data (src/biio2.0/src/mmio_highlevel.h:7-321). */
mf_status mf_read_mtx(const char *path, mf_csr *out);
/* Binary CSR cache: int32 rows, int32 cols, int64 nnz, int64 rowptr[rows+1],
 * int32 col[nnz], double val[nnz].
 * Replaces binary_read_csr / binary_write_csr (src/biio2.0/src/biio.h:22-97). */
mf_status mf_read_cbd(const char *path, mf_csr *out);
mf_status mf_write_cbd(const char *path, const mf_csr *a);
void mf_csr_free(mf_csr *a);

/* ---------------------------------------------------------------------- */
/* Tiled matrix handle                                                     */
/* ---------------------------------------------------------------------- */

typedef struct mf_tiled mf_tiled;

/* Builds the 16x16 tiled CSR of rows [0,rowA) and columns [0,colA) of the
 * CSR (rowptr has at least rowA+1 entries; val_low may be NULL, then the FP32
 * copy is (float)val), uploads the device layout to `device` and packs the
 * FP32 / FP16 / INT8 payloads there.  device = -1 builds a host-only handle
 * (the reference arrays and preprocessing, no HIP calls); device entry points
 * reject such a handle with MF_ERR_ARG.
 * Replaces Tile_create (src/csr2block.h:267-422) plus the H2D phase of the
 * solvers (src/main-cg.cu:441-516, src/main-cg-mixed.cu:2299-2362). */
mf_status mf_tile_create(int rowA, int colA, const int *rowptr,
                         const int *colidx, const double *val,
                         const float *val_low, int device, mf_tiled **out);
void mf_tile_free(mf_tiled *t);

/* Shapes: tilem, tilen, tilenum, csrsize (nnz), csrptrlen, npad (length of
 * padded device vectors = 16*max(tilem,tilen)). */
mf_status mf_tile_info(const mf_tiled *t, int *tilem, int *tilen,
                       int *tilenum, int *csrsize, int *csrptrlen, int *npad);

/* Host views of the reference-identical Tile_create arrays (owned by the
 * handle; valid until mf_tile_free).  Sizes: tile_ptr tilem+1,
 * tile_columnidx tilenum, tile_nnz/blknnz/csr_offset/csrptr_offset
 * tilenum+1, blknnznnz tilenum+1, Blockcsr_Val/_Low/Tile_csr_Col csrsize,
 * Blockcsr_Ptr csrptrlen.  Field meanings: src/format.h:19-45. */
typedef struct mf_tile_arrays {
  const int *tile_ptr, *tile_columnidx, *tile_nnz, *blknnz, *csr_offset,
      *csrptr_offset;
  const unsigned char *blknnznnz, *Tile_csr_Col, *Blockcsr_Ptr;
  const double *Blockcsr_Val;
  const float *Blockcsr_Val_Low;
} mf_tile_arrays;
mf_status mf_tile_arrays_get(const mf_tiled *t, mf_tile_arrays *out);

/* Reference-derived host arrays, for callers that drive the reference's own
 * kernels or its CPU check loop.  All output arrays are caller-allocated.
 *   mf_tile_offsets: ptroffset1/ptroffset2 (tilenum each) and the chunked
 *     row-block list (rowblkblock entries, see mf_tile_rowblkblock);
 *     replaces blockspmv_cpu's preprocessing (src/blockspmv_cpu.h:2-83).
 *   mf_tile_compact: nonzero_row_new (tilenum+1), then blockrowid_new /
 *     blockcsr_ptr_new (cnt+1), block_signal (tilem+1), tile_rowidx
 *     (tilenum); replaces src/main-cg.cu:536-593.
 *   mf_tile_balance: the reference's two-pointer nnz balancer, returns the
 *     warp count in *nwarps; balance_tile_ptr (nwarps+1, sized tilenum+2 is
 *     always enough), index_each_block / row_each_block (tilenum+1);
 *     replaces src/main-cg.cu:596-761.
 *   mf_tile_shared_offsets: from mf_tile_balance's output, per warp the
 *     prefix of its tiles whose nnz fit a 256-entry LDS queue:
 *     non_each_block_offset (tilenum+1; offset of each such tile inside the
 *     queue, 0 for the others, which the reference leaves unset) and
 *     balance_tile_ptr_shared_end (nwarps+1; end of the prefix, entry w+1
 *     for warp w, the warp's own end when everything fits); replaces
 *     src/main-cg-mixed.cu:2681-2708. */
mf_status mf_tile_rowblkblock(const mf_tiled *t, int *rowblkblock);
mf_status mf_tile_offsets(const mf_tiled *t, int *ptroffset1, int *ptroffset2,
                          unsigned int *blkcoostylerowidx, int *colstart,
                          int *colstop);
mf_status mf_tile_compact_count(const mf_tiled *t, int *nonzero_row_new,
                                int *cnt_non_new);
mf_status mf_tile_compact_fill(const mf_tiled *t, const int *nonzero_row_new,
                               unsigned char *blockrowid_new,
                               unsigned char *blockcsr_ptr_new,
                               int *block_signal, int *tile_rowidx);
mf_status mf_tile_balance(const mf_tiled *t, int warp_limit,
                          int *balance_tile_ptr, int *index_each_block,
                          int *row_each_block, int *nwarps);
mf_status mf_tile_shared_offsets(const mf_tiled *t, int nwarps,
                                 const int *balance_tile_ptr,
                                 const int *index_each_block,
                                 int *non_each_block_offset,
                                 int *balance_tile_ptr_shared_end);

/* Nibble codec for tile-local column indices, two per byte, even position in
 * the high nibble (src/encode.h:29-72). */
void mf_encode_nibbles(const unsigned char *cols, unsigned char *packed,
                       int n, int offset);
void mf_decode_nibbles(const unsigned char *packed, unsigned char *cols, int n);

/* ---------------------------------------------------------------------- */
/* Per-tile precision                                                      */
/* ---------------------------------------------------------------------- */

/* Sets every tile's precision level (MF_PREC_*) from a host array of tilenum
 * bytes, or all tiles to `uniform` when levels is NULL. */
mf_status mf_tile_set_precision(mf_tiled *t, const unsigned char *levels,
                                int uniform);
mf_status mf_tile_get_precision(const mf_tiled *t, unsigned char *levels);
/* Per-tile payload scales (tilenum floats each; FP16 power-of-two scale and
 * INT8 max/127 scale) and per-tile max |a - dequant(a)| for FP32/FP16/INT8
 * (3*tilenum doubles, level-major).  Host copies. */
mf_status mf_tile_get_scales(const mf_tiled *t, float *scale16, float *scale8);
mf_status mf_tile_get_quant_error(const mf_tiled *t, double *err);

/* ---------------------------------------------------------------------- */
/* SpMV                                                                    */
/* ---------------------------------------------------------------------- */

/* y = A x on the device.  d_x holds npad doubles (zero beyond colA), d_y
 * receives npad doubles (rows beyond rowA are zero).  mode: 0 = all tiles
 * FP64; 1 = per-tile precision levels set on the handle, low-precision tiles
 * multiply the FP32 copy of x (d_x32, npad floats; required for mode 1) and
 * accumulate in FP32 per tile.
 * Replaces stir_spmv_cuda_kernel_newcsr (src/Mille-feuille-cg.h:97-166) and
 * the tile loop of blockspmv_cpu (src/blockspmv_cpu.h:85-112). */
mf_status mf_spmv(mf_tiled *t, const double *d_x, const float *d_x32,
                  double *d_y, int mode, void *stream);

/* Stand-alone SpMV kernel of a handle: MF_SPMV_STREAM (nnz-balanced tile
 * ranges, LDS row accumulators; any sparsity; the default), MF_SPMV_ROWGROUP
 * (one row per lane, four tile-rows per wave; regular sparsity) or
 * MF_SPMV_AUTO (the default).
 * No reference counterpart (one kernel form there, src/Mille-feuille-cg.h:97-166). */
#define MF_SPMV_AUTO (-1)
#define MF_SPMV_STREAM 0
#define MF_SPMV_ROWGROUP 1
mf_status mf_tile_set_spmv_kernel(mf_tiled *t, int kind);

/* Schedule knob for the autotuner: entries per wave of the stand-alone SpMV
 * (its nnz-balanced tile-row ranges are rebuilt; 0 keeps the current value).
 * No reference counterpart (the reference hardcodes its chunking,
 * src/blockspmv_cpu.h:21-31, and its warp budget, src/main-cg.cu:634). */
mf_status mf_tile_set_schedule(mf_tiled *t, int spmv_wave_nnz);

/* dst[i] = src[i] for n doubles with a plain streaming kernel.  Bandwidth
 * calibration for the SpMV roofline (the HBM byte counters of a copy of known
 * size give the scale factor for the SpMV's counters).  No reference
 * counterpart. */
mf_status mf_stream_copy(const double *d_src, double *d_dst, int64_t n,
                         void *stream);

/* ---------------------------------------------------------------------- */
/* Solvers (single persistent kernel per solve)                            */
/* ---------------------------------------------------------------------- */

/* Precision policies for the mixed solvers. */
#define MF_POLICY_FP64 0    /* every tile FP64: the plain single-kernel CG */
#define MF_POLICY_RULE 1    /* reference rule on |p_new - p_old| per segment */
#define MF_POLICY_FIXED 2   /* use the levels set with mf_tile_set_precision */
#define MF_POLICY_LEARNED 3 /* MLP over per-tile features (mf_policy_*); the
                               FP32 / FP16 / INT8 classes of a solver wave's
                               tiles run at the most precise of them */
#define MF_POLICY_ERRBOUND 4 /* analytic per-tile quantisation error bound;
                                the FP32 / FP16 / INT8 levels of a solver
                                wave's tiles run at the most precise of them */

typedef struct mf_solve_opts {
  double tol;           /* stop at ||b - A x||_2 <= tol * ||b||_2 (FP64) */
  int maxiter;          /* iteration cap (0 -> 10000) */
  int policy;           /* MF_POLICY_* */
  int check_interval;   /* iterations between policy evaluations (0 -> 8) */
  double refresh_ratio; /* recompute r = b - A x in FP64 when the recursive
                           residual fell by this factor (0 -> 0.1) */
  int max_refresh;      /* refinement steps before FP64 fallback (0 -> 64) */
  int refresh_every;    /* mixed policies: FP64 residual check at least every
                           this many iterations; no progress since the last
                           check -> FP64 fallback and restart (0 -> 64) */
  int blocks;           /* persistent grid size override (0: about 1024 rows
                           per block, at least 4 blocks, at most one block of
                           16 waves per CU) */
  int fixed_iters;      /* >0: run exactly this many iterations, no test
                           (reference timing mode, src/Mille-feuille-cg.h:426) */
  int form;             /* solver SpMV form: MF_FORM_MEMORY (every SpMV reads
                           the tile metadata of each wave's windows from HBM),
                           MF_FORM_RESIDENT or MF_FORM_AUTO (each wave keeps
                           its share's tile metadata in LDS for the whole
                           solve when every share fits a window of 40
                           tile-rows, 256 tiles and 1024 chunks with no heavy
                           tile-row; else as MEMORY).  The report's form_used
                           says which ran. */
  int batch_waves;      /* mf_batch_cg_solve: waves per (system, right-hand
                           side) item, 4, 8 or 16 (0 -> 8; mf_tune_batch
                           picks one) */
  void *stream;         /* hipStream_t */
} mf_solve_opts;

#define MF_FORM_AUTO 0
#define MF_FORM_MEMORY 1
#define MF_FORM_RESIDENT 2

typedef struct mf_report {
  int iterations;       /* CG / BiCGSTAB iterations executed */
  int refreshes;        /* FP64 true-residual recomputations */
  int fell_back;        /* 1 if the solver promoted every tile to FP64 */
  int status;           /* same as the return value */
  double rel_residual;  /* device FP64 ||b - A x|| / ||b|| at exit */
  double rel_residual_recursive; /* recursive residual at exit */
  double time_ms;       /* kernel time (HIP events) */
  long long prec_hist[5]; /* tile-iterations per precision level */
  int grid_blocks;      /* persistent grid actually launched */
  int form_used;        /* MF_FORM_MEMORY or MF_FORM_RESIDENT */
  int policy_used;      /* MF_POLICY_* that ran (LEARNED without an installed
                           MLP runs the rule policy) */
} mf_report;

void mf_solve_opts_default(mf_solve_opts *o);

/* Solves A x = b by CG in one persistent cooperative kernel.  d_b, d_x are
 * npad doubles on the device; d_x holds the initial guess on entry (zero
 * padding beyond rowA) and the solution on exit.
 * Replaces cg_solve_sync / cg_solve_inc (src/main-cg.cu:379, 1071) and the
 * mixed cg_solve_inc (src/main-cg-mixed.cu:2238) with their persistent
 * kernels (src/Mille-feuille-cg.h:352-580, src/main-cg-mixed.cu:998-1377). */
mf_status mf_cg_solve(mf_tiled *t, const double *d_b, double *d_x,
                      const mf_solve_opts *opts, mf_report *rep);

/* mf_cg_solve without any device allocation.  No counterpart in the
 * reference: this is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * work vectors, grid-barrier state and the solver schedule live in one
 * library-owned device array of 12 MiB per device (held until the process
 * ends), which limits the system to about 150,000 rows; larger systems are
 * rejected with MF_ERR_ARG.  The caller passes d_b and d_x as for
 * mf_cg_solve.  Every call on a device shares that array, so calls that
 * overlap on one device race; the report's time_ms is 0 (no events are
 * created). */
mf_status mf_cg_solve_noalloc(mf_tiled *t, const double *d_b, double *d_x,
                              const mf_solve_opts *opts, mf_report *rep);

/* BiCGSTAB in one persistent cooperative kernel, same conventions.
 * Replaces bicgstab_sync / bicgstab_inc (src/main-bicg.cu:13, 711), the
 * mixed bicgstab (src/main-bicg-mixed.cu:1355) and their kernels
 * (src/Mille-feuille-bicg.h:202-1185, src/main-bicg-mixed.cu:122-1353). */
mf_status mf_bicgstab_solve(mf_tiled *t, const double *d_b, double *d_x,
                            const mf_solve_opts *opts, mf_report *rep);

/* Multi-kernel CG / BiCGSTAB: one launch per step (SpMV, dot, axpy), the
 * host reads the residual each iteration.
 * Replaces cg_solve_reduce (src/main-cg.cu:6-376) and bicgstab
 * (src/main-bicg.cu:1336-1695). */
mf_status mf_cg_solve_multikernel(mf_tiled *t, const double *d_b, double *d_x,
                                  const mf_solve_opts *opts, mf_report *rep);
mf_status mf_bicgstab_solve_multikernel(mf_tiled *t, const double *d_b,
                                        double *d_x, const mf_solve_opts *opts,
                                        mf_report *rep);

/* ---------------------------------------------------------------------- */
/* Reference-compatible entry points (host buffers in and out)            */
/* ---------------------------------------------------------------------- */

/* Same argument lists and semantics as the reference: rows truncated to a
 * multiple of 16 (n is the truncated count, ori the original column count),
 * x is overwritten with the solution, iter/maxiter/threshold are ignored as
 * in the reference, a CSV row is appended to data/cg_performance.csv.
 * Unlike the reference these iterate to the FP64 relative residual 1e-10
 * (set MF_COMPAT_FIXED_ITERS=100 in the environment for the reference's fixed
 * 100 iterations).  src/main-cg.cu:6, 379, 1071; src/main-cg-mixed.cu:2238. */
void cg_solve_reduce(int *RowPtr, int *ColIdx, double *Val, float *Val_Low,
                     double *x, double *b, int n, int *iter, int maxiter,
                     double threshold, char *filename, int nnzR, int ori);
void cg_solve_sync(int *RowPtr, int *ColIdx, double *Val, float *Val_Low,
                   double *x, double *b, int n, int *iter, int maxiter,
                   double threshold, char *filename, int nnzR, int ori);
void cg_solve_inc(int *RowPtr, int *ColIdx, double *Val, float *Val_Low,
                  double *x, double *b, int n, int *iter, int maxiter,
                  double threshold, char *filename, int nnzR, int ori);
void cg_solve_mixed(int *RowPtr, int *ColIdx, double *Val, float *Val_Low,
                    double *x, double *b, int n, int *iter, int maxiter,
                    double threshold, char *filename, int nnzR, int ori);
/* rhs is overwritten with the row sums of A (A*ones) and x with the
 * solution; a CSV row goes to data/bicg_performance.csv.
 * src/main-bicg.cu:13, 711, 1336; src/main-bicg-mixed.cu:1355. */
int bicgstab(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x,
             int nnzR, char *filename, int n);
int bicgstab_sync(int *RowPtr, int *ColIdx, double *Val, double *rhs,
                  double *x, int nnzR, char *filename, int n);
int bicgstab_inc(int *RowPtr, int *ColIdx, double *Val, double *rhs,
                 double *x, int nnzR, char *filename, int n);
int bicgstab_mixed(int *RowPtr, int *ColIdx, double *Val, double *rhs,
                   double *x, int nnzR, char *filename, int n);
/* src/biio2.0/src/biio.h:153 */
void read_Dmatrix_32(int *row, int *col, int *nnz, int **row_ptr,
                     int **col_idx, double **val, int *isSymmeticeR,
                     char *filename);

/* ---------------------------------------------------------------------- */
/* Learned tile-precision policy (no reference counterpart: the reference */
/* has only the rule of src/main-cg-mixed.cu:1316-1364)                    */
/* ---------------------------------------------------------------------- */

#define MF_MLP_IN 8
#define MF_MLP_H1 32
#define MF_MLP_H2 32
#define MF_MLP_OUT 4
#define MF_MLP_NPARAM                                                     \
  (MF_MLP_H1 * MF_MLP_IN + MF_MLP_H1 + MF_MLP_H2 * MF_MLP_H1 + MF_MLP_H2 + \
   MF_MLP_OUT * MF_MLP_H2 + MF_MLP_OUT)

/* MLP 8->32->32->4 with ReLU; parameters packed as w1[32][8], b1[32],
 * w2[32][32], b2[32], w3[4][32], b3[4] (1476 values).  The device copy is
 * FP16 with FP32 accumulation; training keeps an FP32 master copy.
 *   mf_mlp_forward: logits (nrows x 4 floats) for features (nrows x 8).
 *   mf_mlp_grad:    mean softmax cross-entropy loss and its gradient with
 *                   respect to the FP32 parameters (all device pointers).
 *   mf_mlp_train:   epochs of minibatch Adam on the device; returns the
 *                   final loss. */
mf_status mf_mlp_forward(const float *d_params, const float *d_feat, int nrows,
                         float *d_logits, void *stream);
mf_status mf_mlp_grad(const float *d_params, const float *d_feat,
                      const int *d_label, int nrows, float *d_grad,
                      float *d_loss, void *stream);
mf_status mf_mlp_train(float *d_params, const float *d_feat,
                       const int *d_label, int nrows, int epochs, int batch,
                       float lr, unsigned seed, float *loss_out, void *stream);
/* Installs trained parameters (device FP32, converted to FP16 inside) as the
 * handle's learned policy. */
mf_status mf_policy_set_mlp(mf_tiled *t, const float *d_params);
/* Per-tile features the policy sees (tilenum x 8 floats, device) for the
 * vector d_p the SpMV multiplies (npad doubles) at relative residual
 * res_ratio = ||r|| / ||b||: nnz, value range, FP16 and INT8 payload error,
 * weight of the tile in its tile-row, diagonal dominance of the tile-row,
 * local norm of the segment of d_p the tile reads, iteration progress. */
mf_status mf_policy_features(mf_tiled *t, const double *d_p, double res_ratio,
                             float *d_feat, void *stream);
/* Training labels (tilenum ints, device): for each tile the cheapest level
 * (MF_PREC_INT8, FP16, FP32, else FP64) whose product error on d_p, computed
 * as the SpMV computes it, is <= theta * ||A d_p||_2 / sqrt(tilenum). */
mf_status mf_policy_labels(mf_tiled *t, const double *d_p, double theta,
                           int *d_label, void *stream);

/* ---------------------------------------------------------------------- */
/* Batched solves                                                          */
/* ---------------------------------------------------------------------- */

/* Solves many independent SPD systems, each with nrhs right-hand sides, by
 * CG in one persistent launch captured in a hipGraph: every block takes
 * (system, right-hand side) items, largest system first, from a global queue
 * and solves each with block-local reductions.  Tiles run at the handle's
 * levels (opts->policy FP64, FIXED or ERRBOUND) with FP64 scalars, FP64 x and
 * FP64 residual refresh / fallback as mf_cg_solve.  systems[i] are square
 * tiled handles on one device; d_b[i] / d_x[i] point to npad_i * nrhs
 * doubles (right-hand side j at offset j * npad_i, zero padded; d_x holds the
 * initial guesses).  iterations and rel_residual (optional, host) receive
 * nsys * nrhs entries (system-major); per_system_ms (optional, nsys entries)
 * the time from a system's first item start to its last item end; total_ms
 * the launch time.  Returns MF_NOT_CONVERGED if any item missed the
 * tolerance.
 * No reference counterpart (the reference solves one system per process). */
mf_status mf_batch_cg_solve(mf_tiled **systems, int nsys, int nrhs,
                            const double **d_b, double **d_x,
                            const mf_solve_opts *opts, int *iterations,
                            double *rel_residual, double *per_system_ms,
                            double *total_ms);

/* ---------------------------------------------------------------------- */
/* Autotuner                                                               */
/* ---------------------------------------------------------------------- */

/* Candidate schedules timed on the device; the fastest is kept.  cand[i] is
 * a candidate's value and us[i] its median time in microseconds (SpMV: one
 * launch; batch: the whole timed batch run). */
typedef struct mf_tune_report {
  int ncand;
  int cand[8];
  float us[8];
  int best;             /* chosen value */
} mf_tune_report;

/* Stand-alone SpMV schedule of a handle for mf_spmv `mode` (0 or 1, the
 * handle's current levels): the streamed kernel at 384, 512, 768, 1024, 1536
 * and 2048 entries per wave and, on matrices with regular tile-rows, the
 * row-group kernel (value 0).  The fastest is applied with
 * mf_tile_set_schedule / mf_tile_set_spmv_kernel.  Allocates and frees its
 * own x, x32 and y.
 * No reference counterpart (the reference's warp budget is fixed,
 * src/main-cg.cu:634-675). */
mf_status mf_tune_spmv(mf_tiled *t, int mode, mf_tune_report *rep);

/* Waves per (system, right-hand side) item of mf_batch_cg_solve: 4, 8 and
 * 16 are timed on `iters` CG iterations of the given batch (on copies of the
 * initial guesses d_x, which are left unchanged); rep->best is the value for
 * opts->batch_waves.
 * No reference counterpart. */
mf_status mf_tune_batch(mf_tiled **systems, int nsys, int nrhs, const double **d_b,
                        double **d_x, const mf_solve_opts *opts, int iters,
                        mf_tune_report *rep);

const char *mf_last_error(void);
const char *mf_version(void);

#ifdef __cplusplus
}
#endif
#endif
