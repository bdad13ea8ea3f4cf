/*
 * mf_oracle.h -- CPU oracle for the tiled SpMV / CG / BiCGSTAB hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and never by the product library (mille-feuille_amd/), which must fail
 * loudly rather than fall back to it.
 *
 * Every function restates one piece of the reference's algorithm in plain C
 * and cites the reference file:line it follows.  Paths are relative to the
 * reference repository root.
 *
 * Pinning (what checks each function against something other than itself):
 *   or_mtx_info/or_mtx_read        golden CSR of test/add20.mtx written by the
 *                                  compiled reference reader (oracle/_ref),
 *                                  plus hand-written .mtx known answers.
 *   or_tile_sizes/or_tile_fill     golden Tile_create arrays from oracle/_ref
 *                                  (add20, 2D Laplacian 64x64), plus
 *                                  reconstruction invariants.
 *   or_blockspmv_offsets           golden ptroffset1/2, rowblkblock and the
 *                                  chunk arrays from oracle/_ref.
 *   or_encode/or_decode            exhaustive round trip on small inputs and
 *                                  the closed form of the nibble layout.
 *   or_blockspmv_nibble            golden y of the reference blockspmv_cpu
 *                                  (its all-zero nibble array) from oracle/_ref.
 *   or_blockspmv_tiled             equals or_csr_spmv (independent CSR loop)
 *                                  on every fixture, bit-exact per tile-row
 *                                  order is not required; tolerance 1e-15 rel.
 *   or_compact_rows                brute-force recount on random tiny tiles and
 *                                  the identity cnt = tilenum + nonempty rows.
 *   or_load_balance                invariants only (permutation of all tiles,
 *                                  monotone bounds, warp limit): the reference
 *                                  balancer lives in .cu host code that cannot
 *                                  be compiled here -> PARITY UNPINNED beyond
 *                                  those invariants.
 *   or_shared_offsets              the closed form (per warp, exclusive prefix
 *                                  sums of its tiles' nnz up to the first
 *                                  tile that overflows 256; that tile ends the
 *                                  prefix) recomputed in numpy on every
 *                                  fixture and on hand-built warps at the 256
 *                                  boundary; same CUDA-host-code limitation
 *                                  as or_load_balance -> PARITY UNPINNED
 *                                  against the reference's own run.
 *   or_cg/or_bicgstab              scipy.sparse.linalg on the same systems
 *                                  (test-time only) and the closed-form answer
 *                                  x = ones for b = A*ones.
 *   or_half_from_double            numpy.float16(numpy.float32(v)) over all
 *                                  special classes and random values.
 *   or_quant_int8                  closed form: scale = smallest 2^e with
 *                                  max|a|/2^e <= 127, q = rint(a/scale).
 *   or_rule_policy                 hand-computed segments at each threshold.
 *   or_mlp_*                       central finite differences of or_mlp_loss.
 *   or_mlp_forward16               equals or_mlp_forward wherever every
 *                                  feature and activation is FP16-representable
 *                                  (small dyadic weights), and differs from it
 *                                  by the FP16 rounding otherwise.
 */
#ifndef MF_ORACLE_H
#define MF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Matrix Market input (src/biio2.0/src/mmio_highlevel.h:7-321) ---- */
int or_mtx_info(const char *path, int *m, int *n, int64_t *nnz, int *is_sym);
int or_mtx_read(const char *path, int64_t *rowptr, int *colidx, double *val);

/* ---- 16x16 tiled CSR (src/csr2block.h:267-422) ---- */
int or_tile_sizes(int rowA, int colA, const int *rowptr, const int *colidx,
                  int *tilem, int *tilen, int *tilenum, int *csrsize,
                  int *csrptrlen);
int or_tile_fill(int rowA, int colA, const int *rowptr, const int *colidx,
                 const double *val, const float *val_low,
                 int *tile_ptr, int *tile_columnidx, int *tile_nnz,
                 int *blknnz, unsigned char *blknnznnz, int *csr_offset,
                 int *csrptr_offset, double *Blockcsr_Val,
                 float *Blockcsr_Val_Low, unsigned char *Tile_csr_Col,
                 unsigned char *Blockcsr_Ptr);

/* ---- blockspmv_cpu offsets / chunking (src/blockspmv_cpu.h:2-112) ---- */
int or_rowblkblock_count(int tilem, const int *tile_ptr);
void or_blockspmv_offsets(int rowA, int tilem, const int *tile_ptr,
                          const int *blknnz, int *ptroffset1, int *ptroffset2,
                          unsigned int *blkcoostylerowidx, int *colstart,
                          int *colstop);

/* ---- nibble codec (src/encode.h:29-72) ---- */
void or_encode(const unsigned char *str, unsigned char *res, int str_len,
               int offset);
void or_decode(const unsigned char *str, unsigned char *res, int res_len);

/* ---- tiled SpMV, FP64 ---- */
void or_blockspmv_tiled(int rowA, int colA, int tilem, int tilen,
                        const int *tile_ptr, const int *tile_columnidx,
                        const int *blknnz, const double *Blockcsr_Val,
                        const unsigned char *Tile_csr_Col,
                        const unsigned char *Blockcsr_Ptr,
                        const double *x, double *y, int nthreads);
void or_blockspmv_nibble(int rowA, int colA, int tilem, int tilen,
                         const int *tile_ptr, const int *tile_columnidx,
                         const int *blknnz, const double *Blockcsr_Val,
                         const unsigned char *csr_compressedIdx,
                         const unsigned char *Blockcsr_Ptr, const double *x,
                         double *y);
void or_csr_spmv(int m, const int *rowptr, const int *colidx,
                 const double *val, const double *x, double *y);

/* ---- non-empty-row compaction (src/main-cg.cu:536-593) ---- */
int or_compact_count(int rowA, int tilem, const int *tile_ptr,
                     const int *blknnz, const unsigned char *Blockcsr_Ptr,
                     int *nonzero_row_new /* tilenum+1 */);
void or_compact_fill(int rowA, int tilem, const int *tile_ptr,
                     const int *blknnz, const unsigned char *Blockcsr_Ptr,
                     const int *nonzero_row_new, unsigned char *blockrowid_new,
                     unsigned char *blockcsr_ptr_new, int *block_signal,
                     int *tile_rowidx);

/* ---- nnz load balancer (src/main-cg.cu:596-740) ---- */
void or_shared_offsets(int tilenum, const int *blknnz, int nwarps,
                       const int *balance_tile_ptr, const int *index_each_block,
                       int *non_each_block_offset, int *shared_end);
int or_load_balance(int tilem, int tilenum, int nnzR, const int *tile_ptr,
                    const int *blknnz, int warp_limit, int *balance_tile_ptr,
                    int *index_each_block, int *row_each_block,
                    int *each_block_nnz_out, int *block_per_warp_out);

/* ---- solvers, FP64 ---- */
int or_cg(int n, const int *rowptr, const int *colidx, const double *val,
          const double *b, double *x, double tol, int maxiter,
          double *relres_out);
int or_bicgstab(int n, const int *rowptr, const int *colidx,
                const double *val, const double *b, double *x, double tol,
                int maxiter, double *relres_out);
double or_rel_residual(int n, const int *rowptr, const int *colidx,
                       const double *val, const double *b, const double *x);

/* ---- precision payloads ---- */
uint16_t or_half_from_float(float f);
float or_float_from_half(uint16_t h);
void or_quant_fp16(const double *v, int n, uint16_t *out);
float or_quant_int8(const double *v, int n, int8_t *out);
void or_rule_policy(int nseg, const double *p_new, const double *p_old,
                    unsigned char *level);

/* ---- learned policy MLP (8 -> 32 -> 32 -> 4), FP64 reference ---- */
void or_mlp_forward(int nrows, const double *feat, const double *w1,
                    const double *b1, const double *w2, const double *b2,
                    const double *w3, const double *b3, double *logits);
/* Same network with the device's FP16 rounding of features and hidden
 * activations (parameters passed already FP16-representable). */
void or_mlp_forward16(int nrows, const double *feat, const double *w1,
                      const double *b1, const double *w2, const double *b2,
                      const double *w3, const double *b3, double *logits);
double or_mlp_loss(int nrows, const double *feat, const int *label,
                   const double *w1, const double *b1, const double *w2,
                   const double *b2, const double *w3, const double *b3);
double or_mlp_grad(int nrows, const double *feat, const int *label,
                   const double *w1, const double *b1, const double *w2,
                   const double *b2, const double *w3, const double *b3,
                   double *gw1, double *gb1, double *gw2, double *gb2,
                   double *gw3, double *gb3);

#ifdef __cplusplus
}
#endif
#endif
