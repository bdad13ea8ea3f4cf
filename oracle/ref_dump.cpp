// ref_dump.cpp -- builds the reference's own host-side format code (its
// Matrix Market reader, Tile_create and blockspmv_cpu, compiled from the
// headers where they lie under the reference tree) into a fixture dumper.
//
// TEST INFRASTRUCTURE ONLY.  The binary lands in oracle/_ref/ and is run in
// the build container by tests/golden/gen_golden.py to write golden vectors.
// Nothing on the product path or on the GPU runs it.
//
// Usage: ref_dump <matrix.mtx> <outdir> [notrunc]
// With "notrunc" Tile_create gets the full row count (a short last tile-row).
// Mirrors the driver preamble of src/main-cg.cu:1630-1666 (read, truncate the
// row count to a multiple of 16, FP32 copy of the values) and the preamble of
// cg_solve_sync (src/main-cg.cu:379-424: Tile_create, then blockspmv_cpu).
// blockspmv_cpu is called with a non-trivial x so that its (nibble-decoded)
// product is a fixture too.
// The reference's tile load balancer (src/main-cg.cu:596-769) and its
// shared-queue offsets (src/main-cg-mixed.cu:2681-2709) are plain host C
// inside .cu files; ref_extract.sh copies those line ranges into _ref/ at
// build time and ref_balance() below compiles them as they are.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "biio.h"
#include "csr2block.h"
#include "blockspmv_cpu.h"

static void dump(const std::string &dir, const char *name, const void *p,
                 size_t bytes) {
  std::string path = dir + "/" + name + ".bin";
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) { std::perror(path.c_str()); std::exit(2); }
  if (bytes) std::fwrite(p, 1, bytes, f);
  std::fclose(f);
}

// Runs the reference's balancer and shared-queue offset loops on a tiled
// matrix and writes their outputs; `warps` (the reference's `index`) stays
// -1 when the balancer gives up (it returns when index > 6912).
static void ref_balance(Tile_matrix *matrix, int nnzR, const std::string &out) {
  int tilem = matrix->tilem;
  int tilenum = matrix->tilenum;
  MAT_PTR_TYPE *tile_ptr = matrix->tile_ptr;
  int warps = -1;
  std::string meta = out + "/balance_meta.txt";
  FILE *f = std::fopen(meta.c_str(), "w");
  std::fprintf(f, "warps %d\n", warps);
  std::fclose(f);
  {
#include "ref_balance.inc"
#include "ref_shared.inc"
    warps = index;
    dump(out, "balance_tile_ptr_new", balance_tile_ptr_new, sizeof(int) * (size_t)(index + 1));
    dump(out, "index_each_block_new", index_each_block_new, sizeof(int) * (size_t)(tilenum + 1));
    dump(out, "row_each_block_new", row_each_block_new, sizeof(int) * (size_t)(tilenum + 1));
    dump(out, "non_each_block_new", non_each_block_new, sizeof(int) * (size_t)(tilenum + 1));
    dump(out, "non_each_block_offset", non_each_block_offset, sizeof(int) * (size_t)(tilenum + 1));
    dump(out, "balance_tile_ptr_shared_end", balance_tile_ptr_shared_end,
         sizeof(int) * (size_t)(index + 1));
  }
  f = std::fopen(meta.c_str(), "w");
  std::fprintf(f, "warps %d\n", warps);
  std::fclose(f);
}

int main(int argc, char **argv) {
  if (argc != 3 && argc != 4) { std::fprintf(stderr, "usage: %s A.mtx outdir [notrunc]\n", argv[0]); return 1; }
  char *filename = argv[1];
  std::string out = argv[2];
  int m, n, nnzR, isSym;
  int *RowPtr, *ColIdx;
  double *Val;
  read_Dmatrix_32(&m, &n, &nnzR, &RowPtr, &ColIdx, &Val, &isSym, filename);
  dump(out, "csr_rowptr", RowPtr, sizeof(int) * (size_t)(m + 1));
  dump(out, "csr_colidx", ColIdx, sizeof(int) * (size_t)nnzR);
  dump(out, "csr_val", Val, sizeof(double) * (size_t)nnzR);

  float *Val_Low = (float *)std::malloc(sizeof(float) * (size_t)nnzR);
  for (int i = 0; i < nnzR; ++i) Val_Low[i] = (float)Val[i];
  int ori = n;
  int rowA = argc == 4 ? m : (n / BLOCK_SIZE) * BLOCK_SIZE;
  int colA = ori;

  Tile_matrix *mat = (Tile_matrix *)std::calloc(1, sizeof(Tile_matrix));
  Tile_create(mat, rowA, colA, nnzR, RowPtr, ColIdx, Val, Val_Low);
  int tn = mat->tilenum;
  dump(out, "tile_ptr", mat->tile_ptr, sizeof(int) * (size_t)(mat->tilem + 1));
  dump(out, "tile_columnidx", mat->tile_columnidx, sizeof(int) * (size_t)tn);
  dump(out, "tile_nnz", mat->tile_nnz, sizeof(int) * (size_t)(tn + 1));
  dump(out, "blknnz", mat->blknnz, sizeof(int) * (size_t)(tn + 1));
  dump(out, "blknnznnz", mat->blknnznnz, (size_t)(tn + 1));
  dump(out, "csr_offset", mat->csr_offset, sizeof(int) * (size_t)(tn + 1));
  dump(out, "csrptr_offset", mat->csrptr_offset, sizeof(int) * (size_t)(tn + 1));
  dump(out, "Blockcsr_Val", mat->Blockcsr_Val, sizeof(double) * (size_t)mat->csrsize);
  dump(out, "Blockcsr_Val_Low", mat->Blockcsr_Val_Low, sizeof(float) * (size_t)mat->csrsize);
  dump(out, "Tile_csr_Col", mat->Tile_csr_Col, (size_t)mat->csrsize);
  dump(out, "Blockcsr_Ptr", mat->Blockcsr_Ptr, (size_t)mat->csrptrlen);
  int csize = mat->csrsize % 2 == 0 ? mat->csrsize / 2 : mat->csrsize / 2 + 1;
  dump(out, "csr_compressedIdx", mat->csr_compressedIdx, (size_t)csize);

  int *p1 = (int *)std::calloc((size_t)tn + 1, sizeof(int));
  int *p2 = (int *)std::calloc((size_t)tn + 1, sizeof(int));
  int xn = mat->tilen * BLOCK_SIZE;
  if (xn < rowA) xn = rowA;
  double *x = (double *)std::calloc((size_t)xn, sizeof(double));
  for (int i = 0; i < xn; ++i) x[i] = 1.0 + 0.25 * (double)(i % 7) - 0.125 * (double)(i % 3);
  double *y = (double *)std::calloc((size_t)xn, sizeof(double));
  double *yg = (double *)std::calloc((size_t)xn, sizeof(double));
  int rbb = 0;
  unsigned int *rowidx;
  int *cs, *ce;
  blockspmv_cpu(mat, p1, p2, &rbb, &rowidx, &cs, &ce, rowA, colA, nnzR, RowPtr,
                ColIdx, Val, x, y, yg);
  dump(out, "ptroffset1", p1, sizeof(int) * (size_t)tn);
  dump(out, "ptroffset2", p2, sizeof(int) * (size_t)tn);
  dump(out, "blkcoostylerowidx", rowidx, sizeof(unsigned int) * (size_t)rbb);
  dump(out, "blkcoostylerowidx_colstart", cs, sizeof(int) * (size_t)rbb);
  dump(out, "blkcoostylerowidx_colstop", ce, sizeof(int) * (size_t)rbb);
  dump(out, "spmv_x", x, sizeof(double) * (size_t)xn);
  dump(out, "spmv_y_nibble", y, sizeof(double) * (size_t)rowA);

  if (tn > 1) ref_balance(mat, nnzR, out);

  std::string meta = out + "/meta.txt";
  FILE *f = std::fopen(meta.c_str(), "w");
  std::fprintf(f, "m %d\nn %d\nnnz %d\nis_sym %d\nrowA %d\ncolA %d\n", m, n, nnzR, isSym, rowA, colA);
  std::fprintf(f, "tilem %d\ntilen %d\ntilenum %d\ncsrsize %d\ncsrptrlen %d\nrowblkblock %d\n",
               mat->tilem, mat->tilen, tn, mat->csrsize, mat->csrptrlen, rbb);
  std::fclose(f);
  return 0;
}
