// mf_format.cpp -- host side of the tiled format: Matrix Market / binary CSR
// input, the 16x16 tile build, the reference-derived preprocessing arrays and
// the nnz-balanced tile-row partition used by the persistent kernels.
//
// The tile build produces exactly the arrays of the reference's Tile_create
// (src/csr2block.h:267-422) but is organised differently: one pass counts the
// distinct tile columns of every tile-row, one pass places each entry with a
// per-thread column->slot map (O(nnz) instead of the reference's per-entry
// scan over the tile-row's tiles, src/csr2block.h:204-221), both parallel
// over tile-rows.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <omp.h>

#include "mf_internal.h"

namespace mf {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

mf_status fail(mf_status code, const std::string &msg) {
  g_err = msg;
  return code;
}

static int ceil16(int v) { return v % kTile == 0 ? v / kTile : v / kTile + 1; }

mf_status build_tiles_host(mf_tiled *t, int rowA, int colA, const int *rowptr,
                           const int *colidx, const double *val,
                           const float *val_low) {
  if (rowA < 0 || colA < 0 || !rowptr || (!colidx && rowA > 0) || (!val && rowA > 0))
    return fail(MF_ERR_ARG, "mf_tile_create: bad arguments");
  const int base = rowptr[0];
  for (int i = 0; i < rowA; ++i)
    if (rowptr[i + 1] < rowptr[i])
      return fail(MF_ERR_FORMAT, "mf_tile_create: row pointers decrease at row " + std::to_string(i));
  const int nnz = rowptr[rowA] - base;
  {
    int bad = -1;
#pragma omp parallel for reduction(max : bad)
    for (int k = 0; k < nnz; ++k) {
      int c = colidx[base + k];
      if (c < 0 || c >= colA) bad = std::max(bad, k);
    }
    if (bad >= 0)
      return fail(MF_ERR_FORMAT, "mf_tile_create: column index out of range at entry " + std::to_string(bad));
  }
  const int tm = ceil16(rowA), tn = ceil16(colA);
  t->rowA = rowA;
  t->colA = colA;
  t->tilem = tm;
  t->tilen = tn;
  t->csrsize = nnz;

  // Pass 1: distinct tile columns per tile-row.
  std::vector<int> ntiles(tm, 0);
#pragma omp parallel
  {
    std::vector<int> mark(tn > 0 ? tn : 1, -1);
#pragma omp for schedule(dynamic, 64)
    for (int bi = 0; bi < tm; ++bi) {
      int r0 = bi * kTile, r1 = std::min(rowA, r0 + kTile);
      int c = 0;
      for (int k = rowptr[r0]; k < rowptr[r1]; ++k) {
        int jc = colidx[k] / kTile;
        if (mark[jc] != bi) { mark[jc] = bi; ++c; }
      }
      ntiles[bi] = c;
    }
  }
  t->tile_ptr.assign(tm + 1, 0);
  for (int bi = 0; bi < tm; ++bi) t->tile_ptr[bi + 1] = t->tile_ptr[bi] + ntiles[bi];
  const int T = t->tile_ptr[tm];
  t->tilenum = T;
  const int hlast = tm ? rowA - (tm - 1) * kTile : 0;
  t->csrptrlen = tm ? (t->tile_ptr[tm - 1] * kTile + ntiles[tm - 1] * hlast) : 0;

  t->tile_columnidx.assign(T, 0);
  t->tile_nnz.assign(T + 1, 0);
  t->blknnz.assign(T + 1, 0);
  t->csr_offset.assign(T + 1, 0);
  t->csrptr_offset.assign(T + 1, 0);
  t->blknnznnz.assign(T + 1, 0);
  t->Blockcsr_Val.assign(nnz, 0.0);
  t->Blockcsr_Val_Low.assign(nnz, 0.0f);
  t->Tile_csr_Col.assign(nnz, 0);
  t->Blockcsr_Ptr.assign(t->csrptrlen, 0);

  // Pass 2: per tile-row, ascending tile columns, per-tile counts, scatter.
#pragma omp parallel
  {
    std::vector<int> slot(tn > 0 ? tn : 1, -1);
    std::vector<int> cols, cnt, fill;
    std::vector<unsigned char> rowc;
#pragma omp for schedule(dynamic, 64)
    for (int bi = 0; bi < tm; ++bi) {
      const int r0 = bi * kTile, r1 = std::min(rowA, r0 + kTile), h = r1 - r0;
      const int t0 = t->tile_ptr[bi], nt = ntiles[bi];
      cols.clear();
      for (int k = rowptr[r0]; k < rowptr[r1]; ++k) {
        int jc = colidx[k] / kTile;
        if (slot[jc] < 0) { slot[jc] = 0; cols.push_back(jc); }
      }
      std::sort(cols.begin(), cols.end());
      cnt.assign(nt, 0);
      rowc.assign((size_t)nt * kTile, 0);
      for (int a = 0; a < nt; ++a) slot[cols[a]] = a;
      for (int r = 0; r < h; ++r)
        for (int k = rowptr[r0 + r]; k < rowptr[r0 + r + 1]; ++k) {
          int a = slot[colidx[k] / kTile];
          ++cnt[a];
          ++rowc[(size_t)a * kTile + r];
        }
      fill.assign(nt, 0);
      int acc = rowptr[r0] - base;
      for (int a = 0; a < nt; ++a) {
        const int tt = t0 + a;
        t->tile_columnidx[tt] = cols[a];
        t->tile_nnz[tt] = acc;
        t->blknnz[tt] = acc;
        t->csr_offset[tt] = acc;
        t->blknnznnz[tt] = (unsigned char)cnt[a];
        const int po = t0 * kTile + a * h;
        t->csrptr_offset[tt] = po;
        unsigned char run = 0;
        for (int r = 0; r < h; ++r) {
          t->Blockcsr_Ptr[po + r] = run;
          run = (unsigned char)(run + rowc[(size_t)a * kTile + r]);
        }
        fill[a] = acc;
        acc += cnt[a];
      }
      for (int r = 0; r < h; ++r)
        for (int k = rowptr[r0 + r]; k < rowptr[r0 + r + 1]; ++k) {
          const int a = slot[colidx[k] / kTile];
          const int o = fill[a]++;
          t->Blockcsr_Val[o] = val[k];
          t->Blockcsr_Val_Low[o] = val_low ? val_low[k] : (float)val[k];
          t->Tile_csr_Col[o] = (unsigned char)(colidx[k] - cols[a] * kTile);
        }
      for (int a = 0; a < nt; ++a) slot[cols[a]] = -1;
    }
  }
  t->tile_nnz[T] = nnz;
  t->blknnz[T] = nnz;
  t->csr_offset[T] = nnz;
  t->csrptr_offset[T] = t->csrptrlen;
  t->blknnznnz[T] = 0;
  return MF_OK;
}

// Contiguous tile-row ranges for `nwaves` waves with about equal cost,
// cost(tile-row) = nnz + 8*tiles + 16 (payload, per-tile metadata, rows).
std::vector<int> balance_tilerows(const mf_tiled *t, int nwaves) {
  const int tm = t->tilem;
  std::vector<double> pre(tm + 1, 0.0);
  for (int bi = 0; bi < tm; ++bi) {
    int t0 = t->tile_ptr[bi], t1 = t->tile_ptr[bi + 1];
    double c = (double)(t->csr_offset[t1] - t->csr_offset[t0]) + 8.0 * (t1 - t0) + 16.0;
    pre[bi + 1] = pre[bi] + c;
  }
  std::vector<int> b(nwaves + 1, tm);
  b[0] = 0;
  const double total = pre[tm];
  int bi = 0;
  for (int w = 1; w < nwaves; ++w) {
    double target = total * (double)w / (double)nwaves;
    while (bi < tm && pre[bi] < target) ++bi;
    b[w] = bi;
  }
  b[nwaves] = tm;
  for (int w = 1; w <= nwaves; ++w) b[w] = std::max(b[w], b[w - 1]);
  return b;
}

void window_partition_range(const mf_tiled *t, int r0, int r1, int target_nnz,
                            std::vector<WaveDesc> &out) {
  auto nnz_of = [&](int r) { return t->pay_off[t->tile_ptr[r + 1]] - t->pay_off[t->tile_ptr[r]]; };
  int r = r0;
  while (r < r1) {
    const int rb = r;
    int nz = 0;
    // first tile-row always; then extend while the window limits and the
    // entry target hold
    do {
      nz += nnz_of(r);
      ++r;
    } while (r < r1 && r - rb < kWinRows && t->tile_ptr[r + 1] - t->tile_ptr[rb] <= kWinTiles &&
             nz + nnz_of(r) <= target_nnz && nz + nnz_of(r) <= 4 * kWinChunks);
    WaveDesc d;
    d.tr0 = rb;
    d.tr1 = r;
    d.ta = t->tile_ptr[rb];
    d.T1 = t->tile_ptr[r];
    d.Z0 = t->pay_off[d.ta];
    d.Z1 = t->pay_off[d.T1];
    const bool one = d.tr1 - d.tr0 <= kWinRows && d.T1 - d.ta <= kWinTiles &&
                     d.Z1 - d.Z0 <= 4 * kWinChunks;
    d.flags = one ? kDescOneWindow : 0;
    d.aux = 0;
    out.push_back(d);
  }
}

int cut_pieces(const mf_tiled *t, int r, int h, int max_tiles, int max_slots, int extra_flags,
               std::vector<WaveDesc> &out) {
  int np = 0;
  int ta = t->tile_ptr[r];
  const int tend = t->tile_ptr[r + 1];
  while (ta < tend) {
    int tb = ta + 1;  // a tile holds at most 256 slots
    while (tb < tend && tb - ta < max_tiles && t->pay_off[tb + 1] - t->pay_off[ta] <= max_slots) ++tb;
    WaveDesc d;
    d.tr0 = r;
    d.tr1 = r + 1;
    d.ta = ta;
    d.T1 = tb;
    d.Z0 = t->pay_off[ta];
    d.Z1 = t->pay_off[tb];
    d.flags = kDescPiece | extra_flags | (np << 8);
    d.aux = h;
    out.push_back(d);
    ++np;
    ta = tb;
  }
  return np;
}

std::vector<WaveDesc> window_partition(const mf_tiled *t, int target_nnz, std::vector<int> *hn,
                                       std::vector<int> *hbase) {
  std::vector<WaveDesc> out;
  const int heavy_slots = 4 * std::max(target_nnz, 4 * kWinChunks);
  int r = 0, nh = 0, npieces = 0;
  while (r < t->tilem) {
    int e = r;
    while (e < t->tilem &&
           t->pay_off[t->tile_ptr[e + 1]] - t->pay_off[t->tile_ptr[e]] <= heavy_slots)
      ++e;
    window_partition_range(t, r, e, target_nnz, out);
    if (e == t->tilem) break;
    const int np = cut_pieces(t, e, nh, kWinTiles, 4 * kWinChunks, kDescOneWindow, out);
    if (hn) hn->push_back(np);
    if (hbase) hbase->push_back(npieces);
    npieces += np;
    ++nh;
    r = e + 1;
  }
  return out;
}

}  // namespace mf

using namespace mf;

// ------------------------------------------------------------------------
// Matrix Market / binary CSR input
// ------------------------------------------------------------------------

namespace {

struct Buf {
  char *p = nullptr;
  size_t n = 0;
  ~Buf() { std::free(p); }
};

bool slurp(const char *path, Buf &b) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0) { std::fclose(f); return false; }
  b.p = (char *)std::malloc((size_t)sz + 1);
  b.n = std::fread(b.p, 1, (size_t)sz, f);
  b.p[b.n] = '\0';
  std::fclose(f);
  return true;
}

const char *next_line(const char *s, const char *end) {
  while (s < end && *s != '\n') ++s;
  return s < end ? s + 1 : end;
}

bool parse_long(const char *&s, const char *end, long long &v) {
  while (s < end && (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r')) ++s;
  if (s >= end) return false;
  char *e;
  errno = 0;
  v = std::strtoll(s, &e, 10);
  if (e == s || errno) return false;
  s = e;
  return true;
}

bool parse_double(const char *&s, const char *end, double &v) {
  while (s < end && (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r')) ++s;
  if (s >= end) return false;
  char *e;
  v = std::strtod(s, &e);
  if (e == s) return false;
  s = e;
  return true;
}

}  // namespace

extern "C" mf_status mf_read_mtx(const char *path, mf_csr *out) {
  if (!path || !out) return fail(MF_ERR_ARG, "mf_read_mtx: null argument");
  std::memset(out, 0, sizeof(*out));
  Buf b;
  if (!slurp(path, b)) return fail(MF_ERR_IO, std::string("mf_read_mtx: cannot open ") + path);
  const char *s = b.p, *end = b.p + b.n;
  // Banner: %%MatrixMarket matrix coordinate <field> <symmetry>
  // (src/biio2.0/src/mmio.h:249-332).
  char t0[128] = {0}, t1[128] = {0}, t2[128] = {0}, t3[128] = {0}, t4[128] = {0};
  {
    std::string line(s, next_line(s, end) - s);
    if (std::sscanf(line.c_str(), "%127s %127s %127s %127s %127s", t0, t1, t2, t3, t4) != 5 ||
        std::strncmp(t0, "%%MatrixMarket", 14) != 0)
      return fail(MF_ERR_FORMAT, "mf_read_mtx: missing Matrix Market banner");
  }
  auto lower = [](char *p) { for (; *p; ++p) *p = (char)std::tolower((unsigned char)*p); };
  lower(t1); lower(t2); lower(t3); lower(t4);
  if (std::strcmp(t1, "matrix")) return fail(MF_ERR_FORMAT, "mf_read_mtx: not a matrix");
  if (std::strcmp(t2, "coordinate")) return fail(MF_ERR_UNSUPPORTED, "mf_read_mtx: only coordinate format is supported");
  enum { REAL, CPLX, INTG, PATT } field;
  if (!std::strcmp(t3, "real")) field = REAL;
  else if (!std::strcmp(t3, "complex")) field = CPLX;
  else if (!std::strcmp(t3, "integer")) field = INTG;
  else if (!std::strcmp(t3, "pattern")) field = PATT;
  else return fail(MF_ERR_FORMAT, std::string("mf_read_mtx: unknown field ") + t3);
  int sym;
  if (!std::strcmp(t4, "symmetric") || !std::strcmp(t4, "hermitian")) sym = 1;
  else if (!std::strcmp(t4, "general") || !std::strcmp(t4, "skew-symmetric")) sym = 0;
  else return fail(MF_ERR_FORMAT, std::string("mf_read_mtx: unknown symmetry ") + t4);
  // Skip the banner and every comment line; the size line follows.
  s = next_line(s, end);
  while (s < end && *s == '%') s = next_line(s, end);
  long long m, n, nz;
  if (!parse_long(s, end, m) || !parse_long(s, end, n) || !parse_long(s, end, nz) || m < 0 ||
      n < 0 || nz < 0 || m > 0x7fffffff || n > 0x7fffffff)
    return fail(MF_ERR_FORMAT, "mf_read_mtx: bad size line");
  std::vector<int> ri(nz), ci(nz);
  std::vector<double> vv(nz);
  for (long long k = 0; k < nz; ++k) {
    long long i, j;
    double re = 1.0, im;
    bool ok = parse_long(s, end, i) && parse_long(s, end, j);
    if (ok && field == REAL) ok = parse_double(s, end, re);
    else if (ok && field == CPLX) ok = parse_double(s, end, re) && parse_double(s, end, im);
    else if (ok && field == INTG) { long long iv; ok = parse_long(s, end, iv); re = (double)iv; }
    if (!ok || i < 1 || j < 1 || i > m || j > n)
      return fail(MF_ERR_FORMAT, "mf_read_mtx: bad entry " + std::to_string(k + 1));
    ri[k] = (int)(i - 1);
    ci[k] = (int)(j - 1);
    vv[k] = re;
  }
  // Counting sort in file order; a mirrored entry goes right after its
  // source (src/biio2.0/src/mmio_highlevel.h:253-312).
  std::vector<long long> cnt(m + 1, 0);
  for (long long k = 0; k < nz; ++k) {
    ++cnt[ri[k]];
    if (sym && ri[k] != ci[k]) ++cnt[ci[k]];
  }
  long long total = 0;
  for (long long i = 0; i < m; ++i) total += cnt[i];
  if (total > 0x7fffffffLL) return fail(MF_ERR_UNSUPPORTED, "mf_read_mtx: nnz exceeds 32-bit row pointers");
  out->m = (int)m;
  out->n = (int)n;
  out->nnz = total;
  out->is_sym = sym;
  out->rowptr = (int *)std::malloc(sizeof(int) * (size_t)(m + 1));
  out->colidx = (int *)std::malloc(sizeof(int) * (size_t)(total ? total : 1));
  out->val = (double *)std::malloc(sizeof(double) * (size_t)(total ? total : 1));
  if (!out->rowptr || !out->colidx || !out->val) { mf_csr_free(out); return fail(MF_ERR_NOMEM, "mf_read_mtx: out of memory"); }
  out->rowptr[0] = 0;
  for (long long i = 0; i < m; ++i) out->rowptr[i + 1] = out->rowptr[i] + (int)cnt[i];
  std::fill(cnt.begin(), cnt.end(), 0);
  for (long long k = 0; k < nz; ++k) {
    int i = ri[k], j = ci[k];
    long long o = out->rowptr[i] + cnt[i]++;
    out->colidx[o] = j;
    out->val[o] = vv[k];
    if (sym && i != j) {
      o = out->rowptr[j] + cnt[j]++;
      out->colidx[o] = i;
      out->val[o] = vv[k];
    }
  }
  return MF_OK;
}

extern "C" mf_status mf_read_cbd(const char *path, mf_csr *out) {
  if (!path || !out) return fail(MF_ERR_ARG, "mf_read_cbd: null argument");
  std::memset(out, 0, sizeof(*out));
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(MF_ERR_IO, std::string("mf_read_cbd: cannot open ") + path);
  int rows = 0, cols = 0;
  int64_t nnz = 0;
  bool ok = std::fread(&rows, 4, 1, f) == 1 && std::fread(&cols, 4, 1, f) == 1 &&
            std::fread(&nnz, 8, 1, f) == 1 && rows >= 0 && cols >= 0 && nnz >= 0 &&
            nnz <= 0x7fffffffLL;
  if (!ok) { std::fclose(f); return fail(MF_ERR_FORMAT, "mf_read_cbd: bad header"); }
  std::vector<int64_t> rp(rows + 1);
  out->m = rows;
  out->n = cols;
  out->nnz = nnz;
  out->rowptr = (int *)std::malloc(sizeof(int) * (size_t)(rows + 1));
  out->colidx = (int *)std::malloc(sizeof(int) * (size_t)(nnz ? nnz : 1));
  out->val = (double *)std::malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
  ok = out->rowptr && out->colidx && out->val &&
       std::fread(rp.data(), 8, rows + 1, f) == (size_t)(rows + 1) &&
       std::fread(out->colidx, 4, nnz, f) == (size_t)nnz &&
       std::fread(out->val, 8, nnz, f) == (size_t)nnz;
  std::fclose(f);
  if (!ok) { mf_csr_free(out); return fail(MF_ERR_FORMAT, "mf_read_cbd: truncated file"); }
  for (int i = 0; i <= rows; ++i) out->rowptr[i] = (int)rp[i];
  return MF_OK;
}

extern "C" mf_status mf_write_cbd(const char *path, const mf_csr *a) {
  if (!path || !a || !a->rowptr) return fail(MF_ERR_ARG, "mf_write_cbd: null argument");
  FILE *f = std::fopen(path, "wb");
  if (!f) return fail(MF_ERR_IO, std::string("mf_write_cbd: cannot open ") + path);
  std::vector<int64_t> rp(a->m + 1);
  for (int i = 0; i <= a->m; ++i) rp[i] = a->rowptr[i];
  int64_t nnz = a->nnz;
  bool ok = std::fwrite(&a->m, 4, 1, f) == 1 && std::fwrite(&a->n, 4, 1, f) == 1 &&
            std::fwrite(&nnz, 8, 1, f) == 1 &&
            std::fwrite(rp.data(), 8, a->m + 1, f) == (size_t)(a->m + 1) &&
            std::fwrite(a->colidx, 4, nnz, f) == (size_t)nnz &&
            std::fwrite(a->val, 8, nnz, f) == (size_t)nnz;
  std::fclose(f);
  return ok ? MF_OK : fail(MF_ERR_IO, "mf_write_cbd: write failed");
}

extern "C" void mf_csr_free(mf_csr *a) {
  if (!a) return;
  std::free(a->rowptr);
  std::free(a->colidx);
  std::free(a->val);
  a->rowptr = nullptr;
  a->colidx = nullptr;
  a->val = nullptr;
}

// Reference-named reader: .mtx or .cbd by extension, int32 row pointers
// (src/biio2.0/src/biio.h:153-194).  On failure *row is set to -1.
extern "C" void read_Dmatrix_32(int *row, int *col, int *nnz, int **row_ptr,
                                int **col_idx, double **val, int *isSymmeticeR,
                                char *filename) {
  mf_csr a;
  size_t len = filename ? std::strlen(filename) : 0;
  mf_status st;
  if (len >= 3 && !std::strcmp(filename + len - 3, "cbd")) st = mf_read_cbd(filename, &a);
  else if (len >= 3 && !std::strcmp(filename + len - 3, "mtx")) st = mf_read_mtx(filename, &a);
  else st = fail(MF_ERR_UNSUPPORTED, "read_Dmatrix_32: unsupported file type");
  if (st != MF_OK) {
    *row = -1;
    *col = -1;
    *nnz = 0;
    *row_ptr = nullptr;
    *col_idx = nullptr;
    *val = nullptr;
    return;
  }
  *row = a.m;
  *col = a.n;
  *nnz = (int)a.nnz;
  *row_ptr = a.rowptr;
  *col_idx = a.colidx;
  *val = a.val;
  if (isSymmeticeR) *isSymmeticeR = a.is_sym;
}

// ------------------------------------------------------------------------
// Handle queries and reference-derived arrays
// ------------------------------------------------------------------------

extern "C" mf_status mf_tile_info(const mf_tiled *t, int *tilem, int *tilen, int *tilenum,
                                  int *csrsize, int *csrptrlen, int *npad) {
  if (!t) return fail(MF_ERR_ARG, "mf_tile_info: null handle");
  if (tilem) *tilem = t->tilem;
  if (tilen) *tilen = t->tilen;
  if (tilenum) *tilenum = t->tilenum;
  if (csrsize) *csrsize = t->csrsize;
  if (csrptrlen) *csrptrlen = t->csrptrlen;
  if (npad) *npad = t->dev.npad;
  return MF_OK;
}

extern "C" mf_status mf_tile_arrays_get(const mf_tiled *t, mf_tile_arrays *o) {
  if (!t || !o) return fail(MF_ERR_ARG, "mf_tile_arrays_get: null argument");
  o->tile_ptr = t->tile_ptr.data();
  o->tile_columnidx = t->tile_columnidx.data();
  o->tile_nnz = t->tile_nnz.data();
  o->blknnz = t->blknnz.data();
  o->csr_offset = t->csr_offset.data();
  o->csrptr_offset = t->csrptr_offset.data();
  o->blknnznnz = t->blknnznnz.data();
  o->Tile_csr_Col = t->Tile_csr_Col.data();
  o->Blockcsr_Ptr = t->Blockcsr_Ptr.data();
  o->Blockcsr_Val = t->Blockcsr_Val.data();
  o->Blockcsr_Val_Low = t->Blockcsr_Val_Low.data();
  return MF_OK;
}

// Row-block chunking of blockspmv_cpu: tile-rows with more than 4 tiles
// (PREFETCH_SMEM_TH, src/common.h:60-64) are split into ceil(n/4) chunks of
// ceil(n/chunks) tiles, flagged with bit 31 (src/blockspmv_cpu.h:21-83).
extern "C" mf_status mf_tile_rowblkblock(const mf_tiled *t, int *rowblkblock) {
  if (!t || !rowblkblock) return fail(MF_ERR_ARG, "mf_tile_rowblkblock: null argument");
  int c = 0;
  for (int bi = 0; bi < t->tilem; ++bi) {
    int nt = t->tile_ptr[bi + 1] - t->tile_ptr[bi];
    c += nt <= 4 ? 1 : (nt + 3) / 4;
  }
  *rowblkblock = c;
  return MF_OK;
}

extern "C" mf_status mf_tile_offsets(const mf_tiled *t, int *ptroffset1, int *ptroffset2,
                                     unsigned int *rowidx, int *colstart, int *colstop) {
  if (!t) return fail(MF_ERR_ARG, "mf_tile_offsets: null handle");
  if (rowidx && colstart && colstop) {
    int c = 0;
    for (int bi = 0; bi < t->tilem; ++bi) {
      const int tp = t->tile_ptr[bi], nt = t->tile_ptr[bi + 1] - tp;
      if (nt <= 4) {
        rowidx[c] = (unsigned)bi;
        colstart[c] = colstop[c] = 0;
        ++c;
        continue;
      }
      const int chunks = (nt + 3) / 4, len = (nt + chunks - 1) / chunks;
      for (int q = 0; q < chunks; ++q, ++c) {
        rowidx[c] = (unsigned)bi | 0x80000000u;
        colstart[c] = tp + q * len;
        colstop[c] = q + 1 == chunks ? tp + nt : tp + (q + 1) * len;
      }
    }
  }
  // ptroffset1/2 are the running nnz and row-pointer offsets, identical to
  // csr_offset / csrptr_offset (src/blockspmv_cpu.h:96-110).
  if (ptroffset1) std::memcpy(ptroffset1, t->csr_offset.data(), sizeof(int) * t->tilenum);
  if (ptroffset2) std::memcpy(ptroffset2, t->csrptr_offset.data(), sizeof(int) * t->tilenum);
  return MF_OK;
}

static int tile_height(const mf_tiled *t, int bi) {
  return bi == t->tilem - 1 ? t->rowA - (t->tilem - 1) * kTile : kTile;
}

// Non-empty-row compaction (src/main-cg.cu:536-593): per tile a leading
// (0,0) pair, then (row, 8-bit end pointer) for every non-empty row.
extern "C" mf_status mf_tile_compact_count(const mf_tiled *t, int *nzr, int *cnt_non_new) {
  if (!t || !nzr) return fail(MF_ERR_ARG, "mf_tile_compact_count: null argument");
#pragma omp parallel for schedule(dynamic, 256)
  for (int bi = 0; bi < t->tilem; ++bi) {
    const int h = tile_height(t, bi);
    for (int tt = t->tile_ptr[bi]; tt < t->tile_ptr[bi + 1]; ++tt) {
      const unsigned char *rp = &t->Blockcsr_Ptr[t->csrptr_offset[tt]];
      const int nzt = t->blknnz[tt + 1] - t->blknnz[tt];
      int c = 1;
      for (int r = 0; r < h; ++r) c += (r == h - 1 ? nzt : rp[r + 1]) != rp[r];
      nzr[tt] = c;
    }
  }
  int run = 0;
  for (int tt = 0; tt < t->tilenum; ++tt) {
    int c = nzr[tt];
    nzr[tt] = run;
    run += c;
  }
  nzr[t->tilenum] = run;
  if (cnt_non_new) *cnt_non_new = run;
  return MF_OK;
}

extern "C" mf_status mf_tile_compact_fill(const mf_tiled *t, const int *nzr,
                                          unsigned char *rowid, unsigned char *ptrnew,
                                          int *block_signal, int *tile_rowidx) {
  if (!t || !nzr || !rowid || !ptrnew) return fail(MF_ERR_ARG, "mf_tile_compact_fill: null argument");
  rowid[nzr[t->tilenum]] = 0;
  ptrnew[nzr[t->tilenum]] = 0;
#pragma omp parallel for schedule(dynamic, 256)
  for (int bi = 0; bi < t->tilem; ++bi) {
    const int h = tile_height(t, bi);
    if (block_signal) block_signal[bi] = t->tile_ptr[bi + 1] - t->tile_ptr[bi];
    for (int tt = t->tile_ptr[bi]; tt < t->tile_ptr[bi + 1]; ++tt) {
      if (tile_rowidx) tile_rowidx[tt] = bi;
      const unsigned char *rp = &t->Blockcsr_Ptr[t->csrptr_offset[tt]];
      const int nzt = t->blknnz[tt + 1] - t->blknnz[tt];
      int o = nzr[tt];
      rowid[o] = 0;
      ptrnew[o] = 0;
      ++o;
      for (int r = 0; r < h; ++r) {
        const int stop = r == h - 1 ? nzt : rp[r + 1];
        if (stop != rp[r]) {
          rowid[o] = (unsigned char)r;
          ptrnew[o] = (unsigned char)stop;
          ++o;
        }
      }
    }
  }
  if (block_signal) block_signal[t->tilem] = 0;
  return MF_OK;
}

// The reference's nnz balancer (src/main-cg.cu:596-761): tiles are taken
// alternately from both ends of the tile list into a warp until its nnz
// reaches E or its tile count reaches B; E runs over the reference's list
// and B doubles until the warp count is below warp_limit.  The reference
// loops forever when nothing fits; this gives up after 24 doublings.
namespace {
int balance_sweep(int T, const std::vector<int> &nb, int E, int B, int *btp, int *order) {
  int i = 0, j = T - 1, cnt = 0, idx = 1, step = 0, cb = 0;
  auto take = [&](int tile, bool close) {
    if (order) order[step] = tile;
    ++step;
    if (close) {
      if (btp) btp[idx] = step;
      ++idx;
      cnt = 0;
      cb = 0;
    } else {
      cnt += nb[tile];
      ++cb;
    }
  };
  while (i < j) {
    bool fits = nb[i] + cnt < E && cb + 1 < B;
    take(i++, !fits);
    fits = nb[j] + cnt < E && cb + 1 < B;
    take(j--, !fits);
    if (i == j) {
      if (order) order[step] = j;
      ++step;
      if (btp) btp[idx] = step;
    } else if (i > j) {
      if (order) order[step] = j;
      if (btp) btp[idx] = step;
    }
  }
  return idx;
}
}  // namespace

extern "C" mf_status mf_tile_shared_offsets(const mf_tiled *t, int nwarps,
                                            const int *balance_tile_ptr,
                                            const int *index_each_block,
                                            int *non_each_block_offset,
                                            int *balance_tile_ptr_shared_end) {
  if (!t || !balance_tile_ptr || !index_each_block || !non_each_block_offset ||
      !balance_tile_ptr_shared_end || nwarps < 0)
    return fail(MF_ERR_ARG, "mf_tile_shared_offsets: bad argument");
  const int T = t->tilenum;
  constexpr int kSharedNnz = 256;  // LDS budget per warp in entries
  for (int w = 0; w < nwarps; ++w)
    if (balance_tile_ptr[w] < 0 || balance_tile_ptr[w] > balance_tile_ptr[w + 1] ||
        balance_tile_ptr[w + 1] > T)
      return fail(MF_ERR_ARG, "mf_tile_shared_offsets: warp bounds outside [0, tilenum]");
  std::fill(non_each_block_offset, non_each_block_offset + T + 1, 0);
  for (int w = 0; w <= nwarps; ++w) balance_tile_ptr_shared_end[w] = balance_tile_ptr[w];
  for (int w = 0; w < nwarps; ++w) {
    // the warp's tiles in balancer order while their running nnz fits
    int run = 0;
    for (int j = balance_tile_ptr[w]; j < balance_tile_ptr[w + 1]; ++j) {
      const int tj = index_each_block[j];
      if (tj < 0 || tj >= T) return fail(MF_ERR_ARG, "mf_tile_shared_offsets: tile index out of range");
      const int nz = t->blknnz[tj + 1] - t->blknnz[tj];
      run += nz;
      if (run > kSharedNnz) {
        balance_tile_ptr_shared_end[w + 1] = j;
        break;
      }
      non_each_block_offset[j] = run - nz;
    }
  }
  return MF_OK;
}

extern "C" mf_status mf_tile_balance(const mf_tiled *t, int warp_limit, int *btp,
                                     int *index_each_block, int *row_each_block,
                                     int *nwarps) {
  if (!t || !btp || !index_each_block || !row_each_block || !nwarps)
    return fail(MF_ERR_ARG, "mf_tile_balance: null argument");
  const int T = t->tilenum;
  std::vector<int> nb(T + 1, 0), row(T + 1, 0);
  for (int bi = 0; bi < t->tilem; ++bi)
    for (int tt = t->tile_ptr[bi]; tt < t->tile_ptr[bi + 1]; ++tt) {
      nb[tt] = t->blknnz[tt + 1] - t->blknnz[tt];
      row[tt] = bi;
    }
  const int list[12] = {16, 32, 64, 96, 128, 256, 512, 1024, 2048, 4096, t->csrsize / 6912, 0};
  int B = 180, E = 16, idx = 0;
  bool found = false;
  for (int d = 0; d < 24 && !found; ++d) {
    for (int k = 0; k < 12 && !found; ++k) {
      E = list[k];
      idx = balance_sweep(T, nb, E, B, nullptr, nullptr);
      found = idx < warp_limit;
    }
    if (!found) B *= 2;
  }
  if (!found || idx == 0 || t->tilem == 0)
    return fail(MF_ERR_UNSUPPORTED, "mf_tile_balance: no schedule below the warp limit");
  std::vector<int> order(T + 1, 0);
  std::fill(btp, btp + idx + 1, 0);
  *nwarps = balance_sweep(T, nb, E, B, btp, order.data());
  for (int s = 0; s <= T; ++s) {
    index_each_block[s] = order[s];
    row_each_block[s] = row[order[s]];
  }
  return MF_OK;
}

extern "C" void mf_encode_nibbles(const unsigned char *cols, unsigned char *packed, int n,
                                  int offset) {
  int o = offset;
  for (int i = 0; i + 1 < n; i += 2) packed[o++] = (unsigned char)((cols[i] << 4) | (cols[i + 1] & 15));
  if (n & 1) packed[o] = (unsigned char)(cols[n - 1] << 4);
}

extern "C" void mf_decode_nibbles(const unsigned char *packed, unsigned char *cols, int n) {
  for (int i = 0; i < n; ++i) cols[i] = (unsigned char)(i & 1 ? packed[i >> 1] & 15 : packed[i >> 1] >> 4);
}

extern "C" const char *mf_last_error(void) { return g_err.c_str(); }
extern "C" const char *mf_version(void) { return "millefeuille-hip 0.1 (gfx950)"; }
