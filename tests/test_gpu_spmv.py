"""Tiled SpMV on the GPU through the C ABI, against the oracle.

Tolerances (stated per precision):
  FP64 tiles: |y - y_ref| <= 1e-14 * sum_j |a_ij x_j|  (y_ref = oracle tiled
              FP64 SpMV; different summation order and FMA only).
  FP32/FP16/INT8 tiles: the reference value is computed in FP64 from the
              dequantised payload and the FP32-rounded x; the GPU accumulates
              each tile row in FP32 (at most 16 terms), so
              |y - y_ref| <= 2e-6 * sum_j |deq(a_ij) x32_j| + 1e-14 * sum |a x|.
Payload packing is pinned separately: FP16 scales are 2^floor(log2 max|a|),
INT8 payloads and scales equal the oracle's closed form per tile.
"""
import numpy as np
import pytest

from test_oracle import tiles_from_golden

pytestmark = pytest.mark.gpu


def _mk(mf, rp, ci, va, rowA=None, colA=None):
    n = len(rp) - 1
    return mf.TiledMatrix(rp, ci, va, rowA=n if rowA is None else rowA, colA=n if colA is None else colA)


def _cases():
    from millefeuille import gen
    return {
        "lap2d64": lambda: gen.laplace2d(64),
        "lap2d30_short_tilerow": lambda: gen.laplace2d(30),
        "lap3d_scaled24": lambda: gen.laplace3d_scaled(24, seed=2),
        "convdiff20": lambda: gen.convdiff3d(20),
        "powerlaw": lambda: gen.powerlaw_laplacian(3000, 40000, seed=5),
    }


@pytest.mark.parametrize("name", ["lap2d64", "lap2d30_short_tilerow", "lap3d_scaled24", "convdiff20", "powerlaw"])
def test_spmv_fp64_matches_oracle(mf, oracle, name):
    print(f"case {name}", flush=True)
    rp, ci, va = _cases()[name]()
    n = len(rp) - 1
    tm = _mk(mf, rp, ci, va)
    x = np.random.default_rng(1).uniform(-1, 1, n)
    dx, dy = tm.vector(x), tm.vector()
    tm.spmv(dx, dy, mode=0)
    y = dy.get()
    ot = oracle.tile_create(n, n, rp, ci, va)
    yr = oracle.blockspmv_tiled(ot, x)
    bound = oracle.csr_spmv(rp, ci, np.abs(va), np.abs(x))
    assert np.all(np.abs(y[:n] - yr) <= 1e-14 * bound + 1e-300)
    assert not y[n:].any()  # padding rows are zero


@pytest.mark.parametrize("name", ["add20", "lap2d64", "lap3d_sym_notrunc"])
def test_spmv_fp64_golden_inputs(mf, oracle, golden, name):
    g = golden(name)
    rowA, colA = int(g["meta_rowA"]), int(g["meta_colA"])
    tm = mf.TiledMatrix(g["csr_rowptr"], g["csr_colidx"], g["csr_val"], rowA=rowA, colA=colA)
    x = g["spmv_x"][:colA]
    dx, dy = tm.vector(x), tm.vector()
    tm.spmv(dx, dy)
    ot = tiles_from_golden(oracle, g)
    yr = oracle.blockspmv_tiled(ot, x)
    bound = oracle.blockspmv_tiled(dict(ot, Blockcsr_Val=np.abs(ot["Blockcsr_Val"])), np.abs(x))
    assert np.all(np.abs(dy.get()[:rowA] - yr) <= 1e-14 * bound + 1e-300)


def _dequant_reference(tm, levels, x):
    a = tm.arrays()
    s16, s8 = tm.scales()
    val = a["Blockcsr_Val"]
    off = a["csr_offset"]
    T = tm.tilenum
    nz = np.diff(off)
    tile_of = np.repeat(np.arange(T), nz)
    lv = levels[tile_of]
    deq = val.copy()
    m32 = lv == 1
    deq[m32] = val[m32].astype(np.float32)
    m16 = lv == 2
    sc = s16[tile_of].astype(np.float64)
    deq[m16] = ((val[m16].astype(np.float32) / sc[m16].astype(np.float32)).astype(np.float16)
                .astype(np.float64) * sc[m16])
    m8 = lv == 3
    s8e = s8[tile_of].astype(np.float64)
    deq[m8] = np.clip(np.rint(val[m8] / s8e[m8]), -127, 127) * s8e[m8]
    deq[lv == 4] = 0.0
    # rows and columns of every entry
    rows = np.zeros(len(val), np.int64)
    cols = np.zeros(len(val), np.int64)
    for bi in range(tm.tilem):
        h = min(16, tm.rowA - 16 * bi)
        for t in range(a["tile_ptr"][bi], a["tile_ptr"][bi + 1]):
            rp = a["Blockcsr_Ptr"][a["csrptr_offset"][t]: a["csrptr_offset"][t] + h].astype(np.int64)
            ends = np.append(rp[1:], nz[t])
            for r in range(h):
                rows[off[t] + rp[r]: off[t] + ends[r]] = 16 * bi + r
            cols[off[t]: off[t + 1]] = 16 * a["tile_columnidx"][t] + a["Tile_csr_Col"][off[t]: off[t + 1]]
    xlow = x.astype(np.float32).astype(np.float64)
    xe = np.where(lv == 0, x[cols], xlow[cols])
    y = np.bincount(rows, weights=deq * xe, minlength=16 * tm.tilem)
    blow = np.bincount(rows, weights=np.where(lv == 0, 0.0, np.abs(deq * xe)), minlength=16 * tm.tilem)
    bhigh = np.bincount(rows, weights=np.abs(val * x[cols]), minlength=16 * tm.tilem)
    return y, blow, bhigh


@pytest.mark.parametrize("name", ["lap3d_scaled24", "convdiff20", "powerlaw"])
def test_spmv_mixed_per_tile_levels(mf, oracle, name):
    print(f"case {name}", flush=True)
    rp, ci, va = _cases()[name]()
    n = len(rp) - 1
    tm = _mk(mf, rp, ci, va)
    rng = np.random.default_rng(5)
    levels = rng.integers(0, 5, tm.tilenum).astype(np.uint8)
    tm.set_precision(levels)
    np.testing.assert_array_equal(tm.get_precision(), levels)
    x = rng.uniform(-1, 1, n)
    dx, dy = tm.vector(x), tm.vector()
    dx32 = tm.vector(x.astype(np.float32), dtype=np.float32)
    tm.spmv(dx, dy, mode=1, d_x32=dx32)
    y = dy.get()[:n]
    yr, blow, bhigh = _dequant_reference(tm, levels, x)
    assert np.all(np.abs(y - yr[:n]) <= 2e-6 * blow[:n] + 1e-14 * bhigh[:n] + 1e-300)


def test_payload_scales_and_quant_errors(mf, oracle):
    from millefeuille import gen
    rp, ci, va = gen.laplace3d_scaled(16, seed=2)
    tm = _mk(mf, rp, ci, va)
    a = tm.arrays()
    s16, s8 = tm.scales()
    err = tm.quant_error()
    off = a["csr_offset"]
    for t in range(0, tm.tilenum, 7):
        v = a["Blockcsr_Val"][off[t]:off[t + 1]]
        amax = np.max(np.abs(v))
        assert s16[t] == 2.0 ** (np.frexp(amax)[1] - 1)
        q, s = oracle.quant_int8(v)
        assert s8[t] == s
        assert err[2, t] == np.max(np.abs(v - q * np.float64(s)))
        assert err[0, t] == np.max(np.abs(v - v.astype(np.float32)))
        h = (v.astype(np.float32) / np.float32(s16[t])).astype(np.float16).astype(np.float64) * s16[t]
        assert err[1, t] == np.max(np.abs(v - h))
        # in-range tiles: the scaled FP16 payload equals the reference's (half)(float)a
        np.testing.assert_array_equal(h, oracle.quant_fp16(v).view(np.float16).astype(np.float64))
    # the -1 coupling tiles are exact in INT8 and FP16
    offdiag = [t for t in range(tm.tilenum) if np.all(a["Blockcsr_Val"][off[t]:off[t + 1]] == -1.0)]
    assert offdiag and all(err[2, t] == 0 and err[1, t] == 0 for t in offdiag)


def test_spmv_is_deterministic(mf):
    from millefeuille import gen
    rp, ci, va = gen.powerlaw_laplacian(5000, 80000, seed=9)
    n = len(rp) - 1
    tm = _mk(mf, rp, ci, va)
    x = np.random.default_rng(2).standard_normal(n)
    dx, dy = tm.vector(x), tm.vector()
    tm.spmv(dx, dy)
    y1 = dy.get()
    dy.zero()
    tm.spmv(dx, dy)
    np.testing.assert_array_equal(y1, dy.get())


def test_spmv_empty_and_rectangular(mf, oracle):
    # empty rows and a matrix with more tile columns than tile-rows
    rng = np.random.default_rng(4)
    m, ncol = 40, 100
    mask = rng.random((m, ncol)) < 0.05
    mask[5] = False
    rows, cols = np.nonzero(mask)
    rp = np.zeros(m + 1, np.int32)
    np.cumsum(np.bincount(rows, minlength=m), out=rp[1:])
    va = rng.standard_normal(len(rows))
    tm = mf.TiledMatrix(rp, cols.astype(np.int32), va, rowA=m, colA=ncol)
    assert tm.npad == 16 * 7
    x = rng.uniform(-1, 1, ncol)
    dx, dy = tm.vector(x), tm.vector()
    tm.spmv(dx, dy)
    y = dy.get()
    yr = oracle.csr_spmv(rp, cols, va, x)
    np.testing.assert_allclose(y[:m], yr, rtol=0, atol=1e-13)
    assert y[5] == 0 and not y[m:].any()


def _arrow(n=3000, dense_rows=16, seed=9):
    """Arrow matrix: the first `dense_rows` rows are dense (one tile-row of
    ceil(n/16) tiles and dense_rows*n entries, more than one SpMV window holds:
    128 tiles, 4096 slots), the rest is a tridiagonal band (tile-rows that fit
    one window).  Symmetric pattern, random values."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(dense_rows):
        rows.append(np.full(n, i))
        cols.append(np.arange(n))
    for i in range(dense_rows, n):
        js = [j for j in (i - 1, i, i + 1) if dense_rows <= j < n]
        js += list(range(dense_rows))  # the arrow's column part
        rows.append(np.full(len(js), i))
        cols.append(np.array(js))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    order = np.lexsort((c, r))
    r, c = r[order], c[order]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    rp = np.cumsum(rp)
    return rp, c.astype(np.int32), rng.uniform(-1, 1, len(c))


@pytest.mark.parametrize("n", [600, 3000])
def test_spmv_window_walk_pieces_and_one_window_descriptors(mf, oracle, n):
    """Tile-row 0 of the arrow matrix exceeds one window (128 tiles or 4096
    slots).  At n = 600 (9600 entries) one wave walks it in several windows,
    carrying its partial row sums; at n = 3000 (48000 entries, above the
    heavy-row bound 4*max(wave_nnz, 4096)) it is cut into one-window pieces
    whose last finisher adds the row sums.  The band's tile-rows take the
    one-window staging.  All against the oracle, FP64 tiles and per-tile
    levels, at several wave schedules, and bit-identical on a repeat."""
    rp, ci, va = _arrow(n=n)
    tm = _mk(mf, rp, ci, va)
    a = tm.arrays()
    tiles_per_row = np.diff(a["tile_ptr"])
    row0_entries = a["csr_offset"][tiles_per_row[0]]
    assert (tiles_per_row[0] > 128 or row0_entries > 4096) and tiles_per_row[1:].max() <= 4
    assert (row0_entries > 16384) == (n == 3000)
    x = np.random.default_rng(3).uniform(-1, 1, n)
    dx, dy = tm.vector(x), tm.vector()
    dx32 = tm.vector(x.astype(np.float32), dtype=np.float32)
    ot = oracle.tile_create(n, n, rp, ci, va)
    yr = oracle.blockspmv_tiled(ot, x)
    bound = oracle.csr_spmv(rp, ci, np.abs(va), np.abs(x))
    levels = np.random.default_rng(4).integers(0, 5, tm.tilenum).astype(np.uint8)
    ym, blow, bhigh = _dequant_reference(tm, levels, x)
    for wn in (256, 1024, 4096):
        print(f"n {n} wave_nnz {wn}", flush=True)
        tm.set_schedule(wn)
        tm.spmv(dx, dy, mode=0)
        y0 = dy.get()[:n]
        assert np.all(np.abs(y0 - yr) <= 1e-14 * bound + 1e-300)
        tm.spmv(dx, dy, mode=0)
        np.testing.assert_array_equal(dy.get()[:n], y0)
        tm.set_precision(levels)
        tm.spmv(dx, dy, mode=1, d_x32=dx32)
        assert np.all(np.abs(dy.get()[:n] - ym[:n]) <= 2e-6 * blow[:n] + 1e-14 * bhigh[:n] + 1e-300)
        tm.set_precision(None, uniform=0)
