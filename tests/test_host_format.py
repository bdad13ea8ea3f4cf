"""Product host code (reader, tile build, reference-derived arrays) against the
golden fixtures and the oracle, through the C ABI with host-only handles
(device = -1: no HIP calls).  CPU only."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from test_oracle import FIXTURES, MTX, TILE_KEYS, tiles_from_golden


def host_handle(mf, g):
    return mf.TiledMatrix(g["csr_rowptr"], g["csr_colidx"], g["csr_val"], rowA=int(g["meta_rowA"]),
                          colA=int(g["meta_colA"]), device=-1)


@pytest.mark.parametrize("name", [k for k in FIXTURES if k in MTX])
def test_product_reader_matches_reference(mf, golden, name):
    g = golden(name)
    m, n, rp, ci, va, sym = mf.read_mtx(os.path.join(GOLDEN, MTX[name]))
    assert (m, n, sym) == (int(g["meta_m"]), int(g["meta_n"]), int(g["meta_is_sym"]))
    np.testing.assert_array_equal(rp, g["csr_rowptr"])
    np.testing.assert_array_equal(ci, g["csr_colidx"])
    np.testing.assert_array_equal(va, g["csr_val"])


def test_product_reader_rejects_bad_input(mf, tmp_path):
    for text in ["junk\n", "%%MatrixMarket matrix array real general\n2 2\n",
                 "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n"]:
        p = tmp_path / "bad.mtx"
        p.write_text(text)
        with pytest.raises(mf.MfError):
            mf.read_mtx(p)
    with pytest.raises(mf.MfError):
        mf.read_mtx(tmp_path / "missing.mtx")


@pytest.mark.parametrize("name", FIXTURES)
def test_product_tiles_match_reference(mf, golden, name):
    g = golden(name)
    t = host_handle(mf, g)
    assert (t.tilem, t.tilen, t.tilenum, t.csrsize, t.csrptrlen) == tuple(
        int(g["meta_" + k]) for k in ("tilem", "tilen", "tilenum", "csrsize", "csrptrlen"))
    a = t.arrays()
    for k in TILE_KEYS:
        np.testing.assert_array_equal(a[k], g[k], err_msg=k)
    o = t.offsets()
    np.testing.assert_array_equal(o["ptroffset1"], g["ptroffset1"])
    np.testing.assert_array_equal(o["ptroffset2"], g["ptroffset2"])
    np.testing.assert_array_equal(o["blkcoostylerowidx"], g["blkcoostylerowidx"])
    np.testing.assert_array_equal(o["colstart"], g["blkcoostylerowidx_colstart"])
    np.testing.assert_array_equal(o["colstop"], g["blkcoostylerowidx_colstop"])


@pytest.mark.parametrize("name", FIXTURES)
def test_product_compaction_and_balance_match_oracle(mf, oracle, golden, name):
    g = golden(name)
    t = host_handle(mf, g)
    ot = tiles_from_golden(oracle, g)
    c, oc = t.compact(), oracle.compact(ot)
    for k in oc:
        np.testing.assert_array_equal(c[k], oc[k], err_msg=k)
    if ot["tilenum"] > 1:
        b = t.balance(6912)
        ob = oracle.load_balance(ot, int(g["meta_nnz"]))
        assert b["nwarps"] == ob["nwarps"]
        np.testing.assert_array_equal(b["balance_tile_ptr"], ob["balance_tile_ptr"])
        np.testing.assert_array_equal(b["index_each_block"], ob["index_each_block"])
        np.testing.assert_array_equal(b["row_each_block"], ob["row_each_block"])


@pytest.mark.parametrize("seed", range(4))
def test_product_tiles_random_matrices(mf, oracle, seed):
    """Random sparsity (rows with many tiles, empty rows, a short last
    tile-row, dense tiles) against the oracle restatement."""
    rng = np.random.default_rng(seed)
    m, n = int(rng.integers(1, 300)), int(rng.integers(1, 300))
    dens = rng.uniform(0.001, 0.3)
    mask = rng.random((m, n)) < dens
    if seed == 3:
        mask[:16, :16] = True  # one dense 256-entry tile
    rows, cols = np.nonzero(mask)
    # shuffle the order inside each row (file order is kept, not sorted)
    perm = np.lexsort((rng.random(len(rows)), rows))
    rows, cols = rows[perm], cols[perm]
    vals = rng.standard_normal(len(rows))
    rp = np.zeros(m + 1, np.int32)
    np.cumsum(np.bincount(rows, minlength=m), out=rp[1:])
    rowA = m if seed % 2 else (m // 16) * 16
    t = mf.TiledMatrix(rp, cols.astype(np.int32), vals, rowA=rowA, colA=n, device=-1)
    ot = oracle.tile_create(rowA, n, rp, cols, vals)
    a = t.arrays()
    for k in TILE_KEYS:
        np.testing.assert_array_equal(a[k], ot[k], err_msg=k)


def test_bad_column_is_rejected(mf):
    with pytest.raises(mf.MfError):
        mf.TiledMatrix(np.array([0, 1], np.int32), np.array([5], np.int32), np.ones(1), colA=4, device=-1)


def test_cbd_round_trip(mf, tmp_path):
    import ctypes as C
    from millefeuille._lib import MfCsr, lib
    m, n, rp, ci, va, _ = mf.read_mtx(os.path.join(GOLDEN, "add20.mtx"))
    a = MfCsr(m, n, len(va), rp.ctypes.data_as(C.POINTER(C.c_int)), ci.ctypes.data_as(C.POINTER(C.c_int)),
              va.ctypes.data_as(C.POINTER(C.c_double)), 0)
    p = str(tmp_path / "add20.cbd").encode()
    assert lib.mf_write_cbd(p, C.byref(a)) == 0
    b = MfCsr()
    assert lib.mf_read_cbd(p, C.byref(b)) == 0
    try:
        assert (b.m, b.n, b.nnz) == (m, n, len(va))
        np.testing.assert_array_equal(np.ctypeslib.as_array(b.rowptr, (m + 1,)), rp)
        np.testing.assert_array_equal(np.ctypeslib.as_array(b.colidx, (len(va),)), ci)
        np.testing.assert_array_equal(np.ctypeslib.as_array(b.val, (len(va),)), va)
    finally:
        lib.mf_csr_free(C.byref(b))
    # the layout is the reference's: int32 rows, int32 cols, int64 nnz, int64 rowptr, ...
    raw = (tmp_path / "add20.cbd").read_bytes()
    assert np.frombuffer(raw[:8], np.int32).tolist() == [m, n]
    assert np.frombuffer(raw[8:16], np.int64)[0] == len(va)
    assert len(raw) == 16 + 8 * (m + 1) + 4 * len(va) + 8 * len(va)


def test_nibble_codec_matches_oracle(oracle):
    import ctypes as C
    from millefeuille._lib import lib
    for n in (0, 1, 2, 7, 64, 255):
        cols = np.random.default_rng(n).integers(0, 16, n).astype(np.uint8)
        packed = np.zeros((n + 1) // 2 + 1, np.uint8)
        lib.mf_encode_nibbles(cols.ctypes.data, packed.ctypes.data, n, 0)
        np.testing.assert_array_equal(packed[: (n + 1) // 2], oracle.encode(cols))
        back = np.zeros(max(n, 1), np.uint8)
        lib.mf_decode_nibbles(packed.ctypes.data, back.ctypes.data, n)
        np.testing.assert_array_equal(back[:n], cols)


def test_read_dmatrix_32_compat(mf, golden):
    import ctypes as C
    from millefeuille._lib import lib
    g = golden("add20")
    m, n, nnz, sym = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    rp, ci, va = C.POINTER(C.c_int)(), C.POINTER(C.c_int)(), C.POINTER(C.c_double)()
    lib.read_Dmatrix_32(C.byref(m), C.byref(n), C.byref(nnz), C.byref(rp), C.byref(ci), C.byref(va),
                        C.byref(sym), os.path.join(GOLDEN, "add20.mtx").encode())
    assert (m.value, n.value, nnz.value) == (2395, 2395, 17319)
    np.testing.assert_array_equal(np.ctypeslib.as_array(rp, (m.value + 1,)), g["csr_rowptr"])
    np.testing.assert_array_equal(np.ctypeslib.as_array(ci, (nnz.value,)), g["csr_colidx"])


def test_generators_match_workload_sizes():
    from millefeuille import gen
    rp, ci, va = gen.laplace2d(64)
    assert len(rp) - 1 == 4096 and rp[-1] == 20224
    rp, ci, va = gen.convdiff3d(20)
    n = 20 ** 3
    assert rp[-1] == 7 * n - 6 * 20 * 20
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.all(np.diff(ci)[np.diff(rows) == 0] > 0)  # ascending columns per row
    d = va[rows == ci]
    np.testing.assert_allclose(d, 6 + 2 * 0.5 * 1.75)


def test_abi_exports_every_declared_symbol():
    """The library loads and exports every function include/millefeuille.h
    declares, and the ctypes table covers all of them."""
    import ctypes as C
    from millefeuille._lib import LIB_PATH, SIGNATURES
    hdr = open(os.path.join(ROOT, "include", "millefeuille.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", hdr))
    names -= {"if", "sizeof", "defined"}
    decl = {nm for nm in names if re.search(r"(mf_status|void|int|const char \*)\s*\*?\s*" + nm + r"\s*\(", hdr)}
    assert len(decl) >= 40
    lib = C.CDLL(LIB_PATH)
    for nm in sorted(decl):
        assert hasattr(lib, nm), nm
    assert decl <= set(SIGNATURES), sorted(decl - set(SIGNATURES))


def test_mixed_compat_library_exports_the_reference_names():
    """libmillefeuille_mixed.so (mixed-precision defaults, for the mixed
    executables) loads beside the main library and defines every
    reference-named solver entry point itself."""
    import ctypes as C
    import subprocess
    from millefeuille._lib import MIXED_LIB_PATH
    names = ["cg_solve_reduce", "cg_solve_sync", "cg_solve_inc", "cg_solve_mixed", "bicgstab",
             "bicgstab_sync", "bicgstab_inc", "bicgstab_mixed"]
    lib = C.CDLL(MIXED_LIB_PATH)
    for nm in names:
        assert hasattr(lib, nm), nm
    out = subprocess.run(["nm", "-D", "--defined-only", MIXED_LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    defined = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(names) <= defined, sorted(set(names) - defined)


def test_product_does_not_reference_the_oracle():
    """The shipped library and package never load or call the oracle."""
    pkg = os.path.join(ROOT, "mille-feuille_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", "Makefile")):
                text = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "mforacle" not in text and "import oracle" not in text, f


def _shared_offsets_closed_form(nnz_new, btp):
    """Per warp: exclusive prefix sums of its tiles' nnz while the running
    total stays <= 256; the first tile past 256 ends the warp's prefix."""
    off = np.zeros(len(nnz_new) + 1, np.int32)
    end = np.array(btp, np.int32).copy()
    for w in range(len(btp) - 1):
        run = 0
        for j in range(btp[w], btp[w + 1]):
            if run + nnz_new[j] > 256:
                end[w + 1] = j
                break
            off[j] = run
            run += nnz_new[j]
    return off, end


@pytest.mark.parametrize("name", FIXTURES)
def test_shared_queue_offsets_match_oracle_and_closed_form(mf, oracle, golden, name):
    g = golden(name)
    t = host_handle(mf, g)
    ot = tiles_from_golden(oracle, g)
    if ot["tilenum"] <= 1:
        pytest.skip("balancer needs at least two tiles")
    b = t.balance(6912)
    got = t.shared_offsets(b)
    ref = oracle.shared_offsets(ot, oracle.load_balance(ot, int(g["meta_nnz"])))
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    nb = np.diff(np.asarray(ot["blknnz"], np.int64))
    nnz_new = nb[b["index_each_block"][: ot["tilenum"]]]
    off, end = _shared_offsets_closed_form(nnz_new, b["balance_tile_ptr"])
    np.testing.assert_array_equal(got["non_each_block_offset"], off)
    np.testing.assert_array_equal(got["balance_tile_ptr_shared_end"], end)


def test_shared_queue_offsets_at_the_256_boundary(mf, oracle):
    """Hand-built warps: prefixes that overflow on their second tile and one
    whose first tile fills the 256 entries exactly."""
    # tile-rows of dense 16x16 tiles (256 nnz) and single entries
    rows, cols = [], []
    for r in range(16):
        for c in range(16):
            rows.append(r); cols.append(c)          # tile (0,0): 256 nnz
    for r in range(16, 32):
        rows.append(r); cols.append(r)              # tile (1,1): 16 nnz
    for r in range(16, 32):
        for c in range(32, 48):
            rows.append(r); cols.append(c)          # tile (1,2): 256 nnz
    m = 48
    order = np.lexsort((cols, rows))
    rows, cols = np.array(rows)[order], np.array(cols)[order]
    rp = np.zeros(m + 1, np.int32)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    va = np.ones(len(cols))
    t = mf.TiledMatrix(rp, cols.astype(np.int32), va, rowA=m, colA=m, device=-1)
    ot = oracle.tile_create(m, m, rp, cols, va)
    # warps by hand: [t0 t1 t2] in one warp, then [t1 t2], then [t2 t0]
    nb = np.diff(np.asarray(ot["blknnz"]))
    T = ot["tilenum"]
    assert T == 3 and list(nb[:3]) == [256, 16, 256]
    bal = {"nwarps": 3, "balance_tile_ptr": np.array([0, 3, 5, 7], np.int32),
           "index_each_block": np.array([1, 0, 2, 1, 2, 2, 0, 0], np.int32)}
    # the handle API takes tilenum+1 index entries; pass a 3-tile list per warp
    # through the oracle only for the permuted lists longer than tilenum
    ref = oracle.shared_offsets(dict(ot, tilenum=7), bal)
    off, end = _shared_offsets_closed_form(nb[bal["index_each_block"][:7]], bal["balance_tile_ptr"])
    np.testing.assert_array_equal(ref["non_each_block_offset"], off)
    np.testing.assert_array_equal(ref["balance_tile_ptr_shared_end"], end)
    assert list(end) == [0, 1, 4, 6]
    assert list(off[:7]) == [0, 0, 0, 0, 0, 0, 0]
    # the library rejects warp bounds past tilenum
    with pytest.raises(mf.MfError):
        t.shared_offsets(bal)
