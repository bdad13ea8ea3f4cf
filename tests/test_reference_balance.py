"""The tile load balancer (mf_tile_balance, or_load_balance) and the
shared-queue offsets (mf_tile_shared_offsets, or_shared_offsets) against the
outputs of the reference's own host loops (src/main-cg.cu:596-769 and
src/main-cg-mixed.cu:2681-2709, compiled from the reference tree by
oracle/ref_extract.sh + oracle/ref_dump.cpp and stored as fixtures by
tests/golden/gen_golden.py).  CPU only: host-only handles (device = -1)."""
import numpy as np
import pytest

from millefeuille import gen
from test_oracle import tiles_from_golden

BAL_FIXTURES = ["add20", "lap2d64", "lap3d_sym_notrunc"]


def _host(mf, g):
    return mf.TiledMatrix(g["csr_rowptr"], g["csr_colidx"], g["csr_val"], rowA=int(g["meta_rowA"]),
                          colA=int(g["meta_colA"]), device=-1)


def _check_balance(g, bal):
    w = int(g["meta_warps"])
    assert bal["nwarps"] == w
    np.testing.assert_array_equal(bal["balance_tile_ptr"][: w + 1], g["bal_balance_tile_ptr_new"])
    np.testing.assert_array_equal(bal["index_each_block"], g["bal_index_each_block_new"])
    np.testing.assert_array_equal(bal["row_each_block"], g["bal_row_each_block_new"])


def _check_shared(g, sh):
    np.testing.assert_array_equal(sh["non_each_block_offset"], g["bal_non_each_block_offset"])
    np.testing.assert_array_equal(sh["balance_tile_ptr_shared_end"],
                                  g["bal_balance_tile_ptr_shared_end"])


@pytest.mark.parametrize("name", BAL_FIXTURES)
def test_balancer_matches_reference_run(mf, oracle, golden, name):
    g = golden(name)
    t = _host(mf, g)
    b = t.balance(6912)
    _check_balance(g, b)
    ob = oracle.load_balance(tiles_from_golden(oracle, g), int(g["meta_nnz"]))
    _check_balance(g, ob)
    _check_shared(g, t.shared_offsets(b))
    _check_shared(g, oracle.shared_offsets(tiles_from_golden(oracle, g), ob))


def test_balancer_nnz_target_search_matches_reference_run(mf, oracle, golden):
    """12,032 tiles: the first nnz target (16 per warp) needs more than 6912
    warps, so the search takes the next one; the reference's choice and its
    packing must come out the same."""
    g = golden("balance_lap3d32")
    rp, ci, va = gen.laplace3d(32)
    n = len(rp) - 1
    t = mf.TiledMatrix(rp, ci, va, rowA=n, colA=n, device=-1)
    a = t.arrays()
    np.testing.assert_array_equal(a["tile_ptr"], g["tile_ptr"])
    np.testing.assert_array_equal(a["blknnz"], g["blknnz"])
    b = t.balance(6912)
    _check_balance(g, b)
    _check_shared(g, t.shared_offsets(b))
    ot = oracle.tile_create(n, n, rp, ci, va)
    ob = oracle.load_balance(ot, int(rp[-1]))
    assert ob["each_block_nnz"] == 32
    _check_balance(g, ob)
    _check_shared(g, oracle.shared_offsets(ot, ob))
