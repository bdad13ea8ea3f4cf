"""Autotuner: candidate schedules are timed on the device and the fastest is
kept (mf_tune_spmv, mf_tune_batch).  Whatever it picks, results must stay
those of the oracle: the tuned SpMV against the oracle's tiled SpMV, the
tuned batch to the FP64-checked 1e-10 residual."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.mark.parametrize("mode", [0, 1])
def test_tune_spmv_keeps_fastest_and_stays_exact(mf, oracle, mode):
    from millefeuille import gen
    rp, ci, va = gen.laplace3d_scaled(40, seed=3)
    n = len(rp) - 1
    tm = mf.TiledMatrix(rp, ci, va)
    if mode == 1:
        lv = np.random.default_rng(5).choice([1, 2, 3], size=tm.tilenum).astype(np.uint8)
        tm.set_precision(levels=lv)
    rep = tm.tune_spmv(mode)
    print("tune spmv", mode, rep, flush=True)
    assert len(rep["cand"]) >= 6 and all(u > 0 for u in rep["us"])
    assert rep["best"] in rep["cand"]
    assert rep["us"][rep["cand"].index(rep["best"])] == min(rep["us"])
    x = gen.uniform_vector(n, 8)
    dx, dy = tm.vector(x), tm.vector()
    dx32 = tm.vector(x.astype(np.float32), dtype=np.float32)
    tm.spmv(dx, dy, mode=mode, d_x32=dx32)
    got = dy.get()[:n]
    if mode == 0:
        ot = oracle.tile_create(n, n, rp, ci, va)
        ref = oracle.blockspmv_tiled(ot, x)
        bound = oracle.csr_spmv(rp, ci, np.abs(va), np.abs(x))
        assert np.all(np.abs(got - ref) <= 1e-14 * bound + 1e-300)
    # every other candidate schedule gives the tuned result up to the order
    # of the FP64 additions
    bound = oracle.csr_spmv(rp, ci, np.abs(va), np.abs(x))
    for wn in (384, 2048):
        tm.set_schedule(wn)
        tm.set_spmv_kernel("stream")
        dy2 = tm.vector()
        tm.spmv(dx, dy2, mode=mode, d_x32=dx32)
        assert np.all(np.abs(dy2.get()[:n] - got) <= 1e-14 * bound + 1e-300)


def test_tune_batch_picks_a_width_and_converges(mf, oracle):
    from millefeuille import gen
    sizes = [2000, 700, 5000, 1300]
    mats = [gen.knn_laplacian(n, 8, seed=20 + k) for k, n in enumerate(sizes)]
    tms = [mf.TiledMatrix(rp, ci, va) for rp, ci, va in mats]
    nrhs = 2
    bs = [np.stack([gen.uniform_vector(len(rp) - 1, 7 * k + j) for j in range(nrhs)])
          for k, (rp, ci, va) in enumerate(mats)]
    xs, rep = mf.batch_cg(tms, bs, nrhs=nrhs, tol=TOL, policy="fp64", tune_iters=10)
    print("tune batch", rep["tune"], rep["batch_waves"], flush=True)
    assert rep["tune"]["cand"] == [4, 8, 16] and rep["batch_waves"] == rep["tune"]["best"]
    for k, (rp, ci, va) in enumerate(mats):
        for j in range(nrhs):
            assert oracle.rel_residual(rp, ci, va, bs[k][j], xs[k][j]) <= TOL


@pytest.mark.parametrize("waves", [4, 16])
def test_batch_widths_agree(mf, oracle, waves):
    from millefeuille import gen
    mats = [gen.knn_laplacian(n, 8, seed=40 + k) for k, n in enumerate([1500, 3100])]
    tms = [mf.TiledMatrix(rp, ci, va) for rp, ci, va in mats]
    bs = [gen.uniform_vector(len(rp) - 1, k) for k, (rp, ci, va) in enumerate(mats)]
    xs, rep = mf.batch_cg(tms, bs, nrhs=1, tol=TOL, policy="fp64", batch_waves=waves)
    for k, (rp, ci, va) in enumerate(mats):
        assert oracle.rel_residual(rp, ci, va, bs[k], xs[k][0]) <= TOL
