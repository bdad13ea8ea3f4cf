"""The named workloads at their full sizes, on the GPU through the C ABI.

Each solve is accepted on the reference's own figure, the FP64 relative
residual ||b - A x|| / ||b|| <= 1e-10 (src/main-cg.cu:978-1011), recomputed on
the CPU by the oracle from the returned x:

- 3D 7-point Laplacian 128^3 with log-uniform diagonal scaling (n = 2,097,152):
  FP64 SpMV against the oracle's tiled SpMV within 1e-14 * sum|a x| per row;
  FP64 and learned CG, the FP64 iteration count within 5% of the oracle's CG;
- 3D upwind convection-diffusion 100^3 (n = 1,000,000): BiCGSTAB under the
  learned, errbound and FP64 policies;
- 512 kNN-graph Laplacians (n ~ U[1k, 64k]) x 8 right-hand sides in one
  hipGraph-captured persistent launch, every (system, right-hand side);
- power-law graph Laplacian, n = 4,000,000, 60,000,000 nnz: learned and rule
  CG, each allowed to fall back to FP64 tiles on stagnation.

The file is named so that it runs after the faster GPU tests.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def lap128(mf):
    from millefeuille import gen, policy
    rp, ci, va = gen.laplace3d_scaled(128, seed=2)
    n = len(rp) - 1
    tm = mf.TiledMatrix(rp, ci, va)
    policy.install(tm, policy.load_or_train())
    return rp, ci, va, gen.uniform_vector(n, 2), tm


def test_lap128_fp64_spmv_full_size(oracle, lap128):
    rp, ci, va, _, tm = lap128
    n = len(rp) - 1
    _say(f"case lap128 spmv n={n}")
    x = np.random.default_rng(7).uniform(-1, 1, n)
    dx, dy = tm.vector(x), tm.vector()
    tm.spmv(dx, dy)
    ot = oracle.tile_create(n, n, rp, ci, va)
    yr = oracle.blockspmv_tiled(ot, x)
    bound = oracle.csr_spmv(rp, ci, np.abs(va), np.abs(x))
    err = np.abs(dy.get()[:n] - yr)
    assert np.all(err <= 1e-14 * bound), float((err / bound).max())


@pytest.mark.timeout(600)
@pytest.mark.parametrize("policy", ["fp64", "learned"])
def test_lap128_cg_full_size(oracle, lap128, policy):
    rp, ci, va, b, tm = lap128
    n = len(rp) - 1
    _say(f"case lap128 cg {policy}")
    db, dx = tm.vector(b), tm.vector()
    rep = tm.cg(db, dx, tol=TOL, policy=policy)
    rel = oracle.rel_residual(rp, ci, va, b, dx.get()[:n])
    _say(f"  {rep['iterations']} iterations, {rep['time_ms']:.2f} ms, rel {rel:.3e}")
    assert rep["status"] == 0 and rel <= TOL, (rel, rep)
    if policy == "fp64":
        _, it_ref, rel_ref = oracle.cg(rp, ci, va, b, tol=TOL)
        _say(f"  oracle CG: {it_ref} iterations, rel {rel_ref:.3e}")
        assert abs(rep["iterations"] - it_ref) <= 0.05 * it_ref, (rep["iterations"], it_ref)
    else:
        assert sum(rep["prec_hist"][1:4]) > 0, rep


@pytest.fixture(scope="module")
def convdiff100(mf):
    from millefeuille import gen, policy
    rp, ci, va = gen.convdiff3d(100)
    n = len(rp) - 1
    b = gen.csr_matvec(rp, ci, va, gen.uniform_vector(n, 3))
    tm = mf.TiledMatrix(rp, ci, va)
    policy.install(tm, policy.load_or_train())
    return rp, ci, va, b, tm


@pytest.mark.parametrize("policy", ["learned", "errbound", "fp64"])
def test_convdiff100_bicgstab_full_size(oracle, convdiff100, policy):
    rp, ci, va, b, tm = convdiff100
    n = len(rp) - 1
    _say(f"case convdiff100 bicgstab {policy}")
    db, dx = tm.vector(b), tm.vector()
    rep = tm.bicgstab(db, dx, tol=TOL, policy=policy)
    rel = oracle.rel_residual(rp, ci, va, b, dx.get()[:n])
    _say(f"  {rep['iterations']} iterations, {rep['time_ms']:.2f} ms, rel {rel:.3e}")
    assert rep["status"] == 0 and rel <= TOL, (rel, rep)


@pytest.mark.timeout(900)
def test_knn512_batch_full_size(mf, oracle):
    from millefeuille import gen
    nsys, nrhs = 512, 8
    sizes = np.random.default_rng(4).integers(1000, 65537, nsys)
    _say(f"case knn512 batch: generating {nsys} systems, sum n = {int(sizes.sum())}")
    with ThreadPoolExecutor(8) as ex:
        mats = list(ex.map(lambda k: gen.knn_laplacian(int(sizes[k]), 8, seed=4 + k, shift=1e-2),
                           range(nsys)))
    bs = [np.stack([gen.uniform_vector(int(nk), 1000 * k + j) for j in range(nrhs)])
          for k, nk in enumerate(sizes)]
    tms = [mf.TiledMatrix(*m) for m in mats]
    _say("  solving (one hipGraph launch, errbound policy)")
    xs, rep = mf.batch_cg(tms, bs, nrhs=nrhs, tol=TOL, policy="errbound")
    _say(f"  {rep['total_ms']:.1f} ms, iterations max {int(np.max(rep['iterations']))}")
    assert np.max(rep["rel_residual"]) <= TOL, float(np.max(rep["rel_residual"]))
    worst = 0.0
    for k, (rp, ci, va) in enumerate(mats):
        for j in range(nrhs):
            worst = max(worst, oracle.rel_residual(rp, ci, va, bs[k][j], xs[k][j]))
    _say(f"  worst oracle rel residual {worst:.3e} over {nsys * nrhs} solves")
    assert worst <= TOL
    for tm in tms:
        tm.free()


@pytest.mark.timeout(900)
def test_powerlaw4m_learned_and_rule_full_size(mf, oracle):
    from millefeuille import gen, policy
    _say("case powerlaw4m: generating n = 4,000,000, 60,000,000 nnz target")
    rp, ci, va = gen.powerlaw_laplacian(4_000_000, 60_000_000, seed=5)
    n = len(rp) - 1
    b = gen.uniform_vector(n, 2)
    tm = mf.TiledMatrix(rp, ci, va)
    policy.install(tm, policy.load_or_train())
    _say(f"  nnz = {int(rp[-1])}, tiles = {tm.tilenum}")
    db, dx = tm.vector(b), tm.vector()
    for pol in ("learned", "rule"):
        dx.zero()
        rep = tm.cg(db, dx, tol=TOL, policy=pol)
        rel = oracle.rel_residual(rp, ci, va, b, dx.get()[:n])
        _say(f"  {pol}: {rep['iterations']} iterations, {rep['time_ms']:.1f} ms, "
             f"fell back {rep['fell_back']}, rel {rel:.3e}")
        assert rep["status"] == 0 and rel <= TOL, (pol, rel, rep)
    tm.free()
