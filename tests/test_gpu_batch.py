"""Batched CG: many systems of uneven size, several right-hand sides each,
one persistent launch captured in a hipGraph (mf_batch_cg_solve).

Every solution is checked by the oracle's FP64 relative residual (<= 1e-10,
the reference's acceptance figure), and the batch must agree with the
single-system solver on the same systems."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _systems(mf, sizes, seed=4):
    from millefeuille import gen
    out = []
    for k, n in enumerate(sizes):
        rp, ci, va = gen.knn_laplacian(n, 8, seed=seed + k)
        out.append((rp, ci, va))
    return out


@pytest.mark.parametrize("policy", ["fp64", "errbound"])
def test_batch_matches_single_solves(mf, oracle, policy):
    from millefeuille import gen
    sizes = [1000, 4000, 1777, 9000, 2500, 1000]  # uneven, not multiples of 16
    mats = _systems(mf, sizes)
    tms = [mf.TiledMatrix(rp, ci, va) for rp, ci, va in mats]
    nrhs = 3
    bs = [np.stack([gen.uniform_vector(len(rp) - 1, 100 * k + j) for j in range(nrhs)])
          for k, (rp, ci, va) in enumerate(mats)]
    xs, rep = mf.batch_cg(tms, bs, nrhs=nrhs, tol=TOL, policy=policy)
    print("batch", policy, rep["iterations"].tolist(), f"{rep['total_ms']:.2f} ms", flush=True)
    assert rep["total_ms"] > 0 and np.all(rep["per_system_ms"] > 0)
    for k, (rp, ci, va) in enumerate(mats):
        for j in range(nrhs):
            rel = oracle.rel_residual(rp, ci, va, bs[k][j], xs[k][j])
            assert rel <= TOL and abs(rel - rep["rel_residual"][k, j]) <= 1e-12, (k, j, rel, rep["rel_residual"][k, j])
    # same iteration counts as the single-system FP64 solver (to a few steps)
    if policy == "fp64":
        for k, tm in enumerate(tms):
            db, dx = tm.vector(bs[k][0]), tm.vector()
            r1 = tm.cg(db, dx, tol=TOL, form="memory")
            assert abs(r1["iterations"] - rep["iterations"][k, 0]) <= max(3, 0.05 * r1["iterations"])


@pytest.mark.parametrize("level", [1, 2])
def test_batch_fixed_low_precision_tiles(mf, oracle, level):
    """FP32 / FP16 tiles (the batched workload's precisions) with FP64 refresh
    still reach 1e-10."""
    from millefeuille import gen
    mats = _systems(mf, [3000, 1200, 6000], seed=9)
    tms = [mf.TiledMatrix(rp, ci, va) for rp, ci, va in mats]
    for tm in tms:
        tm.set_precision(uniform=level)
    bs = [gen.uniform_vector(len(rp) - 1, k) for k, (rp, ci, va) in enumerate(mats)]
    xs, rep = mf.batch_cg(tms, bs, nrhs=1, tol=TOL, policy="fixed")
    for k, (rp, ci, va) in enumerate(mats):
        assert oracle.rel_residual(rp, ci, va, bs[k], xs[k][0]) <= TOL


def test_batch_initial_guess_and_zero_rhs(mf, oracle):
    from millefeuille import gen
    mats = _systems(mf, [2000, 3000], seed=20)
    tms = [mf.TiledMatrix(rp, ci, va) for rp, ci, va in mats]
    b0 = gen.uniform_vector(len(mats[0][0]) - 1, 1)
    bs = [b0, np.zeros(len(mats[1][0]) - 1)]
    xs, rep = mf.batch_cg(tms, bs, tol=TOL)
    assert np.all(xs[1] == 0.0) and rep["iterations"][1, 0] == 0
    # warm start from the solution converges at once
    xs2, rep2 = mf.batch_cg(tms, bs, x0=xs, tol=TOL)
    assert rep2["iterations"][0, 0] <= 1
    assert oracle.rel_residual(*mats[0], b0, xs2[0][0]) <= TOL
