"""mf_cg_solve_noalloc (the CG with no device allocation, over the library's
fixed scratch array) against mf_cg_solve on the same inputs.

Both launch the same kernels over the same schedule, so x must be
bit-identical, with the same iteration count, and the oracle's FP64
relative residual must be <= 1e-10.  Two different systems are solved in a
row, the second smaller than the first (the scratch keeps the first call's
contents); a system whose workspace does not fit the scratch is rejected
with MF_ERR_ARG.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _systems():
    from millefeuille import gen
    big = gen.laplace3d_scaled(40, seed=2)            # 64,000 rows
    small = gen.laplace2d(96)                         # 9,216 rows
    return [("lap3d_scaled40", big, gen.uniform_vector(len(big[0]) - 1, 2)),
            ("lap2d96", small, gen.uniform_vector(len(small[0]) - 1, 1))]


@pytest.mark.parametrize("policy", ["fp64", "errbound"])
def test_noalloc_matches_cg_two_calls(mf, oracle, policy):
    handles = []
    for name, (rp, ci, va), b in _systems():
        print(f"case {name} {policy}", flush=True)
        n = len(rp) - 1
        tm = mf.TiledMatrix(rp, ci, va)
        handles.append(tm)
        db, dx = tm.vector(b), tm.vector()
        ref = tm.cg(db, dx, tol=TOL, policy=policy)
        x_ref = dx.get()[:n]
        dx.zero()
        rep = tm.cg_noalloc(db, dx, tol=TOL, policy=policy)
        x = dx.get()[:n]
        assert rep["iterations"] == ref["iterations"]
        np.testing.assert_array_equal(x, x_ref)
        rel = oracle.rel_residual(rp, ci, va, b, x)
        assert rel <= TOL, rel
        for o in (db, dx):
            o.free()
    for tm in handles:
        tm.free()


def test_noalloc_rejects_a_system_beyond_the_scratch(mf):
    from millefeuille import gen
    from millefeuille._lib import MfError
    rp, ci, va = gen.laplace3d_scaled(64, seed=2)     # 262,144 rows
    tm = mf.TiledMatrix(rp, ci, va)
    db, dx = tm.vector(gen.uniform_vector(len(rp) - 1, 2)), tm.vector()
    with pytest.raises(MfError) as ei:
        tm.cg_noalloc(db, dx, tol=TOL)
    assert ei.value.code == -1  # MF_ERR_ARG
    for o in (db, dx, tm):
        o.free()
