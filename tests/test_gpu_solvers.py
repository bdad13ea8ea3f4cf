"""Single-kernel CG / BiCGSTAB on the GPU through the C ABI.

Acceptance is the reference's own figure, the FP64 relative residual
||b - A x|| / ||b||, recomputed here by the oracle on the CPU: every solve
must reach <= 1e-10 (the device-side FP64 value reported by the kernel must
agree with the oracle's to within rounding).  Iteration counts of the FP64
single-kernel CG are compared with the oracle's textbook CG.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _system(name):
    from millefeuille import gen
    if name == "lap2d64":
        rp, ci, va = gen.laplace2d(64)
        b = gen.csr_matvec(rp, ci, va, np.ones(len(rp) - 1))
    elif name == "lap3d_scaled32":
        rp, ci, va = gen.laplace3d_scaled(32, seed=2)
        b = gen.uniform_vector(len(rp) - 1, 2)
    elif name == "lap2d30":
        rp, ci, va = gen.laplace2d(30)
        b = gen.uniform_vector(len(rp) - 1, 1)
    elif name == "convdiff20":
        rp, ci, va = gen.convdiff3d(20)
        xs = gen.uniform_vector(len(rp) - 1, 3)
        b = gen.csr_matvec(rp, ci, va, xs)
    elif name == "knn":
        rp, ci, va = gen.knn_laplacian(3000, 8, seed=4)
        b = gen.uniform_vector(len(rp) - 1, 4)
    elif name == "powerlaw":
        rp, ci, va = gen.powerlaw_laplacian(20000, 300000, seed=5)
        b = gen.uniform_vector(len(rp) - 1, 5)
    else:
        raise KeyError(name)
    return rp, ci, va, b


def _solve(mf, oracle, name, method="cg", **kw):
    print(f"case {name} {method} {kw}", flush=True)
    rp, ci, va, b = _system(name)
    n = len(rp) - 1
    tm = mf.TiledMatrix(rp, ci, va)
    db, dx = tm.vector(b), tm.vector()
    rep = getattr(tm, method)(db, dx, tol=TOL, **kw)
    x = dx.get()[:n]
    rel = oracle.rel_residual(rp, ci, va, b, x)
    return rep, rel, x, (rp, ci, va, b)


@pytest.mark.parametrize("form", ["memory", "resident"])
@pytest.mark.parametrize("name", ["lap2d64", "lap3d_scaled32", "lap2d30", "knn", "powerlaw"])
def test_cg_fp64_reaches_tolerance(mf, oracle, name, form):
    rep, rel, x, (rp, ci, va, b) = _solve(mf, oracle, name, form=form)
    assert rep["status"] == 0 and rel <= TOL, (rel, rep)
    assert abs(rep["rel_residual"] - rel) <= 1e-3 * TOL + 1e-15
    _, it_ref, rel_ref = oracle.cg(rp, ci, va, b, tol=TOL)
    assert abs(rep["iterations"] - it_ref) <= max(3, 0.05 * it_ref)
    if name == "lap2d64":
        np.testing.assert_allclose(x, 1.0, atol=1e-7)


@pytest.mark.parametrize("form", ["memory", "resident"])
@pytest.mark.parametrize("policy", ["rule", "errbound"])
@pytest.mark.parametrize("name", ["lap2d64", "lap3d_scaled32", "powerlaw"])
def test_cg_mixed_policies_reach_tolerance(mf, oracle, name, policy, form):
    rep, rel, _, _ = _solve(mf, oracle, name, policy=policy, form=form)
    assert rep["status"] == 0 and rel <= TOL, (rel, rep)
    assert sum(rep["prec_hist"]) > 0, rep


@pytest.mark.parametrize("level", [1, 2, 3])
def test_cg_fixed_low_precision_refines_to_tolerance(mf, oracle, level):
    """Every tile at FP32 / FP16 / INT8: the FP64 residual recomputation
    (iterative refinement inside the Krylov run) still reaches 1e-10, falling
    back to FP64 tiles if the low-precision operator stagnates."""
    rp, ci, va, b = _system("lap3d_scaled32")
    n = len(rp) - 1
    tm = mf.TiledMatrix(rp, ci, va)
    tm.set_precision(uniform=level)
    db, dx = tm.vector(b), tm.vector()
    rep = tm.cg(db, dx, tol=TOL, policy="fixed")
    rel = oracle.rel_residual(rp, ci, va, b, dx.get()[:n])
    assert rep["status"] == 0 and rel <= TOL, (rel, rep)
    assert rep["refreshes"] >= 1, rep


@pytest.mark.parametrize("name", ["convdiff20", "lap2d30"])
@pytest.mark.parametrize("policy", ["fp64", "rule"])
def test_bicgstab_reaches_tolerance(mf, oracle, name, policy):
    rep, rel, x, (rp, ci, va, b) = _solve(mf, oracle, name, method="bicgstab", policy=policy)
    assert rep["status"] == 0 and rel <= TOL, (rel, rep)


@pytest.mark.parametrize("method", ["cg_multikernel", "bicgstab_multikernel"])
def test_multikernel_solvers(mf, oracle, method):
    name = "lap2d64" if method.startswith("cg") else "convdiff20"
    rep, rel, _, _ = _solve(mf, oracle, name, method=method)
    assert rep["status"] == 0 and rel <= TOL, (rel, rep)


def test_solver_is_deterministic_and_takes_initial_guess(mf, oracle):
    rp, ci, va, b = _system("powerlaw")
    n = len(rp) - 1
    tm = mf.TiledMatrix(rp, ci, va)
    db = tm.vector(b)
    xs = []
    for _ in range(2):
        dx = tm.vector()
        tm.cg(db, dx, tol=TOL)
        xs.append(dx.get())
    np.testing.assert_array_equal(xs[0], xs[1])
    # warm start from the solution: converges immediately
    dx = tm.vector(xs[0][:n])
    rep = tm.cg(db, dx, tol=TOL)
    assert rep["iterations"] <= 1
    assert oracle.rel_residual(rp, ci, va, b, dx.get()[:n]) <= TOL


def _tiles_per_tile_row(rp, ci):
    n = len(rp) - 1
    tr = np.repeat(np.arange(n) // 16, np.diff(rp)).astype(np.int64)
    key = np.unique(tr * (n // 16 + 1) + ci // 16)
    return np.bincount(key // (n // 16 + 1), minlength=(n + 15) // 16)


@pytest.mark.parametrize("method,policy", [("cg", "learned"), ("bicgstab", "learned"), ("cg", "fp64")])
def test_heavy_tile_rows_are_deterministic(mf, oracle, method, policy):
    """A power-law matrix whose tile-rows mostly hold more tiles than one LDS
    window (128): they are cut into pieces spread over the waves.  The owner
    of each such tile-row adds the pieces' partial sums in piece order, and
    each wave infers the levels of exactly the tiles it multiplies, so two
    runs give bit-identical solutions."""
    rp, ci, va, b = _system("powerlaw")
    n = len(rp) - 1
    assert (_tiles_per_tile_row(rp, ci) > 128).sum() > 100
    tm = mf.TiledMatrix(rp, ci, va)
    db = tm.vector(b)
    xs, its = [], []
    for _ in range(2):
        print(f"case powerlaw {method} {policy} run {len(xs)}", flush=True)
        dx = tm.vector()
        rep = getattr(tm, method)(db, dx, tol=TOL, policy=policy)
        assert rep["status"] == 0, rep
        xs.append(dx.get())
        its.append(rep["iterations"])
    assert its[0] == its[1]
    np.testing.assert_array_equal(xs[0], xs[1])
    assert oracle.rel_residual(rp, ci, va, b, xs[0][:n]) <= TOL


@pytest.mark.parametrize("policy", ["fp64", "errbound"])
def test_resident_form_is_deterministic(mf, oracle, policy):
    rp, ci, va, b = _system("lap3d_scaled32")
    n = len(rp) - 1
    tm = mf.TiledMatrix(rp, ci, va)
    db = tm.vector(b)
    xs = []
    for _ in range(2):
        dx = tm.vector()
        rep = tm.cg(db, dx, tol=TOL, policy=policy, form="resident")
        assert rep["form_used"] == mf.FORM_RESIDENT and rep["status"] == 0, rep
        xs.append(dx.get())
    np.testing.assert_array_equal(xs[0], xs[1])
    assert oracle.rel_residual(rp, ci, va, b, xs[0][:n]) <= TOL


def test_fixed_iteration_mode(mf):
    rp, ci, va, b = _system("lap2d64")
    tm = mf.TiledMatrix(rp, ci, va)
    db, dx = tm.vector(b), tm.vector()
    rep = tm.cg(db, dx, fixed_iters=100)
    assert rep["iterations"] == 100


def test_maxiter_reports_not_converged(mf):
    rp, ci, va, b = _system("lap3d_scaled32")
    tm = mf.TiledMatrix(rp, ci, va)
    db, dx = tm.vector(b), tm.vector()
    rep = tm.cg(db, dx, maxiter=5, allow_not_converged=True)
    assert rep["status"] == 1 and rep["iterations"] == 5


def test_zero_rhs(mf):
    rp, ci, va, b = _system("lap2d64")
    tm = mf.TiledMatrix(rp, ci, va)
    db, dx = tm.vector(np.zeros_like(b)), tm.vector()
    rep = tm.cg(db, dx)
    assert rep["iterations"] == 0 and not dx.get().any()


def test_reference_compatible_entry_points(mf, oracle, tmp_path, monkeypatch):
    """cg_solve_sync / bicgstab_sync with the reference's argument lists:
    x overwritten with the solution, CSV row appended under data/."""
    import ctypes as C
    from millefeuille._lib import lib
    from millefeuille import gen
    monkeypatch.chdir(tmp_path)
    (tmp_path / "data").mkdir()
    rp, ci, va = gen.laplace2d(64)
    n = len(rp) - 1
    b = gen.csr_matvec(rp, ci, va, np.ones(n))
    x = np.zeros(n)
    low = va.astype(np.float32)
    it = C.c_int(0)
    p = lambda a: a.ctypes.data  # noqa: E731
    lib.cg_solve_sync(p(rp), p(ci), p(va), p(low), p(x), p(b), n, C.byref(it), 10, 1e-5, b"lap2d64.mtx",
                      int(rp[-1]), n)
    assert oracle.rel_residual(rp, ci, va, b, x) <= TOL
    row = (tmp_path / "data" / "cg_performance.csv").read_text().strip().split(",")
    assert row[0] == "lap2d64.mtx" and int(row[3]) == rp[-1] and float(row[4]) <= TOL
    rp2, ci2, va2 = gen.convdiff3d(16)
    n2 = len(rp2) - 1
    rhs, x2 = np.ones(n2), np.zeros(n2)
    assert lib.bicgstab_sync(p(rp2), p(ci2), p(va2), p(rhs), p(x2), int(rp2[-1]), b"cd.mtx", n2) == 0
    rowsum = np.add.reduceat(va2, rp2[:-1])
    np.testing.assert_allclose(rhs, rowsum)  # rhs overwritten with A * ones
    assert oracle.rel_residual(rp2, ci2, va2, rhs, x2) <= TOL
    assert (tmp_path / "data" / "bicg_performance.csv").exists()


def _libc_flush():
    import ctypes as C
    C.CDLL(None).fflush(None)


def test_mixed_executables_reach_the_mixed_single_kernel_solvers(mf, oracle, tmp_path, monkeypatch,
                                                                 capfd):
    """The calls of the two mixed executables, made through
    libmillefeuille_mixed.so with their argument lists: cg_solve_inc as
    src/main-cg-mixed.cu:3006-3041 (n truncated to a multiple of 16, X = 1 on
    entry, b = A * 1) and bicgstab as src/main-bicg-mixed.cu:2168-2188
    (rhs = 1 and x = 0 on entry, rhs overwritten with A * 1).  stdout names the
    single-kernel solver under the rule policy, the CSV rows carry the
    reference's fields, and the FP64 residual is <= 1e-10."""
    import ctypes as C
    from millefeuille import gen
    from millefeuille._lib import MIXED_LIB_PATH
    monkeypatch.chdir(tmp_path)
    monkeypatch.delenv("MF_COMPAT_POLICY", raising=False)
    monkeypatch.delenv("MF_COMPAT_FIXED_ITERS", raising=False)
    (tmp_path / "data").mkdir()
    lib = C.CDLL(MIXED_LIB_PATH)
    p = lambda a: a.ctypes.data  # noqa: E731

    rp, ci, va = gen.laplace3d_scaled(32, seed=2)
    ori = len(rp) - 1
    n = (ori // 16) * 16
    low = va.astype(np.float32)
    x = np.ones(n)
    b = gen.csr_matvec(rp, ci, va, np.ones(ori))[:n].copy()
    it = C.c_int(0)
    print("case mixed cg_solve_inc lap3d_scaled32", flush=True)
    lib.cg_solve_inc(C.c_void_p(p(rp)), C.c_void_p(p(ci)), C.c_void_p(p(va)), C.c_void_p(p(low)),
                     C.c_void_p(p(x)), C.c_void_p(p(b)), n, C.byref(it), 10, C.c_double(1e-5),
                     b"lap3d.mtx", int(rp[-1]), ori)
    _libc_flush()
    out = capfd.readouterr().out
    line = [ln for ln in out.splitlines() if ln.startswith("mf: cg ")]
    assert line and "single-kernel policy=rule" in line[0], out
    levels = [int(v) for v in line[0].split("levels=")[1].split("/")]
    assert sum(levels[1:4]) > 0, line[0]  # FP32/FP16/INT8 tiles ran
    assert oracle.rel_residual(rp, ci, va, b, x) <= TOL
    row = (tmp_path / "data" / "cg_performance.csv").read_text().strip().split(",")
    assert row[0] == "lap3d.mtx" and int(row[3]) == rp[-1] and float(row[4]) <= TOL
    # field 5: sqrt(r.r), the absolute recursive residual norm
    assert 0.0 < float(row[5]) <= 2 * TOL * np.linalg.norm(b), row

    rp2, ci2, va2 = gen.convdiff3d(20)
    m = len(rp2) - 1
    rhs, x2 = np.ones(m), np.zeros(m)
    print("case mixed bicgstab convdiff20", flush=True)
    ret = lib.bicgstab(C.c_void_p(p(rp2)), C.c_void_p(p(ci2)), C.c_void_p(p(va2)), C.c_void_p(p(rhs)),
                       C.c_void_p(p(x2)), int(rp2[-1]), b"cd.mtx", m)
    _libc_flush()
    out = capfd.readouterr().out
    assert ret == 0, out
    line = [ln for ln in out.splitlines() if ln.startswith("mf: bicgstab ")]
    assert line and "single-kernel policy=rule" in line[0], out
    np.testing.assert_allclose(rhs, np.add.reduceat(va2, rp2[:-1]))
    assert oracle.rel_residual(rp2, ci2, va2, rhs, x2) <= TOL
    row = (tmp_path / "data" / "bicg_performance.csv").read_text().strip().split(",")
    rel, r2 = float(row[4]), float(row[5])
    # fields 4 and 5: ||r|| / ||rhs|| and ||r||^2 / ||r0||, r0 = rhs
    assert 0.0 < rel <= 2 * TOL
    np.testing.assert_allclose(r2, rel * rel * np.linalg.norm(rhs), rtol=1e-5)


def test_compat_policy_switch_reaches_bicgstab(mf, oracle, tmp_path, monkeypatch, capfd):
    """In the FP64 library, bicgstab is the multi-kernel FP64 solver unless
    MF_COMPAT_POLICY asks for a mixed policy."""
    import ctypes as C
    from millefeuille import gen
    from millefeuille._lib import lib
    monkeypatch.chdir(tmp_path)
    (tmp_path / "data").mkdir()
    rp, ci, va = gen.convdiff3d(16)
    m = len(rp) - 1
    p = lambda a: a.ctypes.data  # noqa: E731
    for env, want in ((None, "multi-kernel policy=fp64"), ("errbound", "single-kernel policy=errbound")):
        if env is None:
            monkeypatch.delenv("MF_COMPAT_POLICY", raising=False)
        else:
            monkeypatch.setenv("MF_COMPAT_POLICY", env)
        rhs, x = np.ones(m), np.zeros(m)
        print(f"case bicgstab MF_COMPAT_POLICY={env}", flush=True)
        assert lib.bicgstab(p(rp), p(ci), p(va), p(rhs), p(x), int(rp[-1]), b"cd.mtx", m) == 0
        _libc_flush()
        out = capfd.readouterr().out
        assert any(ln.startswith("mf: bicgstab " + want) for ln in out.splitlines()), out
        assert oracle.rel_residual(rp, ci, va, rhs, x) <= TOL


@pytest.mark.parametrize("mixed", [False, True])
def test_driver_cli(mixed, tmp_path):
    import subprocess
    from conftest import ROOT
    from millefeuille import gen
    exe = os.path.join(ROOT, "mille-feuille_amd", "bin", "mf-solve")
    rp, ci, va = gen.laplace2d(48)
    p = tmp_path / "lap.mtx"
    gen.write_mtx(str(p), rp, ci, va)
    (tmp_path / "data").mkdir()
    out = subprocess.run([exe] + (["--mixed"] if mixed else []) + [str(p)], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "time_cg=" in out.stdout
    err = float(out.stdout.strip().splitlines()[-1].split("=")[1])
    assert err < 1e-7, out.stdout + out.stderr
