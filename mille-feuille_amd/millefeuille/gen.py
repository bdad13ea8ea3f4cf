"""Seeded synthetic systems for the tiled CG / BiCGSTAB workloads.

Every generator returns a CSR triple (rowptr int32, colidx int32, val f64)
with ascending column indices inside each row, which is the entry order the
16x16 tile format preserves (the reference keeps file order inside a row,
src/biio2.0/src/mmio_highlevel.h:303-312, so a .mtx written from these arrays
tiles identically).  Grid points are numbered with x fastest:
row = i + nx*(j + ny*k).

Workloads:
  laplace2d(64)              2D 5-point Dirichlet Laplacian, diag 4, off -1.
  laplace3d_scaled(128, 2)   3D 7-point Laplacian, off -1, diagonal
                             6*10**(3U), U ~ U[0,1) from the seed.
  convdiff3d(100, 0.5, ...)  3D 7-point first-order upwind
                             convection-diffusion: diag 6 + 2*Pe*sum(w),
                             neighbours -1, upstream neighbour in direction d
                             gets an extra -2*Pe*w_d, w = v/max|v|.
  knn_laplacian(n, k, seed)  symmetrised k-nearest-neighbour graph Laplacian
                             on U[0,1]^3 points plus a diagonal shift.
  powerlaw_laplacian(...)    Chung-Lu graph Laplacian (degree exponent 2.5)
                             plus a diagonal shift.
"""
from __future__ import annotations

import numpy as np


def _csr_from_coo(n, rows, cols, vals):
    # row-major, ascending columns (one int64 key; the same order as
    # np.lexsort((cols, rows)), several times faster at 60M entries)
    order = np.argsort(rows.astype(np.int64) * n + cols, kind="stable")
    rows = rows[order]
    cols = cols[order].astype(np.int32)
    vals = vals[order].astype(np.float64)
    counts = np.bincount(rows, minlength=n)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    if rowptr[-1] > np.iinfo(np.int32).max:
        raise ValueError("nnz does not fit 32-bit row pointers")
    return rowptr.astype(np.int32), cols, vals


def _stencil(shape, offsets, diag_vals, off_vals):
    """Structured stencil on an nx*ny*nz grid (x fastest).

    offsets: list of (dx, dy, dz); off_vals: per-offset value (scalar).
    diag_vals: array of length n for the diagonal."""
    nx, ny, nz = shape
    n = nx * ny * nz
    idx = np.arange(n, dtype=np.int64)
    i = idx % nx
    j = (idx // nx) % ny
    k = idx // (nx * ny)
    rows = [idx]
    cols = [idx]
    vals = [np.asarray(diag_vals, dtype=np.float64)]
    for (dx, dy, dz), v in zip(offsets, off_vals):
        ii, jj, kk = i + dx, j + dy, k + dz
        m = (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny) & (kk >= 0) & (kk < nz)
        rows.append(idx[m])
        cols.append((ii + nx * (jj + ny * kk))[m])
        vals.append(np.full(int(m.sum()), v, dtype=np.float64))
    return _csr_from_coo(n, np.concatenate(rows), np.concatenate(cols),
                         np.concatenate(vals))


def laplace2d(nx=64):
    """2D 5-point Dirichlet Laplacian on an nx*nx grid (n = nx^2)."""
    n = nx * nx
    offs = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0)]
    return _stencil((nx, nx, 1), offs, np.full(n, 4.0), [-1.0] * 4)


def laplace3d_scaled(nx=128, seed=2):
    """3D 7-point Laplacian with the diagonal 6*10**(3U), U~U[0,1) (seeded)."""
    n = nx ** 3
    rng = np.random.default_rng(seed)
    diag = 6.0 * 10.0 ** (3.0 * rng.random(n))
    offs = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
    return _stencil((nx, nx, nx), offs, diag, [-1.0] * 6)


def laplace3d(nx=16):
    """Plain 3D 7-point Dirichlet Laplacian (diag 6, off -1)."""
    n = nx ** 3
    offs = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]
    return _stencil((nx, nx, nx), offs, np.full(n, 6.0), [-1.0] * 6)


def convdiff3d(nx=100, pe=0.5, vel=(1.0, 0.5, 0.25)):
    """Nonsymmetric 3D 7-point upwind convection-diffusion operator.

    Diffusion: 6 on the diagonal, -1 on the six neighbours.  Convection with
    w = v / max|v| and cell Peclet number pe, first-order upwind: for a
    positive w_d the upstream neighbour (offset -1 in direction d) gets an
    extra -2*pe*w_d and the diagonal an extra +2*pe*w_d (for negative w_d the
    +1 neighbour is upstream)."""
    n = nx ** 3
    v = np.asarray(vel, dtype=np.float64)
    w = v / np.max(np.abs(v))
    diag = 6.0 + 2.0 * pe * float(np.sum(np.abs(w)))
    offs, vals = [], []
    for d in range(3):
        for s in (-1, 1):
            o = [0, 0, 0]
            o[d] = s
            upstream = (s == -1 and w[d] > 0) or (s == 1 and w[d] < 0)
            vals.append(-1.0 - (2.0 * pe * abs(w[d]) if upstream else 0.0))
            offs.append(tuple(o))
    return _stencil((nx, nx, nx), offs, np.full(n, diag), vals)


def uniform_vector(n, seed, lo=-1.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, n)


def _morton_order(pts, bits=10):
    """Indices that sort 3D points in [0,1)^3 along a Morton (Z-order) curve."""
    q = np.minimum((pts * (1 << bits)).astype(np.int64), (1 << bits) - 1)
    code = np.zeros(len(pts), np.int64)
    for b in range(bits):
        for d in range(3):
            code |= ((q[:, d] >> b) & 1) << (3 * b + d)
    return np.argsort(code, kind="stable")


def knn_laplacian(n, k=8, seed=4, shift=1e-2):
    """Symmetrised kNN graph Laplacian on n uniform points in the unit cube,
    L = D - W + shift*I with unit weights (SPD).  The points are numbered
    along a Morton (Z-order) curve, so graph neighbours fall in nearby rows
    and 16x16 tiles (a symmetric renumbering of the same graph)."""
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 3))
    pts = pts[_morton_order(pts)]
    from scipy.spatial import cKDTree  # generator-only dependency

    tree = cKDTree(pts)
    _, nbr = tree.query(pts, k=k + 1)
    src = np.repeat(np.arange(n), k)
    dst = nbr[:, 1:].reshape(-1)
    a = np.minimum(src, dst)
    b = np.maximum(src, dst)
    key = np.unique(a.astype(np.int64) * n + b)
    a = key // n
    b = key % n
    deg = np.bincount(a, minlength=n) + np.bincount(b, minlength=n)
    rows = np.concatenate([np.arange(n), a, b])
    cols = np.concatenate([np.arange(n), b, a])
    vals = np.concatenate([deg + shift, -np.ones(len(a)), -np.ones(len(a))])
    return _csr_from_coo(n, rows, cols, vals)


def _weighted_choice(rng, size, p):
    """rng.choice(len(p), size=size, p=p) with the same draws and result
    (cdf of p, one uniform per sample, right-sided search), the search done
    on the sorted uniforms so that it walks the cdf in order."""
    cdf = p.cumsum()
    cdf /= cdf[-1]
    u = rng.random(size)
    order = np.argsort(u)
    idx = np.empty(size, np.int64)
    idx[order] = cdf.searchsorted(u[order], side="right")
    return idx


def powerlaw_laplacian(n, target_nnz, exponent=2.5, seed=5, shift=1e-3):
    """Chung-Lu random graph with expected degrees w_i ~ i^(-1/(exponent-1)),
    scaled so that the Laplacian has about target_nnz entries; L = D - A +
    shift*I (SPD).  Edges are sampled by drawing endpoint pairs from the
    degree distribution and removing duplicates and self-loops."""
    rng = np.random.default_rng(seed)
    gamma = 1.0 / (exponent - 1.0)
    w = (np.arange(1, n + 1, dtype=np.float64)) ** (-gamma)
    p = w / w.sum()
    m_edges = max(1, (target_nnz - n) // 2)
    draw = int(m_edges * 1.15) + 16
    a = _weighted_choice(rng, draw, p)
    b = _weighted_choice(rng, draw, p)
    # random relabelling so hubs are spread over the tile rows
    perm = rng.permutation(n)
    a = perm[a]
    b = perm[b]
    keep = a != b
    a, b = a[keep], b[keep]
    lo = np.minimum(a, b).astype(np.int64)
    hi = np.maximum(a, b).astype(np.int64)
    key = np.unique(lo * n + hi)[:m_edges]
    lo = key // n
    hi = key % n
    deg = np.bincount(lo, minlength=n) + np.bincount(hi, minlength=n)
    rows = np.concatenate([np.arange(n), lo, hi])
    cols = np.concatenate([np.arange(n), hi, lo])
    vals = np.concatenate([deg + shift, -np.ones(len(lo)), -np.ones(len(lo))])
    return _csr_from_coo(n, rows, cols, vals)


def write_mtx(path, rowptr, colidx, val, ncols=None, comment=None):
    """Writes a general real coordinate Matrix Market file, rows in order and
    columns in the order given (ascending for these generators)."""
    n = len(rowptr) - 1
    ncols = n if ncols is None else ncols
    rowptr = np.asarray(rowptr)
    rows = np.repeat(np.arange(1, n + 1), np.diff(rowptr))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        if comment:
            for line in comment.splitlines():
                f.write("% " + line + "\n")
        f.write(f"{n} {ncols} {len(val)}\n")
        data = np.column_stack([rows, np.asarray(colidx) + 1])
        lines = [f"{r} {c} {v:.17g}\n" for (r, c), v in zip(data.tolist(), np.asarray(val).tolist())]
        f.writelines(lines)


def csr_matvec(rowptr, colidx, val, x):
    """FP64 y = A x (numpy, used to form b = A x*)."""
    rows = np.repeat(np.arange(len(rowptr) - 1), np.diff(rowptr))
    return np.bincount(rows, weights=val * x[colidx], minlength=len(rowptr) - 1)
