"""Learned tile-precision policy: traces, GPU training and installation.

Traces come from the method's own FP64 runs.  For a system and a list of
iteration counts k, the FP64 single-kernel solver runs k iterations from
x0 = 0, the residual r_k = b - A x_k is formed with the library's SpMV, and
for every tile the library records the features (``mf_policy_features``) and
the label (``mf_policy_labels``: the cheapest level whose product error on
r_k is at most theta * ||A r_k|| / sqrt(tilenum)).  The MLP
(8 -> 32 -> 32 -> 4) is trained on the GPU by ``mf_mlp_train``: Adam on FP32
master weights, forward and backward with the FP16-rounded weights.

The trained parameters ship with the package (``policy_mlp.f32``, 1476
little-endian float32); ``train_default`` regenerates them from the traces of
the three training workloads: the 2D 5-point Laplacian 64x64 (b = A 1), the
3D 7-point Laplacian 128^3 with log-uniform diagonal scaling (seed 2) and the
3D upwind convection-diffusion operator 100^3 (BiCGSTAB).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import gen
from ._lib import MLP_NPARAM, check, lib
from .core import TiledMatrix
from .hip import DeviceArray

THETA = 1e-5
PARAMS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "policy_mlp.f32")


def init_params(seed=0):
    """He-initialised FP32 parameter vector (w1, b1, w2, b2, w3, b3)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(MLP_NPARAM, np.float32)
    o = 0
    for fan_in, fan_out in ((8, 32), (32, 32), (32, 4)):
        p[o:o + fan_in * fan_out] = rng.standard_normal(fan_in * fan_out) * np.sqrt(2.0 / fan_in)
        o += fan_in * fan_out + fan_out
    return p


def forward(params, feat):
    """Logits (n x 4) of the device MLP (FP16 weights, FP32 accumulation)."""
    feat = np.ascontiguousarray(feat, np.float32).reshape(-1, 8)
    n = len(feat)
    d_p = DeviceArray.from_host(np.asarray(params, np.float32))
    d_f = DeviceArray.from_host(feat.ravel())
    d_z = DeviceArray(max(4 * n, 1), np.float32)
    check("mf_mlp_forward", lib.mf_mlp_forward(d_p.ptr, d_f.ptr, n, d_z.ptr, None))
    return d_z.get(4 * n).reshape(n, 4)


def grad(params, feat, labels):
    """(mean cross-entropy, gradient) from the device kernels."""
    feat = np.ascontiguousarray(feat, np.float32).reshape(-1, 8)
    n = len(feat)
    d_p = DeviceArray.from_host(np.asarray(params, np.float32))
    d_f = DeviceArray.from_host(feat.ravel())
    d_l = DeviceArray.from_host(np.asarray(labels, np.int32))
    d_g = DeviceArray(MLP_NPARAM, np.float32)
    d_loss = DeviceArray(1, np.float32)
    check("mf_mlp_grad", lib.mf_mlp_grad(d_p.ptr, d_f.ptr, d_l.ptr, n, d_g.ptr, d_loss.ptr, None))
    return float(d_loss.get()[0]), d_g.get()


def train(feat, labels, params=None, epochs=30, batch=4096, lr=2e-3, seed=0):
    """Minibatch Adam on the device; returns (params, final epoch loss)."""
    feat = np.ascontiguousarray(feat, np.float32).reshape(-1, 8)
    labels = np.ascontiguousarray(labels, np.int32)
    p = init_params(seed) if params is None else np.asarray(params, np.float32)
    d_p = DeviceArray.from_host(p)
    d_f = DeviceArray.from_host(feat.ravel())
    d_l = DeviceArray.from_host(labels)
    loss = C.c_float()
    check("mf_mlp_train", lib.mf_mlp_train(d_p.ptr, d_f.ptr, d_l.ptr, len(labels), int(epochs),
                                           int(batch), float(lr), int(seed), C.byref(loss), None))
    return d_p.get(), float(loss.value)


def features(tm: TiledMatrix, d_v, res_ratio):
    d_f = DeviceArray(max(8 * tm.tilenum, 1), np.float32)
    check("mf_policy_features", lib.mf_policy_features(tm.h, d_v.ptr, float(res_ratio), d_f.ptr, None))
    return d_f.get(8 * tm.tilenum).reshape(tm.tilenum, 8)


def labels(tm: TiledMatrix, d_v, theta=THETA):
    d_l = DeviceArray(max(tm.tilenum, 1), np.int32)
    check("mf_policy_labels", lib.mf_policy_labels(tm.h, d_v.ptr, float(theta), d_l.ptr, None))
    return d_l.get(tm.tilenum)


def install(tm: TiledMatrix, params):
    """Makes `params` the handle's learned policy (MF_POLICY_LEARNED)."""
    d = DeviceArray.from_host(np.asarray(params, np.float32))
    check("mf_policy_set_mlp", lib.mf_policy_set_mlp(tm.h, d.ptr))


def residual(tm: TiledMatrix, b, d_x):
    """r = b - A x (device SpMV, host subtraction) as a device vector."""
    n = tm.rowA
    d_y = tm.vector()
    tm.spmv(d_x, d_y)
    r = np.zeros(tm.npad)
    r[:n] = b[:n] - d_y.get()[:n]
    return tm.vector(r), r


def trace(tm: TiledMatrix, b, iters, method="cg", theta=THETA, sample=None, seed=0):
    """Features and labels of every tile at the residuals after each k in
    `iters` FP64 iterations (x0 = 0).  `sample` caps the tiles kept per
    checkpoint (uniform without replacement)."""
    rng = np.random.default_rng(seed)
    d_b = tm.vector(b)
    bn = np.linalg.norm(b)
    solve = tm.cg if method == "cg" else tm.bicgstab
    F, L = [], []
    for k in iters:
        d_x = tm.vector()
        if k > 0:
            solve(d_b, d_x, allow_not_converged=True, policy="fp64", fixed_iters=int(k))
        d_r, r = residual(tm, b, d_x)
        ratio = np.linalg.norm(r) / bn if bn > 0 else 1.0
        if ratio < 1e-13:
            break
        f = features(tm, d_r, ratio)
        lab = labels(tm, d_r, theta)
        if sample is not None and tm.tilenum > sample:
            idx = rng.choice(tm.tilenum, sample, replace=False)
            f, lab = f[idx], lab[idx]
        F.append(f)
        L.append(lab)
    return np.concatenate(F), np.concatenate(L)


def training_workloads(scale=1.0):
    """The three training systems (matrix CSR, b, method, checkpoints).
    scale < 1 shrinks the grids (quick training in tests)."""
    g2 = max(8, int(round(64 * np.sqrt(scale)))) if scale < 1 else 64
    g3 = max(8, int(round(128 * scale))) if scale < 1 else 128
    g5 = max(8, int(round(100 * scale))) if scale < 1 else 100
    rp, ci, va = gen.laplace2d(g2)
    b1 = gen.csr_matvec(rp, ci, va, np.ones(len(rp) - 1))
    yield "lap2d", (rp, ci, va), b1, "cg", list(range(0, 160, 8))
    rp, ci, va = gen.laplace3d_scaled(g3, seed=2)
    yield "lap3d_scaled", (rp, ci, va), gen.uniform_vector(len(rp) - 1, 2), "cg", list(range(0, 480, 32))
    rp, ci, va = gen.convdiff3d(g5)
    xs = gen.uniform_vector(len(rp) - 1, 3)
    yield "convdiff3d", (rp, ci, va), gen.csr_matvec(rp, ci, va, xs), "bicgstab", list(range(0, 64, 4))


def build_training_set(scale=1.0, sample=20000, theta=THETA, seed=0):
    F, L = [], []
    for name, (rp, ci, va), b, method, iters in training_workloads(scale):
        tm = TiledMatrix(rp, ci, va)
        f, lab = trace(tm, b, iters, method, theta, sample, seed)
        F.append(f)
        L.append(lab)
        tm.free()
    return np.concatenate(F), np.concatenate(L)


def train_default(scale=1.0, epochs=40, seed=0, path=None):
    """Traces of the training workloads -> trained parameters (saved to
    `path` when given).  Returns (params, loss, accuracy, label histogram)."""
    F, L = build_training_set(scale, seed=seed)
    params, loss = train(F, L, epochs=epochs, seed=seed)
    acc = float(np.mean(np.argmax(forward(params, F), axis=1) == L))
    if path:
        params.astype("<f4").tofile(path)
    return params, loss, acc, np.bincount(L, minlength=4)


def load(path=PARAMS_FILE):
    p = np.fromfile(path, dtype="<f4")
    if p.size != MLP_NPARAM:
        raise ValueError(f"{path}: expected {MLP_NPARAM} float32 parameters, found {p.size}")
    return p


def load_or_train(path=PARAMS_FILE):
    """Shipped parameters if present, else train on the GPU (and save)."""
    if os.path.exists(path):
        return load(path)
    params, _, _, _ = train_default(path=path)
    return params

