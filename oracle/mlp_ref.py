"""ORACLE — test infrastructure only (the checker of the learned Barnes-Hut correction).

numpy restatement of the correction network and of its training data, used by ``tests/`` to
check the HIP kernels; the product library never imports it.  No reference counterpart: the
network corrects the dual-monopole approximation of src/interaction/barnes-hut.cc:183-221.

* ``bf16_round``: float32 -> bfloat16 round-to-nearest-even, pinned by known values (ties to even,
  overflow to inf, NaN stays NaN).
* ``forward`` / ``loss_and_grad``: FP64 forward and hand-written backward of
  16 -> 64 -> 64 -> 3 with SiLU and the mean squared error, with optional bf16 rounding of the
  layer inputs and weights (gradients pass straight through the rounding).  Pinned by central
  finite differences with rounding off (the rounded variant applies the pinned ``bf16_round`` at
  the same points).
* ``quadrupole_from_features``: the same quadrupole term from a row's features in label units,
  pinned against ``quadrupole_residual`` on particle clusters.
* ``node_features`` / ``node_residual``: features and exact label of one (target, node, sign) row
  from the node's particles, by direct summation.  Pinned by the multipole identities: the
  residual of a node whose sign-s charges sit at one point is zero, and the quadrupole formula
  reproduces the residual of a far node to O((h/R)^3).
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 16, 64, 3
SHAPES = (("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_H, N_OUT)),
          ("b3", (N_OUT,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(w):
    out, off = {}, 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = np.asarray(w[off:off + size], np.float64).reshape(shape)
        off += size
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16_round(a):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).astype(np.float64)
    return np.where(np.isnan(a), np.nan, out)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def forward(w, X, rounding=True):
    """Network output for normalised inputs X [rows][16]; returns (y, cache)."""
    p = unpack(w)
    r = bf16_round if rounding else (lambda a: np.asarray(a, np.float64))
    x0 = r(X)
    W1, W2, W3 = r(p["W1"]), r(p["W2"]), r(p["W3"])
    z1 = x0 @ W1 + p["b1"]
    a1 = r(_silu(z1))
    z2 = a1 @ W2 + p["b2"]
    a2 = r(_silu(z2))
    y = a2 @ W3 + p["b3"]
    return y, (x0, W1, W2, W3, z1, a1, z2, a2)


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def loss_and_grad(w, X, Y, rounding=True):
    """Mean squared error over rows x 3 outputs and its gradient (flat, FP64)."""
    y, (x0, W1, W2, W3, z1, a1, z2, a2) = forward(w, X, rounding)
    d = y - np.asarray(Y, np.float64)
    count = d.size
    loss = float((d * d).sum() / count)
    dy = 2.0 * d / count
    g = {"W3": a2.T @ dy, "b3": dy.sum(0)}
    dz2 = (dy @ W3.T) * _dsilu(z2)
    g["W2"] = a1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * _dsilu(z1)
    g["W1"] = x0.T @ dz1
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def second_moments(pos, q):
    """Total charge, charge centre and second moments (xx yy zz xy xz yz) about it of charges q > 0
    at pos (units)."""
    Q = q.sum()
    c = (q[:, None] * pos).sum(0) / Q
    d = pos - c
    S = np.array([(q * d[:, 0] * d[:, 0]).sum(), (q * d[:, 1] * d[:, 1]).sum(), (q * d[:, 2] * d[:, 2]).sum(),
                  (q * d[:, 0] * d[:, 1]).sum(), (q * d[:, 0] * d[:, 2]).sum(), (q * d[:, 1] * d[:, 2]).sum()])
    return Q, c, S


def node_residual(target, pos, q, sign):
    """Exact force of the node's sign-s charges on target minus their monopole at the charge centre,
    sum q_j (t - x_j)/|t - x_j|^3 in e/unit^2 (pos in units, q in e; the target itself excluded by
    the caller)."""
    sel = q > 0 if sign == 0 else q < 0
    p, qq = pos[sel], q[sel]
    Q, c, _ = second_moments(p, np.abs(qq))
    d = target - p
    r3 = np.linalg.norm(d, axis=1) ** 3
    exact = (qq[:, None] * d / r3[:, None]).sum(0)
    rc = target - c
    mono = (Q if sign == 0 else -Q) * rc / np.linalg.norm(rc) ** 3
    return exact - mono


def quadrupole_residual(target, pos, q, sign):
    """Leading (quadrupole) term of node_residual for the charges of one sign about their centre:
    the potential phi_2 = sigma r.Qt.r / (2 R^5) with the traceless tensor Qt = 3 S - tr(S) I
    (the dipole vanishes about the charge centre), and F = -grad phi_2 per unit target charge."""
    sel = q > 0 if sign == 0 else q < 0
    p, qq = pos[sel], np.abs(q[sel])
    Q, c, S6 = second_moments(p, qq)
    S = np.array([[S6[0], S6[3], S6[4]], [S6[3], S6[1], S6[5]], [S6[4], S6[5], S6[2]]])
    Qt = 3.0 * S - np.trace(S) * np.eye(3)
    r = target - c
    R = np.linalg.norm(r)
    # phi_2 = r.Qt.r / (2 R^5); grad = Qt.r / R^5 - 5/2 (r.Qt.r) r / R^7
    grad = Qt @ r / R ** 5 - 2.5 * (r @ Qt @ r) * r / R ** 7
    sigma = 1.0 if sign == 0 else -1.0
    return -sigma * grad


def quadrupole_from_features(X):
    """quadrupole_residual in label units (divided by sigma Q h^2 / R^4), from raw features:
    u = X[:, 0:3], S/(Q h^2) = X[:, 7:13]."""
    X = np.asarray(X, np.float64)
    u = X[:, 0:3]
    S6 = X[:, 7:13]
    S = np.zeros((len(X), 3, 3))
    for k, (a, b) in enumerate(((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))):
        S[:, a, b] = S6[:, k]
        S[:, b, a] = S6[:, k]
    Qt = 3.0 * S - np.trace(S, axis1=1, axis2=2)[:, None, None] * np.eye(3)
    Qu = np.einsum("nij,nj->ni", Qt, u)
    uQu = (u * Qu).sum(1)
    return -(Qu - 2.5 * uQu[:, None] * u)


def node_features(target, pos, q, sign, centre, half):
    """The 16 raw features of a row (13 used, 3 zero): r/R, h/R, (c_s - c_geo)/h, S_s/(Q_s h^2)."""
    sel = q > 0 if sign == 0 else q < 0
    Q, c, S = second_moments(pos[sel], np.abs(q[sel]))
    r = target - c
    R = np.linalg.norm(r)
    f = np.zeros(16)
    f[0:3] = r / R
    f[3] = half / R
    f[4:7] = (c - centre) / half
    f[7:13] = S / (Q * half * half)
    return f, Q, R
