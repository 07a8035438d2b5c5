"""Learned far-field correction of Barnes-Hut (am_mlp_* and am_compute_barnes_hut_learned_forces).

The Barnes-Hut walk (src/interaction/barnes-hut.cc:183-221) replaces an accepted node by one point
charge per sign at that sign's charge centre.  The network predicts the residual of that
approximation per (target, node, sign) from 13 features of the node's sign-s charges

    r/R (3), h/R, (c_s - c_geo)/h (3), S_s / (Q_s h^2) (6: xx yy zz xy xz yz)

(r = target - c_s, R = |r|, h = half width, S_s the second moments about c_s), padded to 16.  The
label is the exact residual divided by sigma_s Q_s h^2 / R^4 (units e and 2^-20 m).  Rows whose
prediction exceeds a bound, or whose h/R exceeds amax, take the exact residual over the node's
particles instead.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import Context, Store, _check, _p, lib

N_IN = 16
N_FEATURES = 13
N_HIDDEN = 64
N_OUT = 3


def num_params() -> int:
    return int(lib().am_mlp_num_params())


def param_slices():
    """Offsets of W1 [16][64], b1, W2 [64][64], b2, W3 [64][3], b3 in the flat parameter vector."""
    out, off = {}, 0
    for name, shape in (("W1", (N_IN, N_HIDDEN)), ("b1", (N_HIDDEN,)), ("W2", (N_HIDDEN, N_HIDDEN)),
                        ("b2", (N_HIDDEN,)), ("W3", (N_HIDDEN, N_OUT)), ("b3", (N_OUT,))):
        size = int(np.prod(shape))
        out[name] = (off, shape)
        off += size
    return out


def quadrupole_baseline(X) -> np.ndarray:
    """The analytic quadrupole term y_q of raw feature rows (what the network adds its output to)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.zeros((len(X), N_OUT), np.float32)
    _check(lib().am_mlp_quadrupole_baseline(_p(X), len(X), _p(Y)))
    return Y


class Mlp:
    """The correction network on one device (16 -> 64 -> 64 -> 3, SiLU).  With quadrupole=True it
    predicts the analytic quadrupole term plus the network output (the network then learns the
    remainder); by default the network predicts the whole residual."""

    def __init__(self, ctx: Context, seed: int = 1, quadrupole: bool = False):
        h = ctypes.c_void_p()
        _check(lib().am_mlp_create(ctx.h, int(seed), ctypes.byref(h)))
        self.h = h
        self.ctx = ctx
        self.quadrupole = bool(quadrupole)
        _check(lib().am_mlp_set_baseline(h, int(self.quadrupole)))
        self.in_mean = np.zeros(N_IN, np.float32)
        self.in_std = np.ones(N_IN, np.float32)
        self.out_scale = np.ones(N_OUT, np.float32)
        ctx._children.add(self)

    def close(self):
        if self.h:
            if self.ctx.h:
                lib().am_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        w = np.zeros(num_params(), np.float32)
        _check(lib().am_mlp_get_params(self.h, _p(w)))
        return w

    @params.setter
    def params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != num_params():
            raise ValueError(f"expected {num_params()} parameters")
        _check(lib().am_mlp_set_params(self.h, _p(w)))

    def set_normalization(self, in_mean, in_std, out_scale):
        self.in_mean = np.ascontiguousarray(in_mean, dtype=np.float32)
        self.in_std = np.ascontiguousarray(in_std, dtype=np.float32)
        self.out_scale = np.ascontiguousarray(out_scale, dtype=np.float32)
        _check(lib().am_mlp_set_normalization(self.h, _p(self.in_mean), _p(self.in_std), _p(self.out_scale)))

    def fit_normalization(self, X, Y):
        """Standardise each feature over the data (unit std for constant columns) and scale the
        outputs by the standard deviation of what the network learns (Y - y_q, or Y)."""
        if self.quadrupole:
            Y = np.asarray(Y, np.float64) - quadrupole_baseline(X)
        X = np.asarray(X, np.float64)
        Y = np.asarray(Y, np.float64)
        mean = X.mean(0)
        std = X.std(0)
        std[std < 1e-6] = 1.0
        scale = Y.std(0)
        scale[scale <= 0] = 1.0
        self.set_normalization(mean, std, scale)

    def gradients(self, Xn, Yn):
        """MSE gradient on normalised inputs/targets (no update): (grad, loss)."""
        Xn = np.ascontiguousarray(Xn, dtype=np.float32)
        Yn = np.ascontiguousarray(Yn, dtype=np.float32)
        g = np.zeros(num_params(), np.float32)
        loss = ctypes.c_float()
        _check(lib().am_mlp_gradients(self.h, _p(Xn), _p(Yn), len(Xn), _p(g), ctypes.byref(loss)))
        return g, loss.value

    def train(self, X, Y, epochs=10, batch=1024, lr=1e-3, seed=0):
        """Adam on raw features / labels; returns the mean loss per epoch."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        losses = np.zeros(epochs, np.float32)
        _check(lib().am_mlp_train(self.h, _p(X), _p(Y), len(X), int(epochs), int(batch), float(lr), int(seed),
                                  _p(losses)))
        return losses

    def infer(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), N_OUT), np.float32)
        _check(lib().am_mlp_infer(self.h, _p(X), len(X), _p(Y)))
        return Y


def training_data(store: Store, theta: float, n_targets: int, stride: int = 1, offset: int = 0, rec_cap: int = 1024):
    """(X [rows][16], Y [rows][3], target index, node preorder position, label scale) from the store's
    cloud."""
    cap = 2 * n_targets * rec_cap
    X = np.zeros((cap, N_IN), np.float32)
    Y = np.zeros((cap, N_OUT), np.float32)
    T = np.zeros(cap, np.int32)
    N = np.zeros(cap, np.int32)
    S = np.zeros(cap, np.float32)
    rows = ctypes.c_int64()
    _check(lib().am_bh_learned_training_data(store.h, float(theta), int(n_targets), int(stride), int(offset),
                                             int(rec_cap), _p(X), _p(Y), _p(T), _p(N), _p(S),
                                             ctypes.byref(rows)))
    r = rows.value
    return X[:r].copy(), Y[:r].copy(), T[:r].copy(), N[:r].copy(), S[:r].copy()


def barnes_hut_learned_forces(store: Store, theta: float, net: Mlp, rec_cap: int = 1024, ybound: float = 8.0,
                              amax: float = 1.0):
    """ComputeBarnesHutCoulombForces (FP32 walk) with the learned correction; returns the counts
    (rows by the model, rows that fell back, particles whose node list overflowed rec_cap)."""
    st = np.zeros(3, np.int64)
    _check(lib().am_compute_barnes_hut_learned_forces(store.h, float(theta), net.h, int(rec_cap), float(ybound),
                                                      float(amax), _p(st)))
    return dict(model_rows=int(st[0]), fallback_rows=int(st[1]), overflowed=int(st[2]))


def train_on_clouds(ctx: Context, clouds, theta: float, net: Mlp, targets_per_cloud: int = 1024, rec_cap: int = 1024,
                    epochs: int = 10, batch: int = 1024, lr: float = 1e-3, seed: int = 0):
    """Generates training rows on each (pos, vel, species) cloud with the exact residuals, fits the
    normalisation on them and trains the network.  Returns (rows, per-epoch losses)."""
    xs, ys = [], []
    for pos, vel, sp in clouds:
        s = Store(ctx, len(sp))
        s.upload(pos, vel, sp)
        m = len(sp)
        stride = max(1, m // targets_per_cloud)
        X, Y, _, _, _ = training_data(s, theta, targets_per_cloud, stride, stride // 2, rec_cap)
        xs.append(X)
        ys.append(Y)
        s.close()
    X = np.concatenate(xs)
    Y = np.concatenate(ys)
    net.fit_normalization(X, Y)
    # The error bound of the correction: a prediction larger than almost every label seen in training
    # (|y| above the 99.9th percentile) is not trusted, and the row takes the exact residual.
    mags = np.linalg.norm(Y, axis=1)
    net.label_bound = float(np.quantile(mags, 0.999)) if len(mags) else 0.0
    net.max_h_over_r = float(X[:, 3].max()) if len(X) else 0.0
    losses = net.train(X, Y, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return len(X), losses
