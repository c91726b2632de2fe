"""MLP surrogate of the MPC runs on the MI355X (BF16 MFMA GEMMs, FP32 master
weights, hand-written forward/backward and Adam; see include/mpcx.h).

Inputs are the parameters of a run (8 by default: C and N radii, bond length,
two interaction energies, conversion probability, two rate constants),
standardized with the training set's mean and spread; outputs are the
time-averaged body-frame density field (flattened (nz, nr) grid) followed by
the propulsion velocity, each scaled by the training set's spread.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, ptr, require_gpu, track


def gemm_bf16(A: np.ndarray, B: np.ndarray, reps: int = 1):
    """C = A @ B with BF16 operands and FP32 accumulation on the MFMA units."""
    require_gpu()
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    Cm = np.empty((M, N), np.float32)
    ms = C.c_double(0)
    check(lib().mpcx_gemm_bf16(C.c_int32(M), C.c_int32(N), C.c_int32(K), ptr(A), ptr(B), ptr(Cm), C.c_int32(reps),
                               C.byref(ms)), "mpcx_gemm_bf16")
    return Cm, ms.value


class Surrogate:
    def __init__(self, n_in=8, n_hidden=256, n_layers=3, n_out=2049, seed=1):
        require_gpu()
        self.h = C.c_void_p()
        check(lib().mpcx_mlp_create(C.byref(self.h), C.c_int32(n_in), C.c_int32(n_hidden), C.c_int32(n_layers),
                                    C.c_int32(n_out), C.c_uint64(seed)), "mpcx_mlp_create")
        track(self)
        lib().mpcx_mlp_num_params.restype = C.c_int64
        self.n_in, self.n_hidden, self.n_layers, self.n_out = n_in, n_hidden, n_layers, n_out
        self.nparam = int(lib().mpcx_mlp_num_params(self.h))
        self.x_mean = np.zeros(n_in)
        self.x_std = np.ones(n_in)
        self.y_mean = np.zeros(n_out)
        self.y_scale = np.ones(n_out)

    def close(self):
        if self.h:
            lib().mpcx_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dims(self):
        return [self.n_in] + [self.n_hidden] * self.n_layers + [self.n_out]

    def params(self) -> np.ndarray:
        a = np.empty(self.nparam, np.float32)
        check(lib().mpcx_mlp_get_params(self.h, ptr(a)), "mpcx_mlp_get_params")
        return a

    def set_params(self, flat: np.ndarray):
        a = np.ascontiguousarray(flat, np.float32)
        assert a.size == self.nparam
        check(lib().mpcx_mlp_set_params(self.h, ptr(a)), "mpcx_mlp_set_params")

    # raw (normalized) interface
    def forward_raw(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((x.shape[0], self.n_out), np.float32)
        check(lib().mpcx_mlp_forward(self.h, C.c_int32(x.shape[0]), ptr(x), ptr(y)), "mpcx_mlp_forward")
        return y

    def loss_grad_raw(self, x, t, w):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        w = np.ascontiguousarray(w, np.float32)
        g = np.empty(self.nparam, np.float32)
        loss = C.c_double(0)
        check(lib().mpcx_mlp_loss_grad(self.h, C.c_int32(x.shape[0]), ptr(x), ptr(t), ptr(w), C.byref(loss), ptr(g)),
              "mpcx_mlp_loss_grad")
        return loss.value, g

    def train_raw(self, x, t, w, batch=64, epochs=100, lr=1e-3, seed=0) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        w = np.ascontiguousarray(w, np.float32)
        losses = np.zeros(epochs)
        check(lib().mpcx_mlp_train(self.h, C.c_int64(x.shape[0]), ptr(x), ptr(t), ptr(w), C.c_int32(batch),
                                   C.c_int32(epochs), C.c_float(lr), C.c_uint64(seed), ptr(losses)), "mpcx_mlp_train")
        return losses

    def train_stats(self):
        ms, fl = C.c_double(0), C.c_double(0)
        check(lib().mpcx_mlp_train_stats(self.h, C.byref(ms), C.byref(fl)), "mpcx_mlp_train_stats")
        return ms.value, fl.value

    # physical-units interface
    def fit(self, params: np.ndarray, targets: np.ndarray, out_weight: np.ndarray | None = None,
            scale_floor: float = 1e-2, **kw) -> np.ndarray:
        """Train on (params, targets) in physical units.  Targets are centred
        on the training mean and divided by their spread, floored at
        `scale_floor` times the median spread so that outputs which hardly vary
        (bins inside a sphere) do not turn sampling noise into large values."""
        params = np.asarray(params, np.float64)
        targets = np.asarray(targets, np.float64)
        self.x_mean = params.mean(axis=0)
        self.x_std = params.std(axis=0) + 1e-12
        self.y_mean = targets.mean(axis=0)
        sd = targets.std(axis=0)
        self.y_scale = np.maximum(sd, scale_floor * np.median(sd) + 1e-12)
        w = np.ones(self.n_out) if out_weight is None else np.asarray(out_weight, np.float64)
        return self.train_raw((params - self.x_mean) / self.x_std, (targets - self.y_mean) / self.y_scale,
                              w / w.sum(), **kw)

    def predict(self, params: np.ndarray) -> np.ndarray:
        x = (np.asarray(params, np.float64) - self.x_mean) / self.x_std
        return self.forward_raw(x).astype(np.float64) * self.y_scale + self.y_mean
