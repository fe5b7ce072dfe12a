"""Python binding of the learned-closure MLP (lh_mlp_* in include/littleharsh_amd.h).

The network, its training (Adam) and its inference run on the GPU in the HIP library; this
module only moves host arrays across the C-ABI.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi

_fp = ctypes.POINTER(ctypes.c_float)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a: np.ndarray):
    return a.ctypes.data_as(_fp)


def standardise(x, mean, scale) -> np.ndarray:
    """The network's input standardisation, in float32: (x - mean) * (1 / scale), the
    reciprocal rounded to float32.  The device-side gathers (predict_grid, the closure
    coupling) compute the same operations, so rows standardised here match theirs exactly."""
    r = np.float32(1.0) / _f32(scale)
    return (_f32(x) - _f32(mean)) * r


class MLP:
    """nin -> nhid (tanh) x nlayers -> nout (linear), batch rows per step."""

    def __init__(self, nin: int, nhid: int = 64, nlayers: int = 4, nout: int = 6, batch: int = 4096,
                 seed: int = 1234):
        self.lib = capi.load_library()
        self.nin, self.nhid, self.nlayers, self.nout, self.batch = nin, nhid, nlayers, nout, batch
        self.h = ctypes.c_void_p()
        st = self.lib.lh_mlp_create(nin, nhid, nlayers, nout, batch, seed, ctypes.byref(self.h))
        self._check(st, "lh_mlp_create")

    def _check(self, st: int, what: str):
        if st != 0:
            msg = self.lib.lh_mlp_last_error(self.h).decode() if self.h else ""
            raise RuntimeError(f"{what}: status {st}: {msg}")

    def close(self):
        if self.h:
            self.lib.lh_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nparams(self) -> int:
        return int(self.lib.lh_mlp_param_count(self.h))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        self._check(self.lib.lh_mlp_get_params(self.h, _p(out)), "lh_mlp_get_params")
        return out

    def set_params(self, params) -> None:
        p = _f32(params)
        assert p.size == self.nparams
        self._check(self.lib.lh_mlp_set_params(self.h, _p(p)), "lh_mlp_set_params")

    def set_data(self, x, y) -> None:
        x, y = _f32(x), _f32(y)
        assert x.shape[1] == self.nin and y.shape[1] == self.nout and x.shape[0] == y.shape[0]
        self._keep = (x, y)
        self._check(self.lib.lh_mlp_set_data(self.h, _p(x), _p(y), x.shape[0]), "lh_mlp_set_data")

    def loss_grad(self, x, y):
        x, y = _f32(x), _f32(y)
        assert x.shape == (self.batch, self.nin) and y.shape == (self.batch, self.nout)
        loss = np.zeros(1, dtype=np.float32)
        grad = np.empty(self.nparams, dtype=np.float32)
        self._check(self.lib.lh_mlp_loss_grad(self.h, _p(x), _p(y), _p(loss), _p(grad)), "lh_mlp_loss_grad")
        return float(loss[0]), grad

    def train(self, steps: int, lr: float = 1e-3, seed: int = 7) -> float:
        loss = np.zeros(1, dtype=np.float32)
        self._check(self.lib.lh_mlp_train(self.h, steps, lr, seed, _p(loss)), "lh_mlp_train")
        return float(loss[0])

    def set_fused(self, on: bool):
        """The closure's shape trains by the fused pass (default) or, with False, the per-layer
        kernels; both give bit-identical results."""
        self._check(self.lib.lh_mlp_set_fused(self.h, 1 if on else 0), "lh_mlp_set_fused")

    def predict(self, x) -> np.ndarray:
        x = _f32(x)
        y = np.empty((x.shape[0], self.nout), dtype=np.float32)
        self._check(self.lib.lh_mlp_predict(self.h, _p(x), x.shape[0], _p(y)), "lh_mlp_predict")
        return y

    def predict_grid(self, g, extra, mean, scale) -> np.ndarray:
        """Outputs (ny, nz, nx, nout) from the nch fields g (nch, ny, nz, nx) gathered on the
        3x3x3 neighbourhood on the device, plus the field `extra` (ny, nz, nx) or None, each
        input standardised as standardise(v, mean, scale)."""
        g = _f32(g)
        nch, ny, nz, nx = g.shape
        ex = None if extra is None else _f32(extra)
        if ex is not None:
            assert ex.shape == (ny, nz, nx)
        mean, scale = _f32(mean), _f32(scale)
        assert mean.shape == scale.shape == (self.nin,)
        y = np.empty((ny, nz, nx, self.nout), dtype=np.float32)
        self._check(self.lib.lh_mlp_predict_grid(self.h, _p(g), nch, None if ex is None else _p(ex), ny, nz, nx,
                                                 _p(mean), _p(scale), _p(y)), "lh_mlp_predict_grid")
        return y
