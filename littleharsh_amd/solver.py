"""Host-side driver of one channel on the GPU through the C-ABI (include/littleharsh_amd.h).

Mirrors the reference host loop (littleharsh.f90:160-201): ``step`` is one pass of the
``do kRK = 1,3`` loop plus meanflow_ctP; ``substep`` is one body of that loop.  State
arrays use the reference layout complex(8) f(j, column): numpy complex128 of shape
(ncol, nplanes).  Any non-zero status raises with the library's message, the analogue
of the reference's ``write(*,*) ...; stop``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi
from .channel import ChannelSetup

_dp = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class LHError(RuntimeError):
    """A non-zero status of the C-ABI; ``status`` holds it (LH_ERR_ARG, LH_ERR_HIP,
    LH_ERR_NUMERIC)."""

    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


class Channel:
    def __init__(self, setup: ChannelSetup, precision: int = capi.LH_PREC_FP64):
        self.s = setup
        self._lib = capi.load_library()
        cfg, keep = capi.make_config(setup, precision)
        h = ctypes.c_void_p()
        r = self._lib.lh_create(ctypes.byref(cfg), ctypes.byref(h))
        self._h = h
        if r != capi.LH_OK:
            msg = self._lib.lh_last_error(h).decode() if h.value else "lh_create failed"
            if h.value:
                self._lib.lh_destroy(h)
            self._h = None
            raise LHError(f"lh_create: {msg} (status {r})")
        del keep

    def close(self):
        """Destroy the handle (a batch holding it must be closed first: lh_destroy refuses,
        and the handle then stays alive)."""
        if getattr(self, "_h", None):
            if self._lib.lh_destroy(self._h) == capi.LH_OK:
                self._h = None

    def __del__(self):
        self.close()

    def _chk(self, r, what):
        if r != capi.LH_OK:
            raise LHError(f"{what}: {self._lib.lh_last_error(self._h).decode()} (status {r})", r)

    # ---- state
    def set_state(self, u, v, w, p):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (u, v, w, p)]
        s = self.s
        for a, n in zip(arrs, (s.nyu, s.nyv, s.nyu, s.nyp)):
            if a.shape != (s.ncol, n):
                raise ValueError(f"state array shape {a.shape} != {(s.ncol, n)}")
        self._chk(self._lib.lh_set_state(self._h, *[_p(a) for a in arrs]), "lh_set_state")

    def get_state(self):
        s = self.s
        u = np.zeros((s.ncol, s.nyu), complex)
        v = np.zeros((s.ncol, s.nyv), complex)
        w = np.zeros((s.ncol, s.nyu), complex)
        p = np.zeros((s.ncol, s.nyp), complex)
        self._chk(self._lib.lh_get_state(self._h, _p(u), _p(v), _p(w), _p(p)), "lh_get_state")
        return u, v, w, p

    def nonlinear(self, krk):
        """The nonlinear stage of substep krk (stage-level hook)."""
        self._chk(self._lib.lh_nonlinear(self._h, int(krk)), "lh_nonlinear")

    def solve_u(self, krk, dt):
        """RHS0, the canopy forcing and solveU/V/U of substep krk (stage-level hook)."""
        self._chk(self._lib.lh_solve_u(self._h, int(krk), float(dt)), "lh_solve_u")

    def solve_p(self, krk, dt):
        """meanflow_ctU, solveP and v_corr of substep krk (stage-level hook)."""
        self._chk(self._lib.lh_solve_p(self._h, int(krk), float(dt)), "lh_solve_p")

    def get_nonlinear(self):
        s = self.s
        n1 = np.zeros((s.ncol, s.nyu - 2), complex)
        n2 = np.zeros((s.ncol, s.nyv - 2), complex)
        n3 = np.zeros((s.ncol, s.nyu - 2), complex)
        self._chk(self._lib.lh_get_nonlinear(self._h, _p(n1), _p(n2), _p(n3)), "lh_get_nonlinear")
        return n1, n2, n3

    def set_nonlinear(self, n1, n2, n3):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (n1, n2, n3)]
        self._chk(self._lib.lh_set_nonlinear(self._h, *[_p(a) for a in arrs]), "lh_set_nonlinear")

    def set_forcing(self, f1=None, f2=None, f3=None):
        """Forcing added to Nu_new in every substep (layout of set_nonlinear); no argument
        removes it."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.complex128) for a in (f1, f2, f3)]
        ptrs = [None if a is None else _p(a) for a in arrs]
        self._chk(self._lib.lh_set_forcing(self._h, *ptrs), "lh_set_forcing")

    def get_forcing(self):
        s = self.s
        f1 = np.zeros((s.ncol, s.nyu - 2), complex)
        f2 = np.zeros((s.ncol, s.nyv - 2), complex)
        f3 = np.zeros((s.ncol, s.nyu - 2), complex)
        self._chk(self._lib.lh_get_forcing(self._h, _p(f1), _p(f2), _p(f3)), "lh_get_forcing")
        return f1, f2, f3

    # ---- time stepping
    def substep(self, krk: int, dt: float):
        self._chk(self._lib.lh_substep(self._h, int(krk), float(dt)), "lh_substep")

    def step(self, dt: float):
        self._chk(self._lib.lh_step(self._h, float(dt)), "lh_step")

    def step_cfl(self, cfl: float) -> float:
        dt = ctypes.c_double()
        self._chk(self._lib.lh_step_cfl(self._h, float(cfl), ctypes.byref(dt)), "lh_step_cfl")
        return dt.value

    def scalars(self):
        m, q, d, c = (ctypes.c_double() for _ in range(4))
        self._chk(self._lib.lh_get_scalars(self._h, ctypes.byref(m), ctypes.byref(q),
                                           ctypes.byref(d), ctypes.byref(c)), "lh_get_scalars")
        return dict(mpgx=m.value, qx=q.value, dt=d.value, cflmax=c.value)

    def max_div(self) -> float:
        e = ctypes.c_double()
        self._chk(self._lib.lh_max_div(self._h, ctypes.byref(e)), "lh_max_div")
        return e.value

    def cfl_dt(self, cfl: float) -> float:
        dt = ctypes.c_double()
        self._chk(self._lib.lh_cfl_dt(self._h, float(cfl), ctypes.byref(dt)), "lh_cfl_dt")
        return dt.value

    # ---- statistics (stats.f90) and spectra (spectra.f90)
    def stats_accumulate(self):
        self._chk(self._lib.lh_stats_accumulate(self._h), "lh_stats_accumulate")

    def stats(self, reset: bool = False):
        """(13, nyu) per-plane sums (write_stats row order) and the sample count."""
        a = np.zeros((13, self.s.nyu))
        n = ctypes.c_int()
        self._chk(self._lib.lh_stats_read(self._h, _p(a), ctypes.byref(n), 1 if reset else 0),
                  "lh_stats_read")
        return a, n.value

    def cond_stats_flat(self, reset: bool = False):
        n = ctypes.c_longlong()
        self._chk(self._lib.lh_cond_stats_size(self._h, ctypes.byref(n)), "lh_cond_stats_size")
        a = np.zeros(n.value)
        self._chk(self._lib.lh_cond_stats_read(self._h, _p(a), n.value, 1 if reset else 0),
                  "lh_cond_stats_read")
        return a

    def spectra_accumulate(self):
        self._chk(self._lib.lh_spectra_accumulate(self._h), "lh_spectra_accumulate")

    def spectra(self, reset: bool = False):
        """(5, ncol, nyu): U, V, W, UV, P in the reference layout sp(j, column)."""
        a = np.zeros((5, self.s.ncol, self.s.nyu))
        self._chk(self._lib.lh_spectra_read(self._h, _p(a), 1 if reset else 0), "lh_spectra_read")
        return a

    # ---- launch parameters (results do not depend on them)
    def set_launch(self, xraw_lc: int, zasm_lc: int, col_bs: int):
        self._chk(self._lib.lh_set_launch(self._h, int(xraw_lc), int(zasm_lc), int(col_bs)),
                  "lh_set_launch")

    def get_launch(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(self._lib.lh_get_launch(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                  "lh_get_launch")
        return dict(xraw_lc=a.value, zasm_lc=b.value, col_bs=c.value)

    def set_plans(self, xplan: int, zplan: int):
        """FFT plan of the x and of the z transforms (0 = the default radices; 1 = the
        alternative that lengths 48, 72, 96 and 144 have)."""
        self._chk(self._lib.lh_set_plans(self._h, int(xplan), int(zplan)), "lh_set_plans")

    def get_plans(self):
        v = [ctypes.c_int() for _ in range(4)]
        self._chk(self._lib.lh_get_plans(self._h, *(ctypes.byref(x) for x in v)), "lh_get_plans")
        return dict(xplan=v[0].value, zplan=v[1].value, xcount=v[2].value, zcount=v[3].value)

    def autotune(self, steps: int = 2):
        """Times every pair of FFT plans, then the launch parameters; returns the choice."""
        self._chk(self._lib.lh_autotune(self._h, int(steps)), "lh_autotune")
        p = self.get_plans()
        return dict(self.get_launch(), xplan=p["xplan"], zplan=p["zplan"])

    # ---- per-kernel timing (HIP events on the handle's stream)
    def timing(self, enable):
        """False/0 off, True/1 per-kernel events (direct issue), 2 per-step events around
        the graph replays."""
        self._chk(self._lib.lh_timing_enable(self._h, int(enable)), "lh_timing_enable")

    def timing_reset(self):
        self._chk(self._lib.lh_timing_reset(self._h), "lh_timing_reset")

    def timing_read(self):
        out = {}
        for k, name in enumerate(capi.KERNEL_NAMES):
            ms, n, by = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
            self._chk(self._lib.lh_timing_read(self._h, k, ctypes.byref(ms), ctypes.byref(n),
                                               ctypes.byref(by)), "lh_timing_read")
            out[name] = dict(ms=ms.value, launches=n.value, bytes_per_launch=by.value)
        return out


class Batch:
    """Independent channels advanced together: one CFL step of every instance per call, issued
    as one HIP graph whose branches are the instances (lh_batch_step_cfl)."""

    def __init__(self, channels):
        self.channels = list(channels)
        self._lib = capi.load_library()
        arr = (ctypes.c_void_p * len(self.channels))(*[c._h.value for c in self.channels])
        b = ctypes.c_void_p()
        r = self._lib.lh_batch_create(len(self.channels), arr, ctypes.byref(b))
        if r != capi.LH_OK:
            raise LHError(f"lh_batch_create failed (status {r})", r)
        self._b = b

    def step_cfl(self, cfl: float):
        r = self._lib.lh_batch_step_cfl(self._b, float(cfl))
        if r != capi.LH_OK:
            raise LHError(f"lh_batch_step_cfl: {self._lib.lh_batch_last_error(self._b).decode()} "
                          f"(status {r})", r)

    def close(self):
        if getattr(self, "_b", None):
            self._lib.lh_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        self.close()
