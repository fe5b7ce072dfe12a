"""ctypes front end of the CPU oracle (oracle/lh_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker of the HIP library, never by the product.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from littleharsh_amd.capi import make_config, LHConfig

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liblh_oracle.so")
_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        L = ctypes.CDLL(SO)
        vp = ctypes.c_void_p
        L.lho_create.argtypes = [ctypes.POINTER(LHConfig)]
        L.lho_create.restype = vp
        L.lho_destroy.argtypes = [vp]
        L.lho_set_state.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.lho_get_state.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.lho_set_nonlinear.argtypes = [vp, _dp, _dp, _dp]
        L.lho_set_forcing.argtypes = [vp, _dp, _dp, _dp]
        L.lho_get_nonlinear.argtypes = [vp, _dp, _dp, _dp]
        L.lho_get_scalars.argtypes = [vp, _dp, _dp, _dp]
        L.lho_substep.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.lho_step.argtypes = [vp, ctypes.c_double]
        L.lho_max_div.argtypes = [vp]
        L.lho_max_div.restype = ctypes.c_double
        L.lho_flowrate.argtypes = [vp]
        L.lho_flowrate.restype = ctypes.c_double
        L.lho_four_to_phys.argtypes = [vp, _dp, _dp]
        L.lho_phys_to_four.argtypes = [vp, _dp, _dp]
        L.lho_cfft.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.lho_stats_accumulate.argtypes = [vp]
        L.lho_spectra_accumulate.argtypes = [vp]
        L.lho_stats_read.argtypes = [vp, _dp]
        L.lho_stats_read.restype = ctypes.c_int
        L.lho_cond_stats_read.argtypes = [vp, _dp, ctypes.c_longlong]
        L.lho_cond_stats_read.restype = ctypes.c_longlong
        L.lho_spectra_read.argtypes = [vp, _dp]
        L.lho_hook_rhs0.argtypes = [vp, ctypes.c_int]
        L.lho_get_work.argtypes = [vp, ctypes.c_int, _dp]
        L.lho_set_nonlinear_full.argtypes = [vp, _dp, _dp, _dp]
        L.lho_hook_solvep_vcorr.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.lho_hook_substep_vel.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.lho_hook_substep_pres.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.lho_hook_lu_vel.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.lho_get_dg.argtypes = [vp, _dp]
        L.lho_hook_flowrate_corr.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        L.lho_hook_flowrate_corr.restype = ctypes.c_double
        L.lho_hook_nonlinear.argtypes = [vp, ctypes.c_int]
        L.lho_get_nonlinear_full.argtypes = [vp, _dp, _dp, _dp]
        L.lho_hook_ib.argtypes = [vp, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.lho_hook_solve_vel.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class Oracle:
    """One channel advanced on the CPU by the restatement in lh_oracle.c."""

    def __init__(self, setup):
        self.s = setup
        cfg, keep = make_config(setup)
        self._h = lib().lho_create(ctypes.byref(cfg))
        if not self._h:
            raise RuntimeError("lho_create failed (unsupported FFT length?)")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().lho_destroy(self._h)
            self._h = None

    def set_state(self, u, v, w, p):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (u, v, w, p)]
        lib().lho_set_state(self._h, *[_p(a) for a in arrs])

    def get_state(self):
        s = self.s
        u = np.zeros((s.ncol, s.nyu), complex)
        v = np.zeros((s.ncol, s.nyv), complex)
        w = np.zeros((s.ncol, s.nyu), complex)
        p = np.zeros((s.ncol, s.nyp), complex)
        lib().lho_get_state(self._h, _p(u), _p(v), _p(w), _p(p))
        return u, v, w, p

    def get_nonlinear(self):
        s = self.s
        n1 = np.zeros((s.ncol, s.nyu - 2), complex)
        n2 = np.zeros((s.ncol, s.nyv - 2), complex)
        n3 = np.zeros((s.ncol, s.nyu - 2), complex)
        lib().lho_get_nonlinear(self._h, _p(n1), _p(n2), _p(n3))
        return n1, n2, n3

    def set_nonlinear(self, n1, n2, n3):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (n1, n2, n3)]
        lib().lho_set_nonlinear(self._h, *[_p(a) for a in arrs])

    def set_forcing(self, f1=None, f2=None, f3=None):
        """Forcing added to Nu_new in every substep (layout of set_nonlinear); no argument
        removes it."""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.complex128) for a in (f1, f2, f3)]
        self._forcing = arrs
        lib().lho_set_forcing(self._h, *[_p(a) for a in arrs])

    def scalars(self):
        m, q, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        lib().lho_get_scalars(self._h, ctypes.byref(m), ctypes.byref(q), ctypes.byref(c))
        return dict(mpgx=m.value, qx=q.value, cflmax=c.value)

    def substep(self, krk, dt):
        r = lib().lho_substep(self._h, krk, dt)
        if r:
            raise RuntimeError(f"lho_substep failed ({r})")

    def step(self, dt):
        r = lib().lho_step(self._h, dt)
        if r:
            raise RuntimeError(f"lho_step failed ({r})")

    # ---- routine-level hooks (checks against the reference's own routines)
    def rhs0(self, krk, nu_full):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in nu_full]
        lib().lho_set_nonlinear_full(self._h, *[_p(a) for a in arrs])
        lib().lho_hook_rhs0(self._h, krk)
        s = self.s
        out = []
        for which, n in zip(range(6), (s.nyu, s.nyv, s.nyu, s.nyu, s.nyv, s.nyu)):
            a = np.zeros((s.ncol, n), complex)
            lib().lho_get_work(self._h, which, _p(a))
            out.append(a)
        return out

    def work(self, which):
        s = self.s
        n = {6: s.nyp, 7: s.nyp}[which]
        a = np.zeros((s.ncol, n), complex)
        lib().lho_get_work(self._h, which, _p(a))
        return a

    def solvep_vcorr(self, krk, dt):
        lib().lho_hook_solvep_vcorr(self._h, krk, dt)

    def substep_vel(self, krk, dt):
        """RHS0, nonlinear and solveU/V/U of substep krk (littleharsh.f90:166-175)."""
        lib().lho_hook_substep_vel(self._h, krk, dt)

    def substep_pres(self, krk, dt):
        """meanflow_ctU (flag_ctpress = 0), solveP and v_corr of substep krk
        (littleharsh.f90:178-188)."""
        lib().lho_hook_substep_pres(self._h, krk, dt)

    def lu_vel(self, grid_u, krk, dt):
        nj = self.s.nyu if grid_u else self.s.nyv
        ay, axz = np.zeros((nj, 3)), np.zeros(self.s.ncol)
        lib().lho_hook_lu_vel(self._h, 1 if grid_u else 0, krk, dt, _p(ay), _p(axz))
        return ay, axz

    def dg(self):
        a = np.zeros((self.s.ncol, self.s.nyp, 3))
        lib().lho_get_dg(self._h, _p(a))
        return a

    def flowrate_corr(self, krk, dt, qxt, mpgx):
        m = ctypes.c_double(mpgx)
        g = lib().lho_hook_flowrate_corr(self._h, krk, dt, qxt, ctypes.byref(m))
        return g, m.value

    def nonlinear(self, krk=1):
        """Nu_new (full plane ranges, reference layout) of the current state."""
        s = self.s
        lib().lho_hook_nonlinear(self._h, krk)
        n1 = np.zeros((s.ncol, s.nyu), complex)
        n2 = np.zeros((s.ncol, s.nyv), complex)
        n3 = np.zeros((s.ncol, s.nyu), complex)
        lib().lho_get_nonlinear_full(self._h, _p(n1), _p(n2), _p(n3))
        return n1, n2, n3

    def ib(self, grid_u, beta, x, lu):
        x = np.array(x, dtype=np.complex128, order="C", copy=True)
        lu = np.ascontiguousarray(lu, dtype=np.complex128)
        lib().lho_hook_ib(self._h, 1 if grid_u else 0, beta, _p(x), _p(lu))
        return x

    def solve_vel(self, grid_u, krk, dt, u, du, lu):
        """solveU / solveV (littleharsh.f90:733-824): the new velocity from u, du and the
        y-Laplacian lu (used on canopy tip planes); inputs are not modified."""
        u = np.array(u, dtype=np.complex128, order="C", copy=True)
        du = np.array(du, dtype=np.complex128, order="C", copy=True)
        lu = np.ascontiguousarray(lu, dtype=np.complex128)
        lib().lho_hook_solve_vel(self._h, 1 if grid_u else 0, int(krk), float(dt), _p(u), _p(du), _p(lu))
        return u

    def spectra(self):
        s = self.s
        a = np.zeros((5, s.ncol, s.nyu))
        lib().lho_spectra_read(self._h, _p(a))
        return a

    def spectra_accumulate(self):
        lib().lho_spectra_accumulate(self._h)

    def stats_accumulate(self):
        lib().lho_stats_accumulate(self._h)

    def stats(self):
        """(13, nyu) per-plane sums (rows in write_stats order, slot-indexed per grid) and
        the number of accumulated samples."""
        a = np.zeros((13, self.s.nyu))
        n = lib().lho_stats_read(self._h, _p(a))
        return a, n

    def cond_stats(self):
        """13 conditional arrays, each (nb, dnz, dnx) = Fortran (dnx, dnz, nb)."""
        n = lib().lho_cond_stats_read(self._h, None, 0)
        flat = np.zeros(n)
        lib().lho_cond_stats_read(self._h, _p(flat), n)
        return split_cond(self.s, flat)

    def max_div(self):
        return lib().lho_max_div(self._h)

    def flowrate(self):
        return lib().lho_flowrate(self._h)

    def four_to_phys(self, sp):
        sp = np.ascontiguousarray(sp, dtype=np.complex128)
        ph = np.zeros((self.s.pz, self.s.px))
        lib().lho_four_to_phys(self._h, _p(sp), _p(ph))
        return ph

    def phys_to_four(self, ph):
        ph = np.ascontiguousarray(ph, dtype=np.float64)
        sp = np.zeros(self.s.ncol, complex)
        lib().lho_phys_to_four(self._h, _p(ph), _p(sp))
        return sp


STATS_GRID = (0, 0, 1, 1, 0, 0, 2, 2, 0, 0, 1, 1, 1)   # u, v or p grid of each stats row


def cond_bands(s):
    """Planes per canopy band on the u, v and p grids."""
    n31 = s.grid.band_v[0]
    return (n31 + 2 + s.d, n31 + 1 + s.d, n31 + s.d)


def split_cond(s, flat):
    dnx, dnz = s.px // s.ntilex, s.pz // s.ntilez
    nb = cond_bands(s)
    out, a = [], 0
    for q in range(13):
        n = dnx * dnz * nb[STATS_GRID[q]]
        out.append(flat[a:a + n].reshape(nb[STATS_GRID[q]], dnz, dnx))
        a += n
    return out


def cfft(x: np.ndarray, sign: int) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    lib().lho_cfft(len(x), _p(x), _p(out), sign)
    return out
