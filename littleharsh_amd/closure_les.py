"""The coarse channel (LES) in which the learned closure runs, compared with the box-filtered
DNS.

Coarse set-up (coarse_setup): 96 x 145 x 96 (Ngal 144) for the production channel, with its
Re_tau, box and post layout (48 x 24 posts, height 0.15) at the coarse resolution; its 144
interior u planes sit within 0.007 of the filter's coarse centres.  The closure enters through the solver's forcing
hook (Channel.set_forcing) as d_j tau_ij added to the nonlinear term: x and z derivatives
spectrally (i k), y derivatives by differences between the coarse centres; tau_12 and tau_23
are averaged to the v faces for the v equation.  tau is evaluated once per RK3 step from the
state at its start (the same forcing for the three substeps).

Box-filter positions: the box of fine points fx i .. fx i + fx - 1 is centred (fx - 1) / 2
fine cells after the coarse point x_i (likewise in z).  The network was trained on values at
those centres, so the coarse state is sampled there (modes times exp(i k delta)) and the
predicted stresses are moved back to x_i before they are differentiated.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi
from .channel import ChannelSetup, make_setup
from .closure import (STRESS, Filter, modes_to_coarse, phys_to_modes96, velocity_gradients)
from .initial import apply_bc


def coarse_setup(fine: ChannelSetup, nxc: int = 96, nzc: int = 96, dyq: float = 10.0) -> ChannelSetup:
    """Coarse channel of a fine set-up: nxc x nzc modes (Ngal 3/2 of that), one u plane per
    filter centre (half the fine interior planes, half the canopy planes), the same Re, box and
    post tiles, post widths scaled to the coarse collocation points (at least one).  For the
    384 x 289 x 384 production channel: 96 x 145 x 96, posts 2 x 3 of the 3 x 6 points of a
    tile (the fine posts cover 6 x 12 of 12 x 24)."""
    assert (fine.nyu - 2) % 2 == 0 and fine.d % 2 == 0
    pxc, pzc = 3 * nxc // 2, 3 * nzc // 2
    return make_setup(nx=nxc, nz=nzc, ny_total=(fine.nyu - 2) // 2 + 1, dsty=fine.d // 2, re=fine.re,
                      alp=fine.alp, bet=fine.bet, mpgx=fine.mpgx, flag_ctpress=fine.flag_ctpress,
                      posth=fine.posth, dyq=dyq, ntilex=fine.ntilex, ntilez=fine.ntilez,
                      dstx=max(1, round(fine.dstx * pxc / fine.px)), dstz=max(1, round(fine.dstz * pzc / fine.pz)))


def _yinterp(a: np.ndarray, ysrc: np.ndarray, ydst: np.ndarray) -> np.ndarray:
    """Linear interpolation along axis 0 from heights ysrc to ydst (held constant outside)."""
    i = np.clip(np.searchsorted(ysrc, ydst), 1, len(ysrc) - 1)
    w = np.clip((ydst - ysrc[i - 1]) / (ysrc[i] - ysrc[i - 1]), 0.0, 1.0)
    sh = (-1,) + (1,) * (a.ndim - 1)
    return a[i - 1] * (1.0 - w).reshape(sh) + a[i] * w.reshape(sh)


def _ddy_centres(a: np.ndarray, y: np.ndarray) -> np.ndarray:
    """d/dy along axis 0 at the heights y: central inside, one-sided at the ends."""
    d = np.empty_like(a)
    d[1:-1] = (a[2:] - a[:-2]) / (y[2:] - y[:-2]).reshape(-1, *(1,) * (a.ndim - 1))
    d[0] = (a[1] - a[0]) / (y[1] - y[0])
    d[-1] = (a[-1] - a[-2]) / (y[-1] - y[-2])
    return d


class Coarse:
    """Transforms between the coarse channel's state and the filter's coarse grid."""

    def __init__(self, c: ChannelSetup, filt: Filter):
        self.c, self.filt = c, filt
        self.nz, self.mx = c.nz, c.mx
        assert 2 * (c.mx - 1) == filt.nxc and c.nz == filt.nzc
        self.yc = c.grid.yu[1:-1]
        assert len(self.yc) == len(filt.yc)
        col = np.arange(c.ncol)
        kx = c.k1x[col % c.mx]
        kz = c.k1z[col // c.mx]
        s = filt.s
        dx = 0.5 * (filt.fx - 1) * (2.0 * np.pi / s.alp) / s.px
        dz = 0.5 * (filt.fz - 1) * (2.0 * np.pi / s.bet) / s.pz
        self.dx_shift, self.dz_shift = dx, dz
        self.shift = np.exp(1j * (kx * dx + kz * dz))
        self.ikx, self.ikz = 1j * kx, 1j * kz

    # ---- state <-> coarse-point fields
    def initial_state(self, f: dict):
        """Coarse state (u, v, w, p) from the box-filtered fields ub, vb, wb of a snapshot."""
        c, fy = self.c, self.filt.yc

        def modes(a):                                   # (np, nz, nx) -> (ncol, np)
            return np.ascontiguousarray((phys_to_modes96(a, self.nz, self.mx) / self.shift).T)

        u = modes(_yinterp(f["ub"], fy, c.grid.yu))
        w = modes(_yinterp(f["wb"], fy, c.grid.yu))
        v = modes(_yinterp(f["vb"], fy, c.grid.yv))
        apply_bc(c, u, v, w)
        return u, v, w, np.zeros((c.ncol, c.nyp), dtype=np.complex128)

    def sample(self, u, v, w):
        """ub, vb, wb (144, nz, nx) at the box centres of the coarse u planes (v averaged from
        the two faces of each centre)."""
        def phys(cols):                                 # (ncol, np) -> (np, nz, nx)
            return modes_to_coarse(cols.T * self.shift, self.nz, self.mx)

        return phys(u[:, 1:-1]), phys(0.5 * (v[:, :-1] + v[:, 1:])), phys(w[:, 1:-1])

    def gradients(self, ub, vb, wb) -> np.ndarray:
        return velocity_gradients(ub, vb, wb, self.yc, self.filt.dxc, self.filt.dzc)

    # ---- stresses -> forcing of the nonlinear term
    def forcing(self, tau: np.ndarray):
        """tau (ny, nz, nx, 6) in the order of STRESS at the box centres -> d_j tau_ij in the
        layout of Channel.set_forcing (f1, f3: (ncol, 144) at the centres; f2: (ncol, 143) at
        the interior v faces)."""
        T = {k: phys_to_modes96(np.ascontiguousarray(tau[..., n], dtype=np.float64), self.nz, self.mx)
             / self.shift for n, k in enumerate(STRESS)}
        y = self.yc
        f1 = self.ikx * T["uu"] + self.ikz * T["uw"] + _ddy_centres(T["uv"], y)
        f3 = self.ikx * T["uw"] + self.ikz * T["ww"] + _ddy_centres(T["vw"], y)
        f2 = (0.5 * (self.ikx * (T["uv"][:-1] + T["uv"][1:]) + self.ikz * (T["vw"][:-1] + T["vw"][1:]))
              + (T["vv"][1:] - T["vv"][:-1]) / (y[1:] - y[:-1])[:, None])
        return f1.T.copy(), f2.T.copy(), f3.T.copy()


class DeviceClosure:
    """The learned closure coupled on the GPU (lh_closure_*): apply() computes d_j tau_ij
    from the channel's current state into its forcing buffers, on the channel's stream.
    DeviceClosure.smagorinsky(ch, co, cs2) couples the Smagorinsky stresses the same way."""

    @classmethod
    def smagorinsky(cls, ch, co: Coarse, cs2: float):
        self = cls.__new__(cls)
        self.lib = capi.load_library()
        self._keep = [np.ascontiguousarray(co.yc, dtype=np.float64)]
        self.h = ctypes.c_void_p()
        r = self.lib.lh_closure_create_smagorinsky(
            ch._h, self._keep[0].ctypes.data_as(ctypes.POINTER(ctypes.c_double)), co.filt.dxc, co.filt.dzc,
            co.dx_shift, co.dz_shift, float(cs2), ctypes.byref(self.h))
        self._check(r, "lh_closure_create_smagorinsky")
        self._ch, self._net = ch, None
        return self

    def __init__(self, ch, net, co: Coarse, sf, xmean, xscale, ymean, yscale):
        self.lib = capi.load_library()
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        self._keep = [f32(sf), f32(xmean), f32(xscale), f32(ymean), f32(yscale),
                      np.ascontiguousarray(co.yc, dtype=np.float64)]
        fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        sf_, xm, xs, ym, ys, yc = self._keep
        assert sf_.shape == (len(co.yc), co.filt.nzc, co.filt.nxc)
        self.h = ctypes.c_void_p()
        r = self.lib.lh_closure_create(ch._h, net.h, yc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       co.filt.dxc, co.filt.dzc, co.dx_shift, co.dz_shift, fp(sf_), fp(xm),
                                       fp(xs), fp(ym), fp(ys), ctypes.byref(self.h))
        self._check(r, "lh_closure_create")
        self._ch, self._net = ch, net            # both must outlive the closure

    def _check(self, r, what):
        if r != 0:
            msg = self.lib.lh_closure_last_error(self.h).decode() if self.h else ""
            raise RuntimeError(f"{what}: status {r}: {msg}")

    def set_mean_only(self, on: bool = True):
        """Keep only the plane-mean column of the forcing (the effect on the mean profile)."""
        self._check(self.lib.lh_closure_set_mean_only(self.h, 1 if on else 0), "lh_closure_set_mean_only")
        return self

    def apply(self):
        self._check(self.lib.lh_closure_apply(self.h), "lh_closure_apply")

    def close(self):
        if self.h:
            self.lib.lh_closure_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def smagorinsky_tau(grad: np.ndarray, yc: np.ndarray, dxc: float, dzc: float, cs2: float) -> np.ndarray:
    """cs2 (-2 Delta^2 |S| S_ij) at every point, (ny, nz, nx, 6) in the order of STRESS;
    Delta = (dx dy dz)^(1/3) of the coarse cell."""
    G = grad.reshape(3, 3, *grad.shape[1:])           # G[i, j] = d u_i / d x_j
    S = 0.5 * (G + G.transpose(1, 0, 2, 3, 4))
    mag = np.sqrt(2.0 * np.sum(S * S, axis=(0, 1)))
    delta = (dxc * dzc * np.gradient(yc)) ** (1.0 / 3.0)
    base = -2.0 * cs2 * (delta ** 2)[:, None, None] * mag
    comp = {"uu": (0, 0), "vv": (1, 1), "ww": (2, 2), "uv": (0, 1), "uw": (0, 2), "vw": (1, 2)}
    return np.stack([base * S[comp[k]] for k in STRESS], axis=-1)


# ---- the comparison with the DNS: plane statistics of the filtered fields
def plane_stats(ub, vb, wb) -> dict:
    """Per plane: mean U, rms of u', v', w', mean u'v', and the streamwise spectrum of u'
    summed over the planes of each wall half (canopy side first)."""
    U = ub.mean(axis=(1, 2))
    up = ub - U[:, None, None]
    vp = vb - vb.mean(axis=(1, 2))[:, None, None]
    wp = wb - wb.mean(axis=(1, 2))[:, None, None]
    E = (np.abs(np.fft.rfft(up, axis=2)) ** 2).mean(axis=1)    # (ny, nkx)
    h = ub.shape[0] // 2
    return dict(U=U, urms=np.sqrt((up ** 2).mean(axis=(1, 2))), vrms=np.sqrt((vp ** 2).mean(axis=(1, 2))),
                wrms=np.sqrt((wp ** 2).mean(axis=(1, 2))), uv=(up * vp).mean(axis=(1, 2)),
                Eu=np.stack([E[:h].sum(axis=0), E[h:].sum(axis=0)]))


def compare(les: dict, dns: dict, near: np.ndarray) -> dict:
    """Errors of LES statistics against the filtered DNS: relative L2 over y for each
    profile (all planes and the `near` planes next to the canopies), the mean |log10| ratio of
    the u spectra (kx = 1 .. nx/2 - 1)."""
    out = {}
    for k in ("U", "urms", "vrms", "wrms", "uv"):
        a, b = les[k], dns[k]
        out[k] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        out[k + "_near"] = float(np.linalg.norm((a - b)[near]) / max(np.linalg.norm(b[near]), 1e-300))
    n = les["Eu"].shape[1] - 1
    r = les["Eu"][:, 1:n] / np.maximum(dns["Eu"][:, 1:n], 1e-300)
    out["Eu_log10"] = float(np.mean(np.abs(np.log10(np.maximum(r, 1e-300)))))
    return out
