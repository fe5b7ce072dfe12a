"""Learned canopy closure: training data from the GPU DNS, the coarse-grid features and
targets, and the a-priori evaluation against the classical alternatives.

The closure is project-defined (the reference has no subfilter model).  From DNS snapshots
(spectral state of the GPU solver) the filtered field is formed on a coarse grid:

* x, z: a box of ``fx`` x ``fz`` collocation points (576 -> 96 at 384x289x384: 6 x 6);
* y: coarse centres are every second fine u plane (j = 1, 3, 5, ...), filtered with a box
  over the fine centres j-1, j, j+1 weighted by their cell widths.

The targets are the six subfilter stresses tau_ij = bar(u_i u_j) - bar(u_i) bar(u_j) (v
interpolated to the centres); the features of a coarse point are the nine filtered velocity
gradients on its 3 x 3 x 3 coarse neighbourhood (243 values, x and z periodic, y clamped at
the ends) and the solid fraction of its cell (the share of canopy points of the u grid in
the box), 244 inputs in all.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.fft as sfft

from .channel import ChannelSetup

_WORKERS = max(1, min(16, os.cpu_count() or 1))

STRESS = ("uu", "vv", "ww", "uv", "uw", "vw")
NIN = 244


def modes_to_phys(s: ChannelSetup, a: np.ndarray, workers: int = _WORKERS) -> np.ndarray:
    """Spectral plane (ncol,) -> physical collocation plane (pz, px), the inverse of
    littleharsh_amd.initial.phys_to_modes (inverse transforms unscaled)."""
    b = a.reshape(s.nz, s.mx)
    h = s.nz // 2
    full = np.zeros((s.pz, s.px // 2 + 1), dtype=np.complex128)
    full[:h, : s.mx] = b[:h]
    full[s.pz - h:, : s.mx] = b[h:]
    full = sfft.ifft(full, axis=0, workers=workers) * s.pz
    return sfft.irfft(full, n=s.px, axis=1, workers=workers) * s.px


def phys_to_modes96(a: np.ndarray, nz: int, mx: int) -> np.ndarray:
    """Physical plane(s) (..., nzc, nxc) on the coarse collocation grid -> spectral columns
    (..., nz * mx) of a channel with nz = nzc z modes and mx = nxc / 2 + 1 x modes; the Nyquist
    row and column are dropped (set to zero).  Inverse of modes_to_coarse."""
    nzc, nxc = a.shape[-2:]
    assert nzc == nz and nxc // 2 + 1 == mx
    f = np.fft.rfft2(a, axes=(-2, -1)) / (nzc * nxc)
    f[..., nzc // 2, :] = 0.0
    f[..., :, nxc // 2] = 0.0
    return f.reshape(*a.shape[:-2], nz * mx)


def modes_to_coarse(b: np.ndarray, nz: int, mx: int) -> np.ndarray:
    """Spectral columns (..., nz * mx) -> physical plane(s) (..., nz, 2 (mx - 1)) sampled at the
    coarse points (one per resolved mode, no dealiasing padding)."""
    nxc = 2 * (mx - 1)
    f = b.reshape(*b.shape[:-1], nz, mx)
    return np.fft.irfft2(f, s=(nz, nxc), axes=(-2, -1)) * (nz * nxc)


def velocity_gradients(ub, vb, wb, yc, dxc, dzc) -> np.ndarray:
    """Nine velocity gradients (9, ny, nz, nx) of fields (ny, nz, nx) at the heights yc:
    d/dx, d/dy, d/dz of u, then of v, then of w; central differences, x and z periodic,
    one-sided at the end planes in y."""
    g = []
    for c in (ub, vb, wb):
        ddx = (np.roll(c, -1, axis=2) - np.roll(c, 1, axis=2)) / (2 * dxc)
        ddz = (np.roll(c, -1, axis=1) - np.roll(c, 1, axis=1)) / (2 * dzc)
        ddy = np.empty_like(c)
        ddy[1:-1] = (c[2:] - c[:-2]) / (yc[2:] - yc[:-2])[:, None, None]
        ddy[0] = (c[1] - c[0]) / (yc[1] - yc[0])
        ddy[-1] = (c[-1] - c[-2]) / (yc[-1] - yc[-2])
        g += [ddx, ddy, ddz]
    return np.stack(g)


def stencil_features(grad: np.ndarray, sf: np.ndarray, m, kz, ix) -> np.ndarray:
    """Inputs (n, 244) of the points (m, kz, ix): the nine gradients on the 3x3x3
    neighbourhood, feature ((dm + 1) 9 + (dz + 1) 3 + (dx + 1)) 9 + c (m clamped, z and x
    periodic), then the solid fraction of the point."""
    ny, nz, nx = grad.shape[1:]
    cols = []
    for dm in (-1, 0, 1):
        mm = np.clip(m + dm, 0, ny - 1)
        for dz in (-1, 0, 1):
            zz = (kz + dz) % nz
            for dx in (-1, 0, 1):
                xx = (ix + dx) % nx
                cols.append(grad[:, mm, zz, xx].T)
    return np.concatenate(cols + [sf[m, kz, ix][:, None]], axis=1).astype(np.float32)


def _box(a: np.ndarray, fz: int, fx: int) -> np.ndarray:
    pz, px = a.shape
    return a.reshape(pz // fz, fz, px // fx, fx).mean(axis=(1, 3))


class Filter:
    """Coarse grid of a fine set-up: box sizes, coarse planes, their y, cell weights."""

    def __init__(self, s: ChannelSetup, nxc: int = 96, nzc: int = 96):
        self.s = s
        self.fx, self.fz = s.px // nxc, s.pz // nzc
        assert self.fx * nxc == s.px and self.fz * nzc == s.pz
        self.nxc, self.nzc = nxc, nzc
        g = s.grid
        self.centres = np.arange(1, s.nyu - 1, 2)            # fine u slots of the coarse planes
        self.yc = g.yu[self.centres]
        # width of fine centre cell j: between its faces (v slots j-1, j)
        self.dy = np.zeros(s.nyu)
        self.dy[1:-1] = g.yv[1:] - g.yv[:-1]
        self.dxc = 2.0 * np.pi / s.alp / nxc
        self.dzc = 2.0 * np.pi / s.bet / nzc
        # canopy points of the u grid per fine slot
        self.mask = np.zeros((s.nyu, s.pz, s.px), dtype=bool)
        lst = s.list_u
        if lst.shape[1]:
            self.mask[lst[2] + s.d, lst[1] - 1, lst[0] - 1] = True

    def window(self, j: int):
        js = [q for q in (j - 1, j, j + 1) if 1 <= q <= self.s.nyu - 2]
        w = np.array([self.dy[q] for q in js])
        return js, w / w.sum()

    def filtered(self, u: np.ndarray, v: np.ndarray, w: np.ndarray):
        """State arrays (ncol, nplanes) -> dict of coarse fields (nyc, nzc, nxc): ub, vb, wb,
        the six stresses and the solid fraction."""
        s = self.s
        keys = ("ub", "vb", "wb", "uu", "vv", "ww", "uv", "uw", "vw", "sf")

        def plane(j):                      # one fine u plane -> its box averages
            U = modes_to_phys(s, u[:, j], 1)
            W = modes_to_phys(s, w[:, j], 1)
            V = 0.5 * (modes_to_phys(s, v[:, j - 1], 1) + modes_to_phys(s, v[:, j], 1))
            bx = lambda a: _box(a, self.fz, self.fx)
            return dict(ub=bx(U), vb=bx(V), wb=bx(W), uu=bx(U * U), vv=bx(V * V), ww=bx(W * W),
                        uv=bx(U * V), uw=bx(U * W), vw=bx(V * W), sf=bx(self.mask[j].astype(np.float64)))

        need = sorted({q for j in self.centres for q in self.window(j)[0]})
        with ThreadPoolExecutor(max_workers=_WORKERS) as ex:
            planes = dict(zip(need, ex.map(plane, need)))
        nyc = len(self.centres)
        out = {k: np.zeros((nyc, self.nzc, self.nxc)) for k in keys}
        for m, j in enumerate(self.centres):
            js, wt = self.window(j)
            for q, wq in zip(js, wt):
                for k in keys:
                    out[k][m] += wq * planes[q][k]
        res = {k: out[k] for k in ("ub", "vb", "wb", "sf")}
        for a, b, name in ((0, 0, "uu"), (1, 1, "vv"), (2, 2, "ww"), (0, 1, "uv"), (0, 2, "uw"), (1, 2, "vw")):
            fa = (out["ub"], out["vb"], out["wb"])[a]
            fb = (out["ub"], out["vb"], out["wb"])[b]
            res["tau_" + name] = out[name] - fa * fb
        return res

    def gradients(self, f: dict) -> np.ndarray:
        """Nine filtered velocity gradients (9, nyc, nzc, nxc): d/dx, d/dy, d/dz of u, v, w."""
        return velocity_gradients(f["ub"], f["vb"], f["wb"], self.yc, self.dxc, self.dzc)

    def samples(self, f: dict, n: int | None, rng: np.random.Generator, grad: np.ndarray | None = None):
        """Features (n, 244) and targets (n, 6) at n random coarse points (all when None)."""
        if grad is None:
            grad = self.gradients(f)
        nyc, nzc, nxc = grad.shape[1:]
        if n is None:
            m, kz, ix = (a.ravel() for a in np.meshgrid(np.arange(nyc), np.arange(nzc), np.arange(nxc),
                                                        indexing="ij"))
        else:
            m = rng.integers(0, nyc, n)
            kz = rng.integers(0, nzc, n)
            ix = rng.integers(0, nxc, n)
        x = stencil_features(grad, f["sf"], m, kz, ix)
        y = np.stack([f["tau_" + k][m, kz, ix] for k in STRESS], axis=1)
        return x, y.astype(np.float32), (m, kz, ix)

    def smagorinsky_basis(self, f: dict, idx, grad: np.ndarray | None = None):
        """-2 Delta^2 |S| S_ij at the sample points (n, 6) in the order of STRESS; Delta is the
        cube root of the coarse cell volume.  A fitted constant C_s^2 scales it."""
        if grad is None:
            grad = self.gradients(f)
        m, kz, ix = idx
        G = grad[:, m, kz, ix].reshape(3, 3, -1)        # G[i, j] = d u_i / d x_j
        S = 0.5 * (G + G.transpose(1, 0, 2))
        mag = np.sqrt(2.0 * np.sum(S * S, axis=(0, 1)))
        dyc = np.gradient(self.yc)[m]
        delta = (self.dxc * self.dzc * dyc) ** (1.0 / 3.0)
        base = -2.0 * delta ** 2 * mag
        comp = {"uu": (0, 0), "vv": (1, 1), "ww": (2, 2), "uv": (0, 1), "uw": (0, 2), "vw": (1, 2)}
        return np.stack([base * S[comp[k][0], comp[k][1]] for k in STRESS], axis=1)


def scores(pred: np.ndarray, true: np.ndarray) -> dict:
    """Per stress: relative rms error ||pred - true|| / ||true|| and correlation."""
    out = {}
    for c, k in enumerate(STRESS):
        t, p = true[:, c].astype(np.float64), pred[:, c].astype(np.float64)
        err = float(np.linalg.norm(p - t) / max(np.linalg.norm(t), 1e-300))
        sp, st = np.std(p), np.std(t)
        cor = float(np.mean((p - p.mean()) * (t - t.mean())) / (sp * st)) if sp > 0 and st > 0 else 0.0
        out[k] = {"rel_err": err, "corr": cor}
    return out
