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

import numpy as np

from .channel import ChannelSetup

STRESS = ("uu", "vv", "ww", "uv", "uw", "vw")
NIN = 244


def modes_to_phys(s: ChannelSetup, a: np.ndarray) -> np.ndarray:
    """Spectral plane (ncol,) -> physical collocation plane (pz, px), the inverse of
    littleharsh_amd.initial.phys_to_modes (inverse transforms unscaled)."""
    b = a.reshape(s.nz, s.mx)
    h = s.nz // 2
    full = np.zeros((s.pz, s.px // 2 + 1), dtype=np.complex128)
    full[:h, : s.mx] = b[:h]
    full[s.pz - h:, : s.mx] = b[h:]
    full = np.fft.ifft(full, axis=0) * s.pz
    return np.fft.irfft(full, n=s.px, axis=1) * s.px


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
        cache = {}

        def plane(j):
            if j not in cache:
                U = modes_to_phys(s, u[:, j])
                W = modes_to_phys(s, w[:, j])
                V = 0.5 * (modes_to_phys(s, v[:, j - 1]) + modes_to_phys(s, v[:, j]))
                bx = lambda a: _box(a, self.fz, self.fx)
                cache[j] = dict(ub=bx(U), vb=bx(V), wb=bx(W), uu=bx(U * U), vv=bx(V * V), ww=bx(W * W),
                                uv=bx(U * V), uw=bx(U * W), vw=bx(V * W),
                                sf=bx(self.mask[j].astype(np.float64)))
            return cache[j]

        nyc = len(self.centres)
        keys = ("ub", "vb", "wb", "uu", "vv", "ww", "uv", "uw", "vw", "sf")
        out = {k: np.zeros((nyc, self.nzc, self.nxc)) for k in keys}
        for m, j in enumerate(self.centres):
            js, wt = self.window(j)
            for q, wq in zip(js, wt):
                pl = plane(q)
                for k in keys:
                    out[k][m] += wq * pl[k]
            for q in list(cache):
                if q < j - 1:
                    del cache[q]
        res = {k: out[k] for k in ("ub", "vb", "wb", "sf")}
        for a, b, name in ((0, 0, "uu"), (1, 1, "vv"), (2, 2, "ww"), (0, 1, "uv"), (0, 2, "uw"), (1, 2, "vw")):
            fa = (out["ub"], out["vb"], out["wb"])[a]
            fb = (out["ub"], out["vb"], out["wb"])[b]
            res["tau_" + name] = out[name] - fa * fb
        return res

    def gradients(self, f: dict) -> np.ndarray:
        """Nine filtered velocity gradients (9, nyc, nzc, nxc): d/dx, d/dy, d/dz of u, v, w."""
        y = self.yc
        g = []
        for c in (f["ub"], f["vb"], f["wb"]):
            ddx = (np.roll(c, -1, axis=2) - np.roll(c, 1, axis=2)) / (2 * self.dxc)
            ddz = (np.roll(c, -1, axis=1) - np.roll(c, 1, axis=1)) / (2 * self.dzc)
            ddy = np.empty_like(c)
            ddy[1:-1] = (c[2:] - c[:-2]) / (y[2:] - y[:-2])[:, None, None]
            ddy[0] = (c[1] - c[0]) / (y[1] - y[0])
            ddy[-1] = (c[-1] - c[-2]) / (y[-1] - y[-2])
            g += [ddx, ddy, ddz]
        return np.stack(g)

    def samples(self, f: dict, n: int | None, rng: np.random.Generator):
        """Features (n, 244) and targets (n, 6) at n random coarse points (all when None)."""
        grad = self.gradients(f)
        nyc, nzc, nxc = grad.shape[1:]
        if n is None:
            m, kz, ix = (a.ravel() for a in np.meshgrid(np.arange(nyc), np.arange(nzc), np.arange(nxc),
                                                        indexing="ij"))
        else:
            m = rng.integers(0, nyc, n)
            kz = rng.integers(0, nzc, n)
            ix = rng.integers(0, nxc, n)
        cols = []
        for dm in (-1, 0, 1):
            mm = np.clip(m + dm, 0, nyc - 1)
            for dz in (-1, 0, 1):
                zz = (kz + dz) % nzc
                for dx in (-1, 0, 1):
                    xx = (ix + dx) % nxc
                    cols.append(grad[:, mm, zz, xx].T)
        x = np.concatenate(cols + [f["sf"][m, kz, ix][:, None]], axis=1)
        y = np.stack([f["tau_" + k][m, kz, ix] for k in STRESS], axis=1)
        return x.astype(np.float32), y.astype(np.float32), (m, kz, ix)

    def smagorinsky_basis(self, f: dict, idx):
        """-2 Delta^2 |S| S_ij at the sample points (n, 6) in the order of STRESS; Delta is the
        cube root of the coarse cell volume.  A fitted constant C_s^2 scales it."""
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
