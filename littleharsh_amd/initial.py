"""Seeded initial conditions in the reference's state layout (host side).

The reference starts from a parabolic profile (mblock_ini_parabolic_profile,
start.f90:1644-1784) or restart planes (read_in, start.f90:1786-2235); both end in
planes_to_modes, i.e. a forward x/z transform of physical planes truncated to the
dealiased modes.  These generators do the same with numpy on the host:

* ``laminar_plus_noise``: U = Re*|mpgx|/2*(1-y^2) between the post tips plus i.i.d.
  uniform noise of amplitude ``amp*U_bulk`` on every physical point of u, v, w (counter-
  based SplitMix64 keyed by seed, field, plane, z, x), zero at the canopy points of each
  grid and v = 0 on the walls.
* ``laminar_plus_modes``: the laminar profile plus random Fourier modes 1 <= |k| <= kmax
  (index units) with |k|^(-5/3) amplitudes and a (1-y^2) shape between the tips.

Arrays are complex128 of shape (ncol, nplanes) -- complex(8) f(j, column) in memory.
"""
from __future__ import annotations

import numpy as np

from .channel import ChannelSetup

_GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix64_uniform(seed: int, counters: np.ndarray) -> np.ndarray:
    """Uniform [0, 1) doubles from SplitMix64 applied to seed + (counter+1)*golden."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (counters.astype(np.uint64) + np.uint64(1)) * _GOLD
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def phys_to_modes(s: ChannelSetup, ph: np.ndarray) -> np.ndarray:
    """Physical planes ph[j, z, x] -> spectral (nplanes, ncol) with the rft/cft forward
    conventions (x 1/px then x 1/pz, rft_buff.f:63-73, cft_buff.f:37-46)."""
    a = np.fft.rfft(ph, axis=2) / s.px
    a = a[:, :, : s.mx]
    a[:, :, 0] = a[:, :, 0].real
    b = np.fft.fft(a, axis=1) / s.pz
    h = s.nz // 2
    b = np.concatenate([b[:, :h, :], b[:, s.pz - h:, :]], axis=1)
    return b.reshape(ph.shape[0], s.ncol)


def _mask_planes(s: ChannelSetup, lst: np.ndarray, nplanes: int) -> np.ndarray:
    m = np.zeros((nplanes, s.pz, s.px), dtype=bool)
    if lst.shape[1]:
        m[lst[2] + s.d, lst[1] - 1, lst[0] - 1] = True
    return m


def apply_bc(s: ChannelSetup, u: np.ndarray, v: np.ndarray, w: np.ndarray) -> None:
    """Ghost u/w from the wall weights and v = 0 on the walls (v_corr,
    littleharsh.f90:866-929).  Arrays are (ncol, nplanes)."""
    gw = s.gw
    for f in (u, w):
        f[:, 0] = -gw[0] * f[:, 1]
        f[:, -1] = -gw[1] * f[:, -2]
    v[:, 0] = 0.0
    v[:, -1] = 0.0


def laminar_profile(s: ChannelSetup) -> np.ndarray:
    """0.5*Re*|mpgx|*(1-y^2) at the u centres between the tips, 0 in the canopy
    (cf. start.f90:1743)."""
    yu = s.grid.yu
    j = np.arange(-s.d, s.ny + s.d + 1)
    prof = 0.5 * s.re * abs(s.mpgx) * (1.0 - yu ** 2)
    prof[(j < 1) | (j > s.ny - 1)] = 0.0
    return prof


def laminar_plus_noise(s: ChannelSetup, seed: int = 1234, amp: float = 0.1):
    ub = s.re * abs(s.mpgx) / 3.0
    prof = laminar_profile(s)
    mu = _mask_planes(s, s.list_u, s.nyu)
    mv = _mask_planes(s, s.list_v, s.nyv)
    out = []
    for field, (nplanes, mask) in enumerate(((s.nyu, mu), (s.nyv, mv), (s.nyu, mu))):
        cnt = (np.arange(nplanes, dtype=np.uint64)[:, None, None] * np.uint64(s.pz)
               + np.arange(s.pz, dtype=np.uint64)[None, :, None]) * np.uint64(s.px) \
            + np.arange(s.px, dtype=np.uint64)[None, None, :]
        cnt = cnt + np.uint64(field) * np.uint64(nplanes * s.pz * s.px)
        noise = amp * ub * (2.0 * splitmix64_uniform(seed, cnt) - 1.0)
        noise[mask] = 0.0
        noise[0] = 0.0
        noise[-1] = 0.0
        if field == 0:
            noise += prof[:, None, None]
        out.append(np.ascontiguousarray(phys_to_modes(s, noise).T))
    u, v, w = out
    apply_bc(s, u, v, w)
    p = np.zeros((s.ncol, s.nyp), dtype=np.complex128)
    return u, v, w, p


def laminar_plus_modes(s: ChannelSetup, seed: int = 1234, kmax: int = 16, amp: float = 0.1):
    rng = np.random.default_rng(seed)
    ub = s.re * abs(s.mpgx) / 3.0
    yu, yv = s.grid.yu, s.grid.yv
    ju = np.arange(-s.d, s.ny + s.d + 1)
    jv = np.arange(-s.d, s.ny + s.d)
    shu = np.where((ju >= 1) & (ju <= s.ny - 1), 1.0 - yu ** 2, 0.0)
    shv = np.where((jv >= 0) & (jv <= s.ny - 1), 1.0 - yv ** 2, 0.0)
    fields = [np.zeros((s.ncol, s.nyu), complex), np.zeros((s.ncol, s.nyv), complex),
              np.zeros((s.ncol, s.nyu), complex)]
    for i in range(0, min(kmax, s.nx // 2) + 1):
        for b in range(-kmax, kmax + 1):
            kk = np.hypot(i, b)
            if kk < 1.0 or kk > kmax or abs(b) >= s.nz // 2:
                continue
            if i == 0 and b < 0:
                continue
            kidx = b if b >= 0 else s.nz + b
            c = kidx * s.mx + i
            a = amp * ub * kk ** (-5.0 / 3.0)
            for f, sh in zip(fields, (shu, shv, shu)):
                coef = a * np.exp(2j * np.pi * rng.random())
                f[c, :] += coef * sh
                if i == 0:
                    f[(s.nz - b) * s.mx, :] += np.conj(coef) * sh
    fields[0][0, :] += laminar_profile(s)
    u, v, w = fields
    apply_bc(s, u, v, w)
    p = np.zeros((s.ncol, s.nyp), dtype=np.complex128)
    return u, v, w, p
