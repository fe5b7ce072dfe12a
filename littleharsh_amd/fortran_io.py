"""Stream files exchanged with the Fortran host demo (littleharsh_amd/fortran/lh_demo.f90).

write_case: int32 hdr(8) = nx, nz, px, pz, ny, dsty, flag_ctpress, precision;
            float64 re, alp, bet, mpgx, qxt, ddthetavi;
            float64 yu, yv, dthdyu, dthdyv, dyu2i(3, nyu), dyv2i(3, nyv), gw(2), rk(15),
                    k1x, k2x, k1z, k2z;
            int32 nlist_u, nlist_v, list_u(3, nlist_u), list_v(3, nlist_v),
                  tip_u(2), tip_v(2), ntilex, ntilez, band_v(2);
            complex128 u1(nyu, ncol), u2(nyv, ncol), u3(nyu, ncol), p(nyp, ncol);
            int32 nsteps, float64 dt.
read_result: complex128 u1, u2, u3, p (same layout), float64 max|div|, mpgx.
All values little-endian, no record markers (Fortran access='stream').
"""
from __future__ import annotations

import numpy as np

from .channel import ChannelSetup


def write_case(path: str, s: ChannelSetup, state, nsteps: int, dt: float, precision: int = 0):
    g = s.grid
    with open(path, "wb") as f:
        np.array([s.nx, s.nz, s.px, s.pz, s.ny, s.d, s.flag_ctpress, precision], "<i4").tofile(f)
        np.array([s.re, s.alp, s.bet, s.mpgx, s.qxt, g.ddthetavi], "<f8").tofile(f)
        for a in (g.yu, g.yv, g.dthdyu, g.dthdyv, g.dyu2i, g.dyv2i, s.gw, s.rk.ravel(),
                  s.k1x, s.k2x, s.k1z, s.k2z):
            np.ascontiguousarray(a, "<f8").tofile(f)
        np.array([s.list_u.shape[1], s.list_v.shape[1]], "<i4").tofile(f)
        for lst in (s.list_u, s.list_v):
            np.ascontiguousarray(lst.T, "<i4").tofile(f)
        np.array([*s.tips["u"], *s.tips["v"], s.ntilex, s.ntilez, *g.band_v], "<i4").tofile(f)
        for a in state:
            np.ascontiguousarray(a, "<c16").tofile(f)     # (ncol, nj) C order == f(j, column)
        np.array([nsteps], "<i4").tofile(f)
        np.array([dt], "<f8").tofile(f)


def read_result(path: str, s: ChannelSetup):
    raw = np.fromfile(path, dtype=np.uint8)
    out, off = [], 0
    for nj in (s.nyu, s.nyv, s.nyu, s.nyp):
        n = s.ncol * nj * 16
        out.append(raw[off:off + n].view("<c16").reshape(s.ncol, nj).copy())
        off += n
    err, mpgx = raw[off:off + 16].view("<f8")
    return out, float(err), float(mpgx)
