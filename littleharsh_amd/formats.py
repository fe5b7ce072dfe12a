"""On-disk formats of the reference, written and read on the host.

* y_grid_canopy.dat      stream file of y_grid_canopy (y_grid_canopy.f90:190-194)
* boundary_*.dat         stream file of boundary_canopies (geom_sym_canopy.f90:239-243)
* stats_*.dat            sequential file of write_stats (stats.f90:255-416)
* spU_/spV_/spW_/spUV_/spP_*.dat   sequential files of recvwrspec (spectra.f90:125-182)
* hist_*.dat             header of the history file (start.f90:260-274)
* u1/u2/u3/p restart files   the plane records read_in reads (start.f90:1786-2121)

Sequential ("unformatted") files are Fortran records: a 4-byte little-endian length, the
payload, the length again.  Integers are int32, reals float64.  The planes of a restart
file are the spectral planes at N resolution as in-place rft buffers: (N(1)+2) reals in x
(Re, Im of kx = 0..N(1)/2) by N(2) rows in z (FFT order).  read_in skips the first three
records; buff_to_u, which copied a plane into the solver's buffers, is absent from the
reference, so this reader keeps the modes the two resolutions share and zero-fills the rest.
"""
from __future__ import annotations

import os
import struct

import numpy as np

from .channel import ChannelSetup

NBAND = 3


# ---------------------------------------------------------------- record helpers
def _rec(f, *parts):
    payload = b"".join(parts)
    n = struct.pack("<i", len(payload))
    f.write(n + payload + n)


def _i4(*v):
    return np.asarray(v, "<i4").tobytes()


def _f8(*v):
    return np.asarray(v, "<f8").tobytes()


def read_records(path):
    """All records of a sequential file as raw bytes."""
    out = []
    with open(path, "rb") as f:
        while True:
            h = f.read(4)
            if not h:
                break
            (n,) = struct.unpack("<i", h)
            out.append(f.read(n))
            f.read(4)
    return out


# ---------------------------------------------------------------- size tables
def band_tables(s: ChannelSetup):
    """N and Ngal as the reference holds them after y_grid_canopy: (4, nband+2) int arrays,
    Fortran column-major when written (start.f90:93-118; y_grid_canopy.f90:147-176)."""
    d, ny = s.d, s.ny
    n31, n32 = s.grid.band_v
    n33 = ny - 2 + d
    Ng = np.zeros((4, NBAND + 2), np.int32)
    Ng[0, 1:NBAND + 1] = s.px
    Ng[0, NBAND + 1] = -2
    Ng[1, 1:NBAND + 1] = s.pz
    Ng[1, NBAND + 1] = 0
    y3 = [-d, n31, n32, n33, 0]
    y4 = [-d, n31 + 1, n32, n33 + 1, 0]
    Ng[2, :] = y3
    Ng[3, :] = y4
    N = Ng.copy()
    N[0, 1:NBAND + 1] = s.nx
    N[1, 1:NBAND + 1] = s.nz
    return N, Ng


def file_tag(s: ChannelSetup, t: float | None = None) -> str:
    """ext1 x ext2 x ext3 [_ext4] of start.f90:250-253."""
    tag = f"{s.nx:04d}x{s.nz:04d}x{s.ny:04d}"
    if t is not None:
        tag += f"_{int(10.0 * t):05d}"
    return tag


# ---------------------------------------------------------------- stream files
def write_y_grid(path: str, s: ChannelSetup):
    N, _ = band_tables(s)
    g = s.grid
    with open(path, "wb") as f:
        f.write(N.T.astype("<i4").tobytes())            # N(4, 0:nband+1), first index fastest
        f.write(_i4(s.d))
        for a in (g.yu, g.yv, g.dthdyu, g.dthdyv):
            f.write(np.asarray(a, "<f8").tobytes())
        f.write(np.asarray(g.dyu2i, "<f8").tobytes())   # (nyu, 3) C order == dyu2i(3, j)
        f.write(np.asarray(g.dyv2i, "<f8").tobytes())


def raw_canopy_lists(s: ChannelSetup):
    """list_ib1, list_ib2 (3, nlist, 2) of boundary_canopies before the getbounds plane
    filter (geom_sym_canopy.f90:98-234): grid 1 = v, grid 2 = u; bottom and top walls."""
    d = s.d
    n33, n43 = s.ny - 2 + d, s.ny - 1 + d
    dnx, dnz = s.px // s.ntilex, s.pz // s.ntilez
    ranges = {"v1": (-d, -1, 0), "u1": (-d + 1, 0, 0),
              "v2": (n33 + 2 - d, n33 + 1, -1), "u2": (n43 - d + 1, n43, 0)}

    def stem(jlo, jhi, jshift):
        j, i, k = np.meshgrid(np.arange(jlo, jhi + 1), np.arange(1, s.dstx + 1),
                              np.arange(1, s.dstz + 1), indexing="ij")
        return np.stack([i.ravel(), k.ravel(), j.ravel() + jshift])

    def rep(p):
        out = []
        for ix in range(1, s.ntilex + 1):
            for iz in range(1, s.ntilez + 1):
                out.append(p + np.array([[dnx * (ix - 1)], [dnz * (iz - 1)], [0]]))
        return np.concatenate(out, axis=1)

    l1 = np.stack([rep(stem(*ranges["v1"])), rep(stem(*ranges["u1"]))], axis=2)
    l2 = np.stack([rep(stem(*ranges["v2"])), rep(stem(*ranges["u2"]))], axis=2)
    return l1.astype(np.int32), l2.astype(np.int32)


def write_boundary(path: str, s: ChannelSetup):
    _, Ng = band_tables(s)
    d = s.d
    n33, n43 = s.ny - 2 + d, s.ny - 1 + d
    l1, l2 = raw_canopy_lists(s)
    with open(path, "wb") as f:
        f.write(_f8(2.0 * np.pi / s.alp, 2.0, 2.0 * np.pi / s.bet))
        f.write(Ng.T.astype("<i4").tobytes())
        nl = l1.shape[1]
        f.write(_i4(nl, nl, -d + 1, 0, n43 - d + 1, n43, -d, -1, n33 + 2 - d, n33 + 1))
        for l in (l1, l2):
            f.write(np.ravel(l, order="F").astype("<i4").tobytes())   # list_ib(3, n, 2)


# ---------------------------------------------------------------- sequential files
def _header(f, s, t, it, mpgx):
    _rec(f, _f8(t, s.re, s.alp, s.bet, mpgx), _i4(NBAND, it), _i4(*([0] * 88)))


def _grid_records(f, s):
    g = s.grid
    dthetai = 1.0
    _rec(f, np.asarray(g.yu, "<f8").tobytes(), _f8(dthetai), np.asarray(g.dthdyu, "<f8").tobytes())
    _rec(f, np.asarray(g.yv, "<f8").tobytes(), _f8(dthetai), np.asarray(g.dthdyv, "<f8").tobytes())


def stats_row_lengths(s: ChannelSetup):
    grid = (0, 0, 1, 1, 0, 0, 2, 2, 0, 0, 1, 1, 1)
    n = (s.nyu, s.nyv, s.nyp)
    return [n[g] for g in grid]


def write_stats(path: str, s: ChannelSetup, rows: np.ndarray, cond: np.ndarray, istat: int,
                t: float, it: int, mpgx: float):
    """rows: (13, nyu) per-plane sums (slot-indexed); cond: the conditional arrays one
    after another (each Fortran (dnx, dnz, nb)), as lh_stats_read / lh_cond_stats_read
    return them."""
    _, Ng = band_tables(s)
    with open(path, "wb") as f:
        _header(f, s, t, it, mpgx)
        _rec(f, Ng.T.astype("<i4").tobytes())
        _grid_records(f, s)
        _rec(f, _i4(istat))
        for q, n in enumerate(stats_row_lengths(s)):
            _rec(f, np.asarray(rows[q, :n], "<f8").tobytes())
        _rec_cond(f, s, cond)


def _rec_cond(f, s, cond):
    dnx, dnz = s.px // s.ntilex, s.pz // s.ntilez
    n31 = s.grid.band_v[0]
    nb = {0: n31 + s.d + 2, 1: n31 + s.d + 1, 2: n31 + s.d}
    grid = (0, 0, 1, 1, 0, 0, 2, 2, 0, 0, 1, 1, 1)
    a = 0
    for q in range(13):
        n = dnx * dnz * nb[grid[q]]
        _rec(f, np.asarray(cond[a:a + n], "<f8").tobytes())
        a += n


def spectra_file_names(s: ChannelSetup, dirout: str, t: float):
    tag = file_tag(s, t)
    return [os.path.join(dirout, f"{p}_{tag}.dat") for p in ("spU", "spV", "spW", "spUV", "spP")]


def spectra_buffer(s: ChannelSetup, sp: np.ndarray, nj: int) -> np.ndarray:
    """recvwrspec for one rank (spectra.f90:125-182): buffSP(0:N(1)/2, 1:N(2)/2+1, j) with
    kz folded onto k <= N(2)/2+1 by assignment in column order (a negative kz overwrites
    the positive one of the same |kz|, as the reference does for its own rank's columns),
    then doubled for i >= 1."""
    mx, nz = s.mx, s.nz
    buf = np.zeros((nj, nz // 2 + 1, mx))
    for c in range(s.ncol):
        i, k1 = c % mx, c // mx + 1
        kk = nz + 2 - k1 if k1 > nz // 2 else k1
        buf[:, kk - 1, i] = sp[c, :nj]
    buf[:, :, 1:mx] *= 2.0
    return buf


def write_spectra(dirout: str, s: ChannelSetup, sp: np.ndarray, istat: int, t: float, it: int,
                  mpgx: float):
    """sp: (5, ncol, nyu) as lh_spectra_read returns it (U, V, W, UV, P)."""
    N, _ = band_tables(s)
    njs = (s.nyu, s.nyv, s.nyu, s.nyu, s.nyp)
    paths = spectra_file_names(s, dirout, t)
    for f_, path, nj in zip(range(5), paths, njs):
        with open(path, "wb") as f:
            _header(f, s, t, it, mpgx)
            _rec(f, N.T.astype("<i4").tobytes())
            _grid_records(f, s)
            _rec(f, _i4(istat))
            _rec(f, spectra_buffer(s, sp[f_], nj).astype("<f8").tobytes())
    return paths


def write_hist_header(path: str, s: ChannelSetup, t: float, mpgx: float):
    N, _ = band_tables(s)
    with open(path, "wb") as f:
        _rec(f, _f8(t, s.re, s.alp, s.bet, mpgx), _i4(NBAND))
        _rec(f, N.T.astype("<i4").tobytes())


# ---------------------------------------------------------------- restart planes
_GRIDS = (("u1", "u"), ("u2", "v"), ("u3", "u"), ("p", "p"))


def _plane_range(s, g):
    d = s.d
    if g == "u":
        return -d, s.ny + d
    if g == "v":
        return -d, s.ny + d - 1
    return -d + 1, s.ny + d - 1


def write_restart(dirout: str, name: str, s: ChannelSetup, state, t: float, it: int,
                  mpgx: float):
    """u1<name>, u2<name>, u3<name>, p<name>: three header records (the run header, N, and
    the y positions of the field's planes), then one record per plane j:
    j, band, N(1)+2, N(2), y(j), plane(N(1)+2, N(2))."""
    N, _ = band_tables(s)
    paths = []
    g = s.grid
    ypos = {"u": g.yu, "v": g.yv, "p": g.yu[1:-1]}
    for (fname, grid), arr in zip(_GRIDS, state):
        path = os.path.join(dirout, fname + name)
        lo, hi = _plane_range(s, grid)
        with open(path, "wb") as f:
            _header(f, s, t, it, mpgx)
            _rec(f, N.T.astype("<i4").tobytes())
            _rec(f, np.asarray(ypos[grid], "<f8").tobytes())
            for r, j in enumerate(range(lo, hi + 1)):
                plane = arr[:, r].reshape(s.nz, s.mx)            # [k][i] complex
                buf = np.zeros((s.nz, s.nx + 2))
                buf[:, 0::2] = plane.real
                buf[:, 1::2] = plane.imag
                _rec(f, _i4(j, 2, s.nx + 2, s.nz), _f8(ypos[grid][r]), buf.astype("<f8").tobytes())
        paths.append(path)
    return paths


def read_restart(dirin: str, name: str, s: ChannelSetup):
    """The state (u1, u2, u3, p in the f(j, column) layout) and (t, Re, alp, bet, mpgx, iter)
    from a restart set; checks nband, alp and bet as read_in does (start.f90:1819-1835)."""
    state = []
    meta = None
    for fname, grid in _GRIDS:
        recs = read_records(os.path.join(dirin, fname + name))
        t, re, alp, bet, mpgx = struct.unpack("<5d", recs[0][:40])
        nband, it = struct.unpack("<2i", recs[0][40:48])
        if nband != NBAND:
            raise ValueError(f"wrong startfile? nband={nband}")
        if alp != s.alp or bet != s.bet:
            raise ValueError(f"wrong startfile? alp,bet = {alp},{bet}")
        meta = dict(t=t, re=re, alp=alp, bet=bet, mpgx=mpgx, iter=it)
        lo, hi = _plane_range(s, grid)
        out = np.zeros((s.ncol, hi - lo + 1), complex)
        for r, rec in enumerate(recs[3:3 + hi - lo + 1]):
            j, band, nxin, nzin = struct.unpack("<4i", rec[:16])
            buf = np.frombuffer(rec[24:], "<f8").reshape(nzin, nxin)
            pl = buf[:, 0::2] + 1j * buf[:, 1::2]                # (nzin, nxin/2)
            mxi = nxin // 2
            dst = np.zeros((s.nz, s.mx), complex)
            mi = min(mxi, s.mx)
            hz = min(nzin, s.nz) // 2
            dst[:hz, :mi] = pl[:hz, :mi]
            dst[s.nz - hz:, :mi] = pl[nzin - hz:, :mi]
            out[:, r] = dst.ravel()
        state.append(out)
    return state, meta
