"""Host-side set-up of one Littleharsh channel: grid metrics, wavenumbers, RK3 tables,
canopy point lists and the index conventions shared by the HIP library and its checks.

This is the caller side of the substep boundary.  In the reference it is Fortran host
code (start.f90, y_grid_canopy.f90, geom_sym_canopy.f90) that runs once at start-up; the
arrays it produces are plain data handed to ``lh_create`` (include/littleharsh_amd.h).

Index conventions (single resolution: Ngal band 1 == band 2, so every spectral column
lives in the reference's middle band, start.f90:1038-1040):

* ``ny`` is the reference input ``ny`` (input.in:19): ``N(3,3) = ny-2`` before the canopy
  is added, ``nn = ny-1`` intervals between the two post tips (y_grid_canopy.f90:31).
* ``d = dsty`` canopy planes per wall.
* u/w ("ugrid") planes j = -d .. ny+d      (nyu = ny+2d+1, ghosts at both ends)
* v   ("vgrid") planes j = -d .. ny+d-1    (nyv = ny+2d, walls at both ends)
* p   ("pgrid") planes j = -d+1 .. ny+d-1  (nyp = ny+2d-1)
  (jlim, start.f90:1152-1175, with N(3,0)=N(4,0)=-d, N(3,3)=ny-2+d, N(4,3)=ny-1+d.)
* Spectral column c = k*(nx//2+1) + i, i = 0..nx/2 fastest, k = 0..nz-1 in FFT order
  (reference column order at np=1, start.f90:1105-1119, with k = columns_k-1).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


# ---------------------------------------------------------------------------------------
# RK3 (Le & Moin) coefficients, start.f90:412-427
# ---------------------------------------------------------------------------------------
def rk3_tables() -> np.ndarray:
    """Return [aRK, bRK, gRK, cRK, dRK] as a (5, 3) float64 array (start.f90:413-427)."""
    a = [4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0]
    b = [4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0]
    g = [8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0]
    c = [0.0, -17.0 / 60.0, -5.0 / 12.0]
    dd = [8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0]
    return np.array([a, b, g, c, dd], dtype=np.float64)


# ---------------------------------------------------------------------------------------
# Wavenumbers, def_k start.f90:655-704
# ---------------------------------------------------------------------------------------
def def_k(nx: int, nz: int, alp: float, bet: float):
    """Imaginary parts of k1F_x, k1F_z and the real k2F_x, k2F_z (start.f90:688-701).

    k1F_x(i) = im*alp*i, i = 0..nx/2;  k1F_z(k) = im*bet*(k-1) for k <= nz/2, else
    -im*bet*(nz-k+1) (1-based k).  Returned arrays use 0-based k.
    """
    i = np.arange(nx // 2 + 1, dtype=np.float64)
    k1x = alp * i
    k2x = -(alp * i) ** 2
    k1z = np.empty(nz)
    k2z = np.empty(nz)
    for k in range(1, nz // 2 + 1):
        k1z[k - 1] = bet * (k - 1)
        k2z[k - 1] = -((bet * (k - 1)) ** 2)
    for k in range(nz // 2 + 1, nz + 1):
        k1z[k - 1] = -(bet * (nz - k + 1))
        k2z[k - 1] = -((bet * (nz - k + 1)) ** 2)
    return k1x, k2x, k1z, k2z


# ---------------------------------------------------------------------------------------
# y grid, y_grid_canopy.f90:1-197
# ---------------------------------------------------------------------------------------
@dataclasses.dataclass
class YGrid:
    d: int
    ny: int
    yu: np.ndarray      # j = -d .. ny+d
    yv: np.ndarray      # j = -d .. ny+d-1
    dthdyu: np.ndarray
    dthdyv: np.ndarray
    dyu2i: np.ndarray   # (nyu, 3): columns = dyu2i(1..3, j)
    dyv2i: np.ndarray   # (nyv, 3)
    ddthetavi: float
    band_v: tuple       # N(3,1), N(3,2): band limits on the v grid (y_grid_canopy.f90:152-163)


def y_grid_canopy(ny: int, dsty: int, posth: float, dyq: float, ppp: int = 5) -> YGrid:
    """Degree-``ppp`` polynomial map between the post tips plus the quintic canopy
    extension to the walls at y = -1-posth and 1+posth (y_grid_canopy.f90:34-145).

    Grids are uniform in theta (dtheta = 1).  ``dsty`` must be positive (posts present).
    """
    if dsty <= 0:
        raise ValueError("dsty must be positive: the grid is built for canopy walls")
    dny = dsty
    nn = ny - 1                          # y_grid_canopy.f90:31, nn = N(3,nband)+1
    q = nn / 2.0
    aaa = 15.0 * (dyq - 1.0) / ((8.0 * dyq + 7.0) * (q ** ppp))
    bbb = -30.0 * (dyq - 1.0) / ((8.0 * dyq + 7.0) * (q ** (ppp - 2)))
    ccc = 15.0 * dyq / ((8.0 * dyq + 7.0) * q)
    nyu = nn + 2 * dny + 2               # j = -dny .. nn+dny+1
    nyv = nn + 2 * dny + 1               # j = -dny .. nn+dny
    yu = np.zeros(nyu)
    yv = np.zeros(nyv)
    dtu = np.zeros(nyu)
    dtv = np.zeros(nyv)
    ou, ov = dny, dny                    # array offset of j = 0

    def pw(x, n):                        # Fortran integer power: repeated products
        r = 1.0
        for _ in range(n):
            r *= x
        return r

    for j in range(0, nn):               # y_grid_canopy.f90:56-62
        s = j - q
        h = j + 0.5 - q
        yv[ov + j] = aaa * pw(s, ppp) / 5.0 + bbb * pw(s, ppp - 2) / 3.0 + ccc * s
        yu[ou + j + 1] = aaa * pw(h, ppp) / 5.0 + bbb * pw(h, ppp - 2) / 3.0 + ccc * h
        dtv[ov + j] = 1.0 / (aaa * pw(s, ppp - 1) + bbb * pw(s, ppp - 3) + ccc)
        dtu[ou + j + 1] = 1.0 / (aaa * pw(h, ppp - 1) + bbb * pw(h, ppp - 3) + ccc)

    grid_min = yv[ov + 1] - yv[ov + 0]   # y_grid_canopy.f90:77
    aa1 = -(15.0 * (posth - grid_min * dny)) / (7.0 * dny ** 5.0)
    cc1 = (30.0 * (posth - grid_min * dny)) / (7.0 * dny ** 3.0)
    ee1 = grid_min
    h = -dny - 0.5                       # y_grid_canopy.f90:83-84
    yu[ou - dny] = aa1 * pw(h, 5) / 5.0 + cc1 * pw(h, 3) / 3.0 + ee1 * h - 1.0
    dtu[ou - dny] = 1.0 / (aa1 * pw(h, 4) + cc1 * pw(h, 2) + ee1)
    for j in range(-dny, 0):             # y_grid_canopy.f90:86-92
        h = j + 0.5
        yv[ov + j] = aa1 * float(j) ** 5.0 / 5.0 + cc1 * float(j) ** 3.0 / 3.0 + ee1 * j - 1.0
        yu[ou + j + 1] = aa1 * pw(h, 5) / 5.0 + cc1 * pw(h, 3) / 3.0 + ee1 * h - 1.0
        dtv[ov + j] = 1.0 / (aa1 * pw(float(j), 4) + cc1 * pw(float(j), 2) + ee1)
        dtu[ou + j + 1] = 1.0 / (aa1 * pw(h, 4) + cc1 * pw(h, 2) + ee1)
    for j in range(0, dny + 1):          # y_grid_canopy.f90:94-100
        h = j + 0.5
        yv[ov + j + nn] = aa1 * float(j) ** 5.0 / 5.0 + cc1 * float(j) ** 3.0 / 3.0 + ee1 * j + 1.0
        yu[ou + j + nn + 1] = aa1 * pw(h, 5) / 5.0 + cc1 * pw(h, 3) / 3.0 + ee1 * h + 1.0
        dtv[ov + j + nn] = 1.0 / (aa1 * pw(float(j), 4) + cc1 * pw(float(j), 2) + ee1)
        dtu[ou + j + nn + 1] = 1.0 / (aa1 * pw(h, 4) + cc1 * pw(h, 2) + ee1)

    # second-derivative weights, y_grid_canopy.f90:114-145
    du = np.zeros((nyu, 3))
    dv = np.zeros((nyv, 3))

    def U(j):
        return yu[ou + j]

    def V(j):
        return yv[ov + j]

    j = -dny
    du[ou + j] = [0.0, -2.0 / (U(j + 1) - U(j)) ** 2, 1.0 / (U(j + 1) - U(j)) ** 2]
    dv[ov + j] = [0.0, -2.0 / (V(j + 1) - V(j)) ** 2, 1.0 / (V(j + 1) - V(j)) ** 2]
    for j in range(-dny + 1, nn + dny):
        du[ou + j] = [2.0 / ((U(j) - U(j - 1)) * (U(j + 1) - U(j - 1))),
                      -2.0 / ((U(j + 1) - U(j)) * (U(j) - U(j - 1))),
                      2.0 / ((U(j + 1) - U(j)) * (U(j + 1) - U(j - 1)))]
        dv[ov + j] = [2.0 / ((V(j) - V(j - 1)) * (V(j + 1) - V(j - 1))),
                      -2.0 / ((V(j + 1) - V(j)) * (V(j) - V(j - 1))),
                      2.0 / ((V(j + 1) - V(j)) * (V(j + 1) - V(j - 1)))]
    j = nn + dny
    du[ou + j] = [2.0 / ((U(j) - U(j - 1)) * (U(j + 1) - U(j - 1))),
                  -2.0 / ((U(j + 1) - U(j)) * (U(j) - U(j - 1))),
                  2.0 / ((U(j + 1) - U(j)) * (U(j + 1) - U(j - 1)))]
    dv[ov + j] = [1.0 / (V(j) - V(j - 1)) ** 2, -2.0 / (V(j) - V(j - 1)) ** 2, 0.0]
    j = nn + dny + 1
    du[ou + j] = [1.0 / (U(j) - U(j - 1)) ** 2, -2.0 / (U(j) - U(j - 1)) ** 2, 0.0]

    # band limits used by the conditional statistics (y_grid_canopy.f90:147-163, with
    # h_ny(1) = 1.5*posth for canopies, start.f90:233-236)
    n33 = ny - 2
    hny = yu[ou + 0] + 1.5 * posth
    jb = -dny
    while yu[ou + jb] < hny:
        jb += 1
    band_v = (jb, n33 - jb)
    return YGrid(d=dny, ny=ny, yu=yu, yv=yv, dthdyu=dtu, dthdyv=dtv, dyu2i=du, dyv2i=dv,
                 ddthetavi=1.0, band_v=band_v)


def gridweighting(g: YGrid) -> np.ndarray:
    """gridweighting(midband,1:2) (start.f90:500-501): ghost-point extrapolation weights."""
    d, ny = g.d, g.ny
    yu = lambda j: g.yu[j + d]
    yv = lambda j: g.yv[j + d]
    n30 = n40 = -d
    n33, n43 = ny - 2 + d, ny - 1 + d
    gw1 = -(yv(n30) - yu(n40)) / (yv(n30) - yu(n40 + 1))
    gw2 = (yu(n43 + 1) - yv(n33 + 1)) / (yv(n33 + 1) - yu(n43))
    return np.array([gw1, gw2])


# ---------------------------------------------------------------------------------------
# Canopy point lists, geom_sym_canopy.f90:63-234 and the plane filter of getbounds
# (start.f90:767-801)
# ---------------------------------------------------------------------------------------
def boundary_canopies(px: int, pz: int, ny: int, dsty: int, dstx: int, dstz: int,
                      ntilex: int, ntilez: int):
    """Return (list_u, list_v, tips) with effective point lists (3, n) of 1-based (i, k)
    and plane index j, one post of dstx*dstz points per tile on each wall.

    The reference builds the lists per wall (geom_sym_canopy.f90:98-234) and getbounds
    keeps only the points on the planes between the IB plane limits, which exclude the
    walls (start.f90:770-779, 788-801).  That drops the bottom v wall plane j=-d and the
    lowest plane of the top v list, j = ny-1; what remains is mirror-symmetric.
    The tip planes are those of tridLU_3D_IB.f90:237 (u: 0 and Ngal(4,3)-dsty+1) and
    tridLU_3D_IB.f90:317 (v: -1 and Ngal(3,3)-dsty+2).
    """
    if px % ntilez != 0:                  # the reference check, geom_sym_canopy.f90:54
        raise ValueError("nx not multiple of ntilex")
    if pz % ntilez != 0:
        raise ValueError("nz not multiple of ntilez")
    if px % ntilex != 0:
        raise ValueError("px not multiple of ntilex")
    d = dsty
    n33, n43 = ny - 2 + d, ny - 1 + d    # Ngal(3,3), Ngal(4,3) after y_grid_canopy
    dnx, dnz = px // ntilex, pz // ntilez
    nyu11, nyu21 = -d + 1, 0
    nyv11, nyv21 = -d, -1
    nyu12, nyu22 = n43 - d + 1, n43
    nyv12, nyv22 = n33 + 1 - d + 1, n33 + 1

    def stem(jlo, jhi, jshift):
        # loop order j, i, k as geom_sym_canopy.f90:103-121
        j, i, k = np.meshgrid(np.arange(jlo, jhi + 1), np.arange(1, dstx + 1),
                              np.arange(1, dstz + 1), indexing="ij")
        return np.stack([i.ravel(), k.ravel(), j.ravel() + jshift])

    def replicate(pts):
        # tile order ix outer, iz inner as geom_sym_canopy.f90:219-233
        out = []
        for ix in range(1, ntilex + 1):
            for iz in range(1, ntilez + 1):
                out.append(pts + np.array([[dnx * (ix - 1)], [dnz * (iz - 1)], [0]]))
        return np.concatenate(out, axis=1)

    lu = np.concatenate([replicate(stem(nyu11, nyu21, 0)), replicate(stem(nyu12, nyu22, 0))], 1)
    lv = np.concatenate([replicate(stem(nyv11, nyv21, 0)), replicate(stem(nyv12, nyv22, -1))], 1)
    # getbounds plane filter: IB ranges intersected with the non-wall planes
    u_ranges = [(max(nyu11, -d + 1), min(nyu21, n43)), (max(nyu12, -d + 1), min(nyu22, n43))]
    v_ranges = [(max(nyv11, -d + 1), min(nyv21, n33)), (max(nyv12, -d + 1), min(nyv22, n33))]

    def keep(lst, ranges):
        m = np.zeros(lst.shape[1], dtype=bool)
        for lo, hi in ranges:
            m |= (lst[2] >= lo) & (lst[2] <= hi)
        return np.ascontiguousarray(lst[:, m].astype(np.int32))

    lu = keep(lu, u_ranges)
    lv = keep(lv, v_ranges)
    tips = dict(u=(0, n43 - d + 1), v=(-1, n33 - d + 2))
    return lu, lv, tips


# ---------------------------------------------------------------------------------------
# The whole set-up
# ---------------------------------------------------------------------------------------
@dataclasses.dataclass
class ChannelSetup:
    nx: int
    nz: int
    px: int
    pz: int
    ny: int
    d: int
    re: float
    alp: float
    bet: float
    mpgx: float
    flag_ctpress: int
    qxt: float
    grid: YGrid
    gw: np.ndarray
    rk: np.ndarray
    k1x: np.ndarray
    k2x: np.ndarray
    k1z: np.ndarray
    k2z: np.ndarray
    list_u: np.ndarray
    list_v: np.ndarray
    tips: dict
    ntilex: int
    ntilez: int
    dstx: int
    dstz: int
    posth: float

    @property
    def mx(self) -> int:
        return self.nx // 2 + 1

    @property
    def ncol(self) -> int:
        return self.mx * self.nz

    @property
    def nyu(self) -> int:
        return self.ny + 2 * self.d + 1

    @property
    def nyv(self) -> int:
        return self.ny + 2 * self.d

    @property
    def nyp(self) -> int:
        return self.ny + 2 * self.d - 1

    @property
    def grid_points(self) -> int:
        """G = Nx*Nz*Ny with Ny the number of v-grid planes (wall to wall)."""
        return self.nx * self.nz * self.nyv


def make_setup(nx: int, nz: int, ny_total: int, dsty: int, re: float, alp: float = 1.0,
               bet: float = 2.0, mpgx: float = -1.0, flag_ctpress: int = 1,
               posth: float = 0.2, dyq: float = 8.0, ntilex: int = 8, ntilez: int = 4,
               dstx: int = 2, dstz: int = 5, qxt: float = 0.0) -> ChannelSetup:
    """Build a channel whose v grid has ``ny_total`` planes wall to wall, i.e. reference
    input ``ny = ny_total - 2*dsty``.  Physical counts follow the 3/2 rule,
    Ngal = 3N/2 (start.f90:115-118)."""
    if nx % 2 or nz % 2:
        raise ValueError("nx and nz must be even")
    px, pz = 3 * nx // 2, 3 * nz // 2
    ny = ny_total - 2 * dsty
    g = y_grid_canopy(ny, dsty, posth, dyq)
    k1x, k2x, k1z, k2z = def_k(nx, nz, alp, bet)
    lu, lv, tips = boundary_canopies(px, pz, ny, dsty, dstx, dstz, ntilex, ntilez)
    return ChannelSetup(nx=nx, nz=nz, px=px, pz=pz, ny=ny, d=dsty, re=re, alp=alp, bet=bet,
                        mpgx=mpgx, flag_ctpress=flag_ctpress, qxt=qxt, grid=g,
                        gw=gridweighting(g), rk=rk3_tables(), k1x=k1x, k2x=k2x, k1z=k1z,
                        k2z=k2z, list_u=lu, list_v=lv, tips=tips, ntilex=ntilex,
                        ntilez=ntilez, dstx=dstx, dstz=dstz, posth=posth)


def config_small() -> ChannelSetup:
    """Re_tau=180 channel 2pi x 2 x pi, Nx=Nz=64 (Ngal 96), 97 v planes with 12 canopy planes
    per wall, 8x4 posts of height 0.2 and side 0.05pi (2x5 collocation points)."""
    return make_setup(nx=64, nz=64, ny_total=97, dsty=12, re=180.0, alp=1.0, bet=2.0,
                      mpgx=-1.0, flag_ctpress=1, posth=0.2, dyq=6.0, ntilex=8, ntilez=4,
                      dstx=2, dstz=5)


def config_production() -> ChannelSetup:
    """Re_tau=550 channel 2pi x 2 x pi, Nx=Nz=384 (Ngal 576), 289 v planes with 24 canopy
    planes per wall, 48x24 posts of height 0.15 and side 0.02pi (6x12 collocation points)."""
    return make_setup(nx=384, nz=384, ny_total=289, dsty=24, re=550.0, alp=1.0, bet=2.0,
                      mpgx=-1.0, flag_ctpress=1, posth=0.15, dyq=10.0, ntilex=48,
                      ntilez=24, dstx=6, dstz=12)


def _nearest_divisor(n: int, target: float, lo: int, hi: int, also: int | None = None):
    """Divisor of n (and of `also`, when given) in [lo, hi] nearest to target, or None."""
    best = None
    for d in range(max(1, lo), hi + 1):
        if n % d or (also is not None and also % d):
            continue
        if best is None or abs(d - target) < abs(best - target):
            best = d
    return best


def random_instance(seed: int) -> ChannelSetup:
    """One channel of the batched multi-instance workload: nx, nz from {32, 48, 64, 96}, ny
    (v planes wall to wall) from {65, 97, 129}, Re_tau from {180, 300, 395}, canopy post
    spacing uniform in [0.03pi, 0.1pi] (rounded to a tile count that divides the collocation
    grid), post height uniform in [0.1h, 0.25h], post side a third of the spacing."""
    rng = np.random.default_rng(seed)
    for _ in range(100):
        nx = int(rng.choice([32, 48, 64, 96]))
        nz = int(rng.choice([32, 48, 64, 96]))
        ny_total = int(rng.choice([65, 97, 129]))
        re = float(rng.choice([180.0, 300.0, 395.0]))
        posth = float(rng.uniform(0.1, 0.25))
        sx = float(rng.uniform(0.03, 0.1)) * np.pi
        sz = float(rng.uniform(0.03, 0.1)) * np.pi
        px, pz = 3 * nx // 2, 3 * nz // 2
        ntx = _nearest_divisor(px, 2.0 * np.pi / sx, 2, px // 2)
        ntz = _nearest_divisor(pz, np.pi / sz, 2, pz // 2, also=px)
        if ntx is None or ntz is None:
            continue
        dnx, dnz = px // ntx, pz // ntz
        dstx = max(1, min(dnx - 1, round(dnx / 3)))
        dstz = max(1, min(dnz - 1, round(dnz / 3)))
        dsty = max(3, int(round(0.6 * posth * ny_total)))
        try:
            return make_setup(nx=nx, nz=nz, ny_total=ny_total, dsty=dsty, re=re, posth=posth,
                              dyq=6.0, ntilex=ntx, ntilez=ntz, dstx=dstx, dstz=dstz)
        except ValueError:
            continue
    raise ValueError(f"no valid instance for seed {seed}")
