"""The CPU oracle and the host set-up against vectors produced by the reference's own
Fortran routines (tests/golden/ref_small.npz, written by tests/golden/make_golden.py from
the reference sources).  No GPU needed."""
import os

import numpy as np
import pytest

from littleharsh_amd.channel import make_setup, def_k, y_grid_canopy, gridweighting, boundary_canopies
from oracle.oracle import Oracle, cfft

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_small.npz")
CFG = dict(nx=16, nz=8, ny_total=19, dsty=3, posth=0.2, dyq=3.0, re=180.0, alp=1.0, bet=2.0,
           mpgx=-1.0, ntilex=4, ntilez=2, dstx=2, dstz=3)


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD)


@pytest.fixture(scope="module")
def setup():
    c = CFG
    return make_setup(nx=c["nx"], nz=c["nz"], ny_total=c["ny_total"], dsty=c["dsty"], re=c["re"],
                      alp=c["alp"], bet=c["bet"], mpgx=c["mpgx"], posth=c["posth"], dyq=c["dyq"],
                      ntilex=c["ntilex"], ntilez=c["ntilez"], dstx=c["dstx"], dstz=c["dstz"])


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def test_grid_matches_y_grid_canopy(g, setup):
    gr = setup.grid
    assert rel(gr.yu, g["yu"]) < 1e-14
    assert rel(gr.yv, g["yv"]) < 1e-14
    assert rel(gr.dthdyu, g["dthdyu"]) < 1e-13
    assert rel(gr.dthdyv, g["dthdyv"]) < 1e-13
    assert rel(gr.dyu2i, g["dyu2i"]) < 1e-12
    assert rel(gr.dyv2i, g["dyv2i"]) < 1e-12
    assert rel(setup.gw, g["gw"]) < 1e-13
    n40, n43, n30, n33, nn, ncol, b1, b2 = g["sizes"]
    assert (n40, n43, n30, n33) == (-setup.d, setup.ny - 1 + setup.d, -setup.d, setup.ny - 2 + setup.d)
    assert gr.band_v == (b1, b2)


def test_wavenumbers_match_def_k(g):
    k1x, k2x, k1z, k2z = def_k(CFG["nx"], CFG["nz"], CFG["alp"], CFG["bet"])
    for a, name in zip((k1x, k2x, k1z, k2z), ("k1x", "k2x", "k1z", "k2z")):
        assert np.array_equal(a, g[name]), name


def test_canopy_lists_match_boundary_canopies(g, setup):
    """Effective lists = the reference lists restricted to the non-wall IB planes that
    getbounds keeps (start.f90:767-801)."""
    d, ny = setup.d, setup.ny
    n33, n43 = ny - 2 + d, ny - 1 + d
    for grid, lo, hi in (("u", -d + 1, n43), ("v", -d + 1, n33)):
        ref = np.concatenate([g[f"list1_{grid}"], g[f"list2_{grid}"]], axis=1)
        # IB plane ranges of geom_sym_canopy.f90:67-76 intersected with the interior planes
        if grid == "u":
            ranges = [(-d + 1, 0), (n43 - d + 1, n43)]
        else:
            ranges = [(-d, -1), (n33 + 2 - d, n33 + 1)]
        keep = np.zeros(ref.shape[1], bool)
        for a, b in ranges:
            keep |= (ref[2] >= max(a, lo)) & (ref[2] <= min(b, hi))
        ref = ref[:, keep]
        ours = setup.list_u if grid == "u" else setup.list_v
        assert sorted(map(tuple, ref.T.tolist())) == sorted(map(tuple, ours.T.tolist())), grid


def test_velocity_lu_factors(g, setup):
    o = Oracle(setup)
    for grid_u, name in ((True, "u"), (False, "v")):
        ay, axz = o.lu_vel(grid_u, 2, 3e-4)
        ref = g[f"ay_{name}"].copy()
        # the reference leaves ay(1,nystart) and ay(3,nyend) at their built values (unused)
        assert rel(ay[1:, 0], ref[1:, 0]) < 1e-14
        assert rel(ay[:, 1], ref[:, 1]) < 1e-14
        assert rel(ay[:-1, 2], ref[:-1, 2]) < 1e-14
        assert rel(axz, g[f"axz_{name}"]) < 1e-15


def _setup_without_canopy():
    c = CFG
    s = make_setup(nx=c["nx"], nz=c["nz"], ny_total=c["ny_total"], dsty=c["dsty"], re=c["re"],
                   alp=c["alp"], bet=c["bet"], mpgx=c["mpgx"], posth=c["posth"], dyq=c["dyq"],
                   ntilex=c["ntilex"], ntilez=c["ntilez"], dstx=c["dstx"], dstz=c["dstz"])
    s.list_u = s.list_u[:, :0]
    s.list_v = s.list_v[:, :0]
    return s


@pytest.mark.parametrize("canopy", [False, True])
@pytest.mark.parametrize("grid", ["u", "v"])
def test_velocity_solve_matches_dense_solve(g, setup, grid, canopy):
    """solveU / solveV with LUsolU / LUsolV (littleharsh.f90:733-824, tridLU_3D_IB.f90:50-194):
    the boundary rows of the rhs are 0, rows 1..nj-2 are (u + dt du) / axz, the canopy
    planes get the immersed-boundary forcing (pinned by its own brute-force check), and the
    new velocity solves the system whose rows the reference's LU_build wrote (golden, before
    LU_dec).  Checked by a dense solve per column against the oracle's Thomas sweeps."""
    s = setup if canopy else _setup_without_canopy()
    o = Oracle(s)
    krk, dt = 2, 3e-4            # the golden LU_build was called with these
    beta = (1.0 / 15.0) * dt / s.re
    nj = s.nyu if grid == "u" else s.nyv
    rng = np.random.default_rng(101 if grid == "u" else 102)
    cr = lambda: rng.standard_normal((s.ncol, nj)) + 1j * rng.standard_normal((s.ncol, nj))
    u, du, lu = cr(), cr(), cr()
    ours = o.solve_vel(grid == "u", krk, dt, u, du, lu)
    rows, axz = g[f"ayb_{grid}"], g[f"axzb_{grid}"]
    assert rel(axz, g[f"axz_{grid}"]) == 0.0
    M = np.diag(rows[:, 1]) + np.diag(rows[1:, 0], -1) + np.diag(rows[:-1, 2], 1)
    b = np.zeros((s.ncol, nj), complex)
    b[:, 1:-1] = (u[:, 1:-1] + dt * du[:, 1:-1]) / axz[:, None]
    if canopy:
        b = o.ib(grid == "u", beta, b, lu)
    ref = np.linalg.solve(M, b.T).T
    assert rel(ours, ref) < 1e-13
    # the boundary rows hold the wall conditions: u rows 0 and nj-1 are tied to their
    # neighbours by gridweighting, v is zero at the walls
    if grid == "u":
        assert np.abs(ours[:, 0] + rows[0, 2] * ours[:, 1]).max() < 1e-14 * np.abs(ours).max()
    else:
        assert np.abs(ours[:, 0]).max() == 0.0 and np.abs(ours[:, -1]).max() == 0.0
    if canopy:
        # a planted mistake must show: leaving out the division by axz
        wrong = b.copy()
        wrong[:, 1:-1] *= axz[:, None]
        assert rel(ours, np.linalg.solve(M, wrong.T).T) > 1e-9


def test_pressure_lu_factors(g, setup):
    o = Oracle(setup)
    dg = o.dg()
    ref = g["dg"]
    assert rel(dg[:, :, 1], ref[:, :, 1]) < 1e-13
    assert rel(dg[:, 1:, 0], ref[:, 1:, 0]) < 1e-14
    assert rel(dg[:, :-1, 2], ref[:, :-1, 2]) < 1e-13


@pytest.mark.parametrize("krk", [1, 2, 3])
def test_rhs0(g, setup, krk):
    o = Oracle(setup)
    o.set_state(g["in_u1"], g["in_u2"], g["in_u3"], g["in_p"])
    full = []
    for name, n in (("in_nu1", setup.nyu), ("in_nu2", setup.nyv), ("in_nu3", setup.nyu)):
        a = np.zeros((setup.ncol, n), complex)
        a[:, 1:-1] = g[name]
        full.append(a)
    du1, du2, du3, lu1, lu2, lu3 = o.rhs0(krk, full)
    for ours, nm in zip((du1, du2, du3, lu1, lu2, lu3), ("du1", "du2", "du3", "lu1", "lu2", "lu3")):
        ref = g[f"rhs0_k{krk}_{nm}"]
        assert rel(ours[:, 1:-1], ref[:, 1:-1]) < 1e-13, nm


@pytest.mark.parametrize("krk", [1, 3])
def test_solvep_vcorr(g, setup, krk):
    o = Oracle(setup)
    o.set_state(g["in_u1"], g["in_u2"], g["in_u3"], g["in_p"])
    o.solvep_vcorr(krk, 2e-4)
    u1, u2, u3, p = o.get_state()
    for ours, nm in zip((u1, u2, u3, p), ("u1", "u2", "u3", "p")):
        assert rel(ours, g[f"sp_k{krk}_{nm}"]) < 1e-12, nm
    assert rel(o.work(6), g[f"sp_k{krk}_psi"]) < 1e-12
    # div after v_corr is round-off sized on both sides: compare absolutely against the
    # velocity scale
    scale = np.abs(g["in_u1"]).max() * np.abs(setup.k1x).max()
    assert np.abs(o.work(7) - g[f"sp_k{krk}_div"]).max() < 1e-12 * scale


def test_flowrate_and_correction(g, setup):
    o = Oracle(setup)
    u1 = np.zeros((setup.ncol, setup.nyu), complex)
    u1[0] = g["fr_col"]
    z = np.zeros((setup.ncol, setup.nyv), complex)
    o.set_state(u1, z, np.zeros_like(u1), np.zeros((setup.ncol, setup.nyp), complex))
    assert rel(o.flowrate(), g["fr_q"][0]) < 1e-14
    gg, mp = o.flowrate_corr(2, 2e-4, 120.0, CFG["mpgx"])
    assert rel(gg, g["frc_g"][0]) < 1e-12
    assert rel(mp, g["frc_mpgx"][0]) < 1e-12
    assert rel(o.get_state()[0][0], g["frc_u"]) < 1e-13


def test_spectra_accumulation(g, setup):
    o = Oracle(setup)
    o.set_state(g["in_u1"], g["in_u2"], g["in_u3"], g["in_p"])
    o.spectra_accumulate()
    sp = o.spectra()
    assert rel(sp[0], g["spec_u"]) < 1e-15
    assert rel(sp[1][:, :setup.nyv], g["spec_v"]) < 1e-15
    assert rel(sp[4][:, :setup.nyp], g["spec_p"]) < 1e-15
    u, v = g["in_u1"], g["in_u3"]
    assert rel(u.real * v.real + u.imag * v.imag, g["spec_uv"]) < 1e-15


@pytest.mark.parametrize("n", [12, 24, 48, 96])
def test_complex_fft_against_fftpack(g, n):
    z = g[f"cft_in_{n}"]
    assert rel(cfft(z, -1) / n, g[f"cft_fwd_{n}"]) < 1e-14
    assert rel(cfft(z, +1), g[f"cft_inv_{n}"]) < 1e-14


@pytest.mark.parametrize("n", [12, 24, 48, 96])
def test_real_fft_convention(g, n):
    """rft forward: cc(k) = c(2k+1) + i c(2k+2) = (1/n) sum x e^{-2 pi i jk/n}, Im cc(0) =
    Im cc(n/2) = 0 (rft_buff.f:63-73); inverse unscaled (rft_buff.f:77-84)."""
    x = g[f"rft_in_{n}"]
    f = g[f"rft_fwd_{n}"]
    cc = f[0::2] + 1j * f[1::2]
    assert rel(cc, np.fft.rfft(x) / n) < 1e-14
    assert f[1] == 0.0 and f[n + 1] == 0.0
    assert rel(g[f"rft_inv_{n}"][:n], x) < 1e-14


def test_plane_transforms_against_fftpack(g, setup):
    o = Oracle(setup)
    sp = o.phys_to_four(g["plane_phys"]).reshape(setup.nz, setup.mx)
    assert rel(sp, g["plane_spec"]) < 1e-14
    back = o.four_to_phys(g["plane_spec"].ravel())
    assert rel(back, g["plane_back"]) < 1e-13


def test_statistics_against_stats_f90(g, setup):
    """Plain and conditional statistics (stats.f90:1-253) of the seeded state, against the
    reference's stats/conddown/condup run on planes built with rft/cft.  The conditional
    arrays are the sum of the bottom and the mirrored top band, as recvXC writes them
    (stats.f90:474-508)."""
    from oracle.oracle import split_cond
    o = Oracle(setup)
    o.set_state(g["in_u1"], g["in_u2"], g["in_u3"], g["in_p"])
    o.stats_accumulate()
    st, n = o.stats()
    assert n == 1
    ref = g["stats"]
    npts = setup.px * setup.pz
    for q in range(13):
        # sums of O(1) values over npts points: compare against the sum of magnitudes
        scale = npts * max(np.abs(ref[q]).max() / npts, 1.0)
        assert np.abs(st[q] - ref[q]).max() < 1e-12 * scale, q
    cond = o.cond_stats()
    both = split_cond(setup, g["cstats_bot"] + g["cstats_top"])
    for q in range(13):
        scale = setup.ntilex * setup.ntilez * max(np.abs(both[q]).max(), 1.0)
        assert np.abs(cond[q] - both[q]).max() < 1e-12 * scale, q
