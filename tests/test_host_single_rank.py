"""The reference Fortran host's set-up on one MPI rank: the reference's own start() (its
set-up sources compiled where they lie) with the single-rank proc_lims_planes and getbounds
of littleharsh_amd/fortran/lh_single_rank.f90, followed by lh_config_from_start
(littleharsh_amd/fortran/lh_host.f90).  What the host would hand to lh_create must equal the
Python host's set-up (littleharsh_amd.channel.make_setup) for the same input.in.  CPU only; it
needs the build of littleharsh_amd/fortran/build_host.sh, which needs the reference sources."""
import os
import subprocess

import numpy as np
import pytest

from littleharsh_amd.channel import make_setup

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SETUP = os.path.join(ROOT, "littleharsh_amd", "fortran", "lh_host_setup")

pytestmark = pytest.mark.skipif(not os.path.exists(SETUP),
                                reason="lh_host_setup not built (needs the reference sources)")


def write_input(path, c):
    """input.in in the fixed columns start() reads (start.f90:63-206): formats 10 (7X,D10.1),
    20 (7X,I10), 21 (12X,I5) and 30 (7X,A90), one value or blank line per read."""
    # always with a decimal point: D10.1 reads a field without one as scaled by 10^-1
    d = lambda name, v: f"{name:<7.7}{v:>10.3e}\n"
    i = lambda name, v: f"{name:<7.7}{v:>10d}\n"
    a = lambda name, v: f"{name:<7.7}{v}\n"
    blank = "\n"
    lines = ["! input.in written by the single-rank host test\n"] + [blank] * 8
    lines += [d("Re=", c["re"]), blank, blank, d("alp=", c["alp"]), d("bet=", c["bet"])]
    px, pz = 3 * c["nx"] // 2, 3 * c["nz"] // 2
    lines += [i("Ngalx1=", px), i("Ngalx2=", px), i("Ngalz1=", pz), i("Ngalz2=", pz)]
    lines += [i("ny=", c["ny_total"] - 2 * c["dsty"]), d("dyq=", c["dyq"]), i("pp=", 5), blank, blank]
    lines += [i("flgin=", 3), blank, blank, blank, a("fnmeb=", "none"), blank, a("dirin=", "./"),
              a("dirout=", "./"), blank, blank]
    lines += [d("CFL=", 0.5), i("nwrite=", 10), f"{'flpress=':<12}{c['flag_ctpress']:>5d}\n", blank]
    lines += [d("mpgx=", c["mpgx"]) if c["flag_ctpress"] == 1 else blank, blank, blank, blank]
    lines += [d("Kib=", 1.0), d("maxerr=", 1e-8), d("maxt=", 1.0), d("maxA=", 4.0), i("FFTbot=", 8),
              i("FFTtop=", 8), blank, blank]
    lines += [i("type=", 3), blank, blank, blank, blank, i("ntilex=", c["ntilex"]), i("ntilez=", c["ntilez"])]
    lines += [d("posth=", c["posth"]), i("npeakx=", 1), i("npeakz=", 1), d("Lfracx=", 3.0), d("Lfracz=", 3.0)]
    lines += [i("dsty=", c["dsty"]), i("dstx=", c["dstx"]), i("dstz=", c["dstz"])]
    with open(path, "w") as f:
        f.writelines(lines)


def read_config(path, s):
    raw = open(path, "rb").read()
    off = 0

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(raw, dtype, n, off)
        off += a.nbytes
        return a
    hdr = take("<i4", 8)
    phys = take("<f8", 6)
    arrs = {}
    for name, n in (("yu", s.nyu), ("yv", s.nyv), ("dthdyu", s.nyu), ("dthdyv", s.nyv),
                    ("dyu2i", 3 * s.nyu), ("dyv2i", 3 * s.nyv), ("gw", 2), ("rk", 15),
                    ("k1x", s.mx), ("k2x", s.mx), ("k1z", s.nz), ("k2z", s.nz)):
        arrs[name] = take("<f8", n)
    nlu, nlv = take("<i4", 2)
    lu = take("<i4", 3 * nlu).reshape(nlu, 3).T
    lv = take("<i4", 3 * nlv).reshape(nlv, 3).T
    rest = take("<i4", 8)
    assert off == len(raw)
    return hdr, phys, arrs, lu, lv, rest


CASES = {
    # the 32 x 41 x 32 canopy channel of the GPU tests
    "tiny": dict(nx=32, nz=32, ny_total=41, dsty=6, re=180.0, alp=1.0, bet=2.0, mpgx=-1.0,
                 flag_ctpress=1, posth=0.2, dyq=4.0, ntilex=4, ntilez=2, dstx=3, dstz=5),
    # the Re_tau = 180 channel, 64 x 97 x 64 with 8 x 4 posts
    "small": dict(nx=64, nz=64, ny_total=97, dsty=12, re=180.0, alp=1.0, bet=2.0, mpgx=-1.0,
                  flag_ctpress=1, posth=0.2, dyq=6.0, ntilex=8, ntilez=4, dstx=2, dstz=5),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_start_on_one_rank_hands_over_the_python_setup(tmp_path, case):
    c = CASES[case]
    write_input(tmp_path / "input.in", c)
    r = subprocess.run([SETUP, "config.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    s = make_setup(**c)
    hdr, phys, arrs, lu, lv, rest = read_config(tmp_path / "config.bin", s)
    assert list(hdr) == [s.nx, s.nz, s.px, s.pz, s.ny, s.d, s.flag_ctpress, 0]
    assert np.allclose(phys, [s.re, s.alp, s.bet, s.mpgx, 0.0, s.grid.ddthetavi], rtol=1e-15, atol=0)
    g = s.grid
    for name, ours in (("yu", g.yu), ("yv", g.yv), ("dthdyu", g.dthdyu), ("dthdyv", g.dthdyv),
                       ("dyu2i", g.dyu2i.ravel()), ("dyv2i", g.dyv2i.ravel()), ("gw", s.gw),
                       ("rk", s.rk.ravel()), ("k1x", s.k1x), ("k2x", s.k2x), ("k1z", s.k1z), ("k2z", s.k2z)):
        ref = arrs[name]
        assert np.abs(ref - ours).max() <= 1e-12 * max(np.abs(ref).max(), 1.0), name
    for ref, ours in ((lu, s.list_u), (lv, s.list_v)):
        assert ref.shape == ours.shape
        assert sorted(map(tuple, ref.T.tolist())) == sorted(map(tuple, ours.T.tolist()))
    assert list(rest) == [*s.tips["u"], *s.tips["v"], s.ntilex, s.ntilez, *g.band_v]
