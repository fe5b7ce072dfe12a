"""The file formats of the reference (littleharsh_amd/formats.py) against files written by
the reference's own writers (tests/golden/ref_small.npz, made by tests/golden/make_golden.py
from y_grid_canopy, boundary_canopies, write_stats and recvwrspec)."""
import os

import numpy as np
import pytest

from littleharsh_amd import formats
from littleharsh_amd.channel import make_setup

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_small.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD)


@pytest.fixture(scope="module")
def s():
    return make_setup(nx=16, nz=8, ny_total=19, dsty=3, re=180.0, alp=1.0, bet=2.0, mpgx=-1.0,
                      posth=0.2, dyq=3.0, ntilex=4, ntilez=2, dstx=2, dstz=3)


def _bytes(p):
    return open(p, "rb").read()


def test_y_grid_file(g, s, tmp_path):
    p = tmp_path / "y_grid_canopy.dat"
    formats.write_y_grid(str(p), s)
    ref = g["file_y_grid"].tobytes()
    ours = _bytes(p)
    assert len(ours) == len(ref)
    # the grid itself is pinned to 1e-13 (test_oracle_golden); the header bytes exactly
    assert ours[:84] == ref[:84]
    a = np.frombuffer(ours[84:], "<f8")
    b = np.frombuffer(ref[84:], "<f8")
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_boundary_file(g, s, tmp_path):
    p = tmp_path / "boundary.dat"
    formats.write_boundary(str(p), s)
    ref = g["file_boundary"].tobytes()
    ours = _bytes(p)
    assert len(ours) == len(ref)
    assert np.allclose(np.frombuffer(ours[:24], "<f8"), np.frombuffer(ref[:24], "<f8"), rtol=1e-15)
    assert ours[24:] == ref[24:]


def test_spectra_files(g, s, tmp_path):
    t, it, istat, mpgx = g["wr_meta"]
    sp = np.zeros((5, s.ncol, s.nyu))
    sp[0] = g["spec_u"]
    sp[1][:, :s.nyv] = g["spec_v"]
    sp[2] = g["spec_w"]
    sp[3] = g["spec_uv"]
    sp[4][:, :s.nyp] = g["spec_p"]
    paths = formats.write_spectra(str(tmp_path), s, sp, int(istat), float(t), int(it), float(mpgx))
    for p, tag in zip(paths, ("U", "V", "W", "UV", "P")):
        ours = formats.read_records(p)
        ref_path = tmp_path / f"ref_{tag}.dat"
        ref_path.write_bytes(g[f"file_sp{tag}"].tobytes())
        ref = formats.read_records(str(ref_path))
        assert len(ours) == len(ref), tag
        assert ours[0] == ref[0] and ours[1] == ref[1] and ours[4] == ref[4], tag
        for q in (2, 3, 5):
            a, b = np.frombuffer(ours[q], "<f8"), np.frombuffer(ref[q], "<f8")
            assert a.shape == b.shape and np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1.0), (tag, q)


def test_stats_file(g, s, tmp_path):
    """write_stats of one rank holding the bottom-band planes (plain rows there, zero
    elsewhere) and that band's conditional arrays."""
    t, it, istat, mpgx = g["wr_meta"]
    rows = np.array(g["stats"])
    n31 = s.grid.band_v[0]
    nb = {0: n31 + s.d + 2, 1: n31 + s.d + 1, 2: n31 + s.d}
    grid = (0, 0, 1, 1, 0, 0, 2, 2, 0, 0, 1, 1, 1)
    for q in range(13):
        rows[q, nb[grid[q]]:] = 0.0
    name = bytes(g["file_stats_name"]).decode()
    assert name == f"stats_{formats.file_tag(s, float(t))}.dat"
    p = tmp_path / name
    formats.write_stats(str(p), s, rows, g["cstats_bot"], int(istat), float(t), int(it), float(mpgx))
    ref_path = tmp_path / "ref_stats.dat"
    ref_path.write_bytes(g["file_stats"].tobytes())
    ours, ref = formats.read_records(str(p)), formats.read_records(str(ref_path))
    assert len(ours) == len(ref) == 5 + 13 + 13
    assert ours[0] == ref[0] and ours[1] == ref[1] and ours[4] == ref[4]
    for q in list(range(2, 4)) + list(range(5, 31)):
        a, b = np.frombuffer(ours[q], "<f8"), np.frombuffer(ref[q], "<f8")
        assert a.shape == b.shape and np.abs(a - b).max() <= 1e-12 * max(np.abs(b).max(), 1.0), q


def test_restart_round_trip_and_resolution_change(s, tmp_path):
    rng = np.random.default_rng(3)
    st = [rng.standard_normal((s.ncol, n)) + 1j * rng.standard_normal((s.ncol, n))
          for n in (s.nyu, s.nyv, s.nyu, s.nyp)]
    formats.write_restart(str(tmp_path), ".001", s, st, 1.5, 42, -1.0)
    back, meta = formats.read_restart(str(tmp_path), ".001", s)
    assert meta["iter"] == 42 and meta["t"] == 1.5
    for a, b in zip(back, st):
        assert np.array_equal(a, b)
    # read into a finer grid: shared modes copied, new modes zero
    s2 = make_setup(nx=32, nz=16, ny_total=19, dsty=3, re=180.0, alp=1.0, bet=2.0, mpgx=-1.0,
                    posth=0.2, dyq=3.0, ntilex=4, ntilez=2, dstx=2, dstz=3)
    fine, _ = formats.read_restart(str(tmp_path), ".001", s2)
    u = fine[0].reshape(s2.nz, s2.mx, s2.nyu)
    uc = st[0].reshape(s.nz, s.mx, s.nyu)
    assert np.array_equal(u[:s.nz // 2, :s.mx], uc[:s.nz // 2])
    assert np.array_equal(u[s2.nz - s.nz // 2:, :s.mx], uc[s.nz // 2:])
    assert not u[s.nz // 2:s2.nz - s.nz // 2].any() and not u[:, s.mx:].any()
