"""GPU statistics, spectra and CFL time step against the CPU oracle (FP64, 1e-10)."""
import numpy as np
import pytest

from littleharsh_amd.initial import laminar_plus_noise
from oracle.oracle import Oracle, split_cond

from cases import tiny

pytestmark = pytest.mark.gpu


def make_pair(setup, seed):
    from littleharsh_amd.solver import Channel
    st = laminar_plus_noise(setup, seed=seed)
    g = Channel(setup)
    o = Oracle(setup)
    g.set_state(*st)
    o.set_state(*st)
    return g, o


def test_statistics_match_oracle(require_gpu):
    s = tiny()
    g, o = make_pair(s, seed=21)
    for _ in range(2):
        g.stats_accumulate()
        o.stats_accumulate()
        g.step(2e-4)
        o.step(2e-4)
    gs, gn = g.stats()
    os_, on = o.stats()
    assert gn == on == 2
    npts = s.px * s.pz
    for q in range(13):
        # sums over npts points: the error bar is set by the sum of magnitudes
        scale = max(np.abs(os_[q]).max(), npts * 1e-3)
        e = np.abs(gs[q] - os_[q]).max() / scale
        print("stats row", q, f"{e:.3e}", flush=True)
        assert e < 1e-10
    gc = split_cond(s, g.cond_stats_flat())
    oc = o.cond_stats()
    for q in range(13):
        scale = max(np.abs(oc[q]).max(), 1e-3 * s.ntilex * s.ntilez)
        e = np.abs(gc[q] - oc[q]).max() / scale
        print("cond row", q, f"{e:.3e}", flush=True)
        assert e < 1e-10
    # reset clears both
    g.stats(reset=True)
    z, n = g.stats()
    assert n == 0 and not z.any() and not g.cond_stats_flat().any()


def test_spectra_match_oracle(require_gpu):
    s = tiny()
    g, o = make_pair(s, seed=22)
    for _ in range(2):
        g.spectra_accumulate()
        o.spectra_accumulate()
        g.step(2e-4)
        o.step(2e-4)
    a, b = g.spectra(), o.spectra()
    for f in range(5):
        e = np.abs(a[f] - b[f]).max() / np.abs(b[f]).max()
        print("spectrum", f, f"{e:.3e}", flush=True)
        assert e < 1e-10


def test_cfl_dt_matches_oracle(require_gpu):
    s = tiny()
    g, o = make_pair(s, seed=23)
    o.nonlinear(1)                     # evaluates the CFL maximum at kRK = 1
    cmax = o.scalars()["cflmax"]
    dt = g.cfl_dt(0.5)
    assert abs(dt - 0.5 / cmax) < 1e-12 * (0.5 / cmax)
    # the device dt is what a substep with dt <= 0 uses
    assert abs(g.scalars()["dt"] - dt) == 0.0


def test_step_cfl_matches_oracle_with_its_dt(require_gpu):
    s = tiny()
    g, o = make_pair(s, seed=24)
    for n in range(3):
        o.nonlinear(1)
        dto = 0.5 / o.scalars()["cflmax"]
        dtg = g.step_cfl(0.5)
        assert abs(dtg - dto) < 1e-11 * dto
        o.step(dto)
    a, b = g.get_state(), o.get_state()
    for x, y in zip(a, b):
        assert np.abs(x - y).max() / np.abs(y).max() < 1e-10
