"""GPU parity: the HIP library (through the C-ABI) against the CPU oracle on the same
seeded inputs.  FP64 bar: max-norm relative error per field <= 1e-10."""
import numpy as np
import pytest

from littleharsh_amd.initial import laminar_plus_noise
from oracle.oracle import Oracle

from cases import tiny, small

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)


def report(tag, gpu, cpu):
    names = ("u", "v", "w", "p")
    errs = {n: rel(g, c) for n, g, c in zip(names, gpu, cpu)}
    print(tag, " ".join(f"{k}={v:.3e}" for k, v in errs.items()), flush=True)
    return errs


def make_pair(setup, seed=1234):
    from littleharsh_amd.solver import Channel
    st = laminar_plus_noise(setup, seed=seed)
    g = Channel(setup)
    o = Oracle(setup)
    g.set_state(*st)
    o.set_state(*st)
    return g, o, st


def test_state_roundtrip(require_gpu):
    s = tiny()
    from littleharsh_amd.solver import Channel
    st = laminar_plus_noise(s, seed=7)
    g = Channel(s)
    g.set_state(*st)
    back = g.get_state()
    for a, b in zip(back, st):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("ib", [False, True])
def test_one_substep_tiny(require_gpu, ib):
    s = tiny(ib=ib)
    g, o, _ = make_pair(s)
    dt = 2e-4
    g.substep(1, dt)
    o.substep(1, dt)
    gn, on = g.get_nonlinear(), o.get_nonlinear()
    for name, a, b in zip(("Nu1", "Nu2", "Nu3"), gn, on):
        e = rel(a, b)
        print(name, f"{e:.3e}", flush=True)
        assert e < TOL
    errs = report(f"substep1 ib={ib}", g.get_state(), o.get_state())
    assert max(errs.values()) < TOL


def test_three_substeps_tiny(require_gpu):
    s = tiny()
    g, o, _ = make_pair(s)
    dt = 2e-4
    for k in (1, 2, 3):
        g.substep(k, dt)
        o.substep(k, dt)
        errs = report(f"substep{k}", g.get_state(), o.get_state())
        assert max(errs.values()) < TOL


def test_steps_graph_tiny(require_gpu):
    s = tiny()
    g, o, _ = make_pair(s)
    dt = 2e-4
    for n in range(5):
        g.step(dt)
        o.step(dt)
    errs = report("5 steps", g.get_state(), o.get_state())
    assert max(errs.values()) < TOL
    assert abs(g.scalars()["qx"] - o.flowrate()) < 1e-10 * abs(o.flowrate())
    assert g.max_div() < 1e-9


def test_constant_flow_rate_tiny(require_gpu):
    s = tiny(flag_ctpress=0)
    g, o, _ = make_pair(s)
    dt = 2e-4
    for n in range(3):
        g.step(dt)
        o.step(dt)
    errs = report("ctU 3 steps", g.get_state(), o.get_state())
    assert max(errs.values()) < TOL
    gm, om = g.scalars()["mpgx"], o.scalars()["mpgx"]
    # Each substep adds g = (QxT - Qx0)/Qx1 to mpgx (flowrate_corr, littleharsh.f90:453-485).
    # Qx0 is a sum of n = ny-1 terms dy*u between the tips (flowrateRe) of total ~QxT; the
    # GPU and the oracle add them in different orders, so their g differ by at most
    # n*eps*QxT/|Qx1|.  The unit response u11 is -dRK*dt between the tips up to O(beta), so
    # |Qx1| = dRK*dt*(yv(ny-1) - yv(0)).  The bound is that worst case summed over the nine
    # substeps (about 7e-8 here; a correction error of 1e-7 would fail).
    eps = np.finfo(float).eps
    span = s.grid.yv[s.ny - 1 + s.d] - s.grid.yv[s.d]
    dRK = s.rk.ravel()[12:15]
    bound = 3 * sum((s.ny - 1) * eps * s.qxt / (d * dt * span) for d in dRK)
    print(f"mpgx gpu {gm:.17e} oracle {om:.17e} |diff| {abs(gm - om):.2e} bound {bound:.2e}", flush=True)
    assert abs(gm - om) < bound


def test_one_step_production(require_gpu):
    """The 384 x 289 x 384 channel (Ngal 576, 48 x 24 posts): one full RK3 step, each of its
    three substeps compared."""
    from littleharsh_amd.channel import config_production
    s = config_production()
    g, o, _ = make_pair(s, seed=5)
    dt = 1e-4
    for k in (1, 2, 3):
        g.substep(k, dt)
        o.substep(k, dt)
        errs = report(f"production substep{k}", g.get_state(), o.get_state())
        assert max(errs.values()) < TOL


def test_200_steps_small(require_gpu):
    """64 x 97 x 64 canopy channel, 200 RK3 steps replayed from the HIP graph, with the
    statistics of stats.f90:1-108 (plain and canopy-conditioned) sampled every 10 steps on
    both sides and compared after the last step."""
    from oracle.oracle import split_cond
    s = small()
    g, o, _ = make_pair(s, seed=11)
    dt = 2e-4
    for n in range(200):
        if n % 10 == 0:
            g.stats_accumulate()
            o.stats_accumulate()
        g.step(dt)
        o.step(dt)
        if (n + 1) % 50 == 0:
            errs = report(f"small step {n + 1}", g.get_state(), o.get_state())
            assert max(errs.values()) < TOL
    assert abs(g.scalars()["qx"] - o.flowrate()) < 1e-10 * abs(o.flowrate())
    gs, gn = g.stats()
    os_, on = o.stats()
    assert gn == on == 20
    npts = s.px * s.pz
    worst = 0.0
    for q in range(13):
        # sums of gn samples over npts points: the error bar is the sum of magnitudes
        scale = max(np.abs(os_[q]).max(), gn * npts * 1e-3)
        worst = max(worst, np.abs(gs[q] - os_[q]).max() / scale)
    gc = split_cond(s, g.cond_stats_flat())
    oc = o.cond_stats()
    for q in range(13):
        scale = max(np.abs(oc[q]).max(), gn * 1e-3 * s.ntilex * s.ntilez)
        worst = max(worst, np.abs(gc[q] - oc[q]).max() / scale)
    print(f"small statistics after 200 steps, 20 samples: worst row error {worst:.3e}",
          flush=True)
    assert worst < TOL
