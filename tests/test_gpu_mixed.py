"""Mixed precision (FP32 nonlinear stage: z/x transforms and products; FP64 solves,
Poisson, (0,0) mode, flow rate and statistics) against the FP64 GPU run."""
import numpy as np
import pytest

from littleharsh_amd import capi
from littleharsh_amd.initial import laminar_plus_noise

from cases import small

pytestmark = pytest.mark.gpu


# plans (1, 1): the FP32 transform kernels of the alternative radix sequences (8 x 12 at
# 96 points), which the autotuner may pick for a mixed-precision channel
@pytest.mark.parametrize("plans", [(0, 0), (1, 1)])
def test_mixed_precision_tracks_fp64(require_gpu, plans):
    from littleharsh_amd.solver import Channel
    s = small()
    st = laminar_plus_noise(s, seed=31)
    g64 = Channel(s)
    g32 = Channel(s, precision=capi.LH_PREC_MIXED)
    g32.set_plans(*plans)
    g64.set_state(*st)
    g32.set_state(*st)
    for _ in range(20):
        g64.step(2e-4)
        g32.step(2e-4)
    a, b = g64.get_state(), g32.get_state()
    for name, x, y in zip("uvwp", a, b):
        e = np.abs(x - y).max() / np.abs(x).max()
        print(f"mixed vs fp64 {name}: {e:.3e}", flush=True)
        # FP32 products carry ~1e-7 relative error per substep into Nu; over 60
        # substeps the fields stay within 1e-4 of the FP64 run
        assert e < 1e-4
    # mean profile U(y) (the (0,0) mode) within 0.5 per mille
    mu64, mu32 = a[0][0].real, b[0][0].real
    assert np.abs(mu64 - mu32).max() < 5e-4 * np.abs(mu64).max()
    assert g32.max_div() < 1e-9


def _rel(a, b):
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)


def test_mixed_production_substeps_against_fp64_oracle(require_gpu):
    """Mixed precision at 384x289x384 (Ngal 576, 48x24 posts), two substeps against the FP64
    oracle.  The FP32 transforms and products carry a relative error of order
    u32 log2(576) ~ 6e-7 per transform into Nu (u32 = 6e-8); it enters the velocity as
    dt gRK dNu and the pressure through the Poisson solve.  Near the walls the y-difference
    m (P(j) - P(j-1)) of Nu subtracts FP32 products much larger than Nu itself, which lifts
    the max-norm error of Nu1 to about 1e-5.  Measured in substep 2: 9.96e-6 and 1.03e-5 with
    the block x transform (two builds that differ only in how A = ikx P0 + m (Ph - Pl) is
    contracted into FMAs, a change of one ulp per product), 1.60e-5 with the one-wave-per-line
    transform, which rounds more often per point (one point in nine passes six radix-2 stages
    with twiddles).  Stated bound: 3e-5 relative in max norm on Nu and on u, v, w, p, about
    twice the largest observed and far below the O(1e-3..1) error of a wrong row, mode or
    twiddle."""
    from littleharsh_amd.channel import config_production
    from littleharsh_amd.initial import laminar_plus_noise as lpn
    from littleharsh_amd.solver import Channel
    from oracle.oracle import Oracle
    s = config_production()
    st = lpn(s, seed=5)
    g = Channel(s, precision=capi.LH_PREC_MIXED)
    o = Oracle(s)
    g.set_state(*st)
    o.set_state(*st)
    dt = 1e-4
    for k in (1, 2):
        g.substep(k, dt)
        o.substep(k, dt)
        nerr = [_rel(a, b) for a, b in zip(g.get_nonlinear(), o.get_nonlinear())]
        serr = [_rel(a, b) for a, b in zip(g.get_state(), o.get_state())]
        print(f"mixed production substep {k}: Nu", " ".join(f"{e:.2e}" for e in nerr),
              "| u v w p", " ".join(f"{e:.2e}" for e in serr), flush=True)
        assert max(nerr) < 3e-5
        assert max(serr) < 3e-5


def test_mixed_production_statistics_track_fp64(require_gpu):
    """Mixed against FP64 at 384x289x384 over 150 RK3 steps at the FP64 run's CFL-0.5 time
    step: the mean profile U(y) within 0.5 %, the rms profiles u', v', w' and uv within 2 %
    (the mixed-precision limits), from statistics sampled every 10 steps."""
    from littleharsh_amd.channel import config_production
    from littleharsh_amd.initial import laminar_plus_modes
    from littleharsh_amd.precision_report import profiles, rel_err
    from littleharsh_amd.solver import Channel
    s = config_production()
    st = laminar_plus_modes(s, seed=1234)
    res = {}
    dt = None
    for prec in (capi.LH_PREC_FP64, capi.LH_PREC_MIXED):
        g = Channel(s, precision=prec)
        g.set_state(*st)
        if dt is None:
            dt = g.cfl_dt(0.5)
        for n in range(150):
            g.step(dt)
            if (n + 1) % 10 == 0:
                g.stats_accumulate()
        a, ns = g.stats()
        res[prec] = profiles(s, a, ns)
        del g
    p64, p32 = res[capi.LH_PREC_FP64], res[capi.LH_PREC_MIXED]
    errs = {k: rel_err(p32[k], p64[k]) for k in p64}
    print("mixed vs fp64 after 150 steps, dt", dt, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert errs["U"] <= 0.005
    assert max(errs.values()) <= 0.02
