"""The divergence monitor (error, littleharsh.f90:323-352) against the oracle on a field with
O(1) divergence, and non-finite detection: a NaN in the state shows in lh_max_div and makes
the step calls return LH_ERR_NUMERIC, with every kernel running to completion."""
import numpy as np
import pytest

from littleharsh_amd import capi
from littleharsh_amd.initial import laminar_plus_noise
from oracle.oracle import Oracle

from cases import tiny

pytestmark = pytest.mark.gpu


def random_state(s, seed):
    rng = np.random.default_rng(seed)
    cr = lambda n: rng.standard_normal((s.ncol, n)) + 1j * rng.standard_normal((s.ncol, n))
    return cr(s.nyu), cr(s.nyv), cr(s.nyu), cr(s.nyp)


@pytest.mark.parametrize("ib", [False, True])
def test_max_div_matches_oracle_on_divergent_field(require_gpu, ib):
    from littleharsh_amd.solver import Channel
    s = tiny(ib=ib)
    st = random_state(s, seed=61)
    g, o = Channel(s), Oracle(s)
    g.set_state(*st)
    o.set_state(*st)
    a, b = g.max_div(), o.max_div()
    print(f"max|div| gpu {a:.17e} oracle {b:.17e}", flush=True)
    assert b > 1.0                      # a field far from solenoidal
    assert abs(a - b) <= 1e-12 * b
    # after one step both are divergence-free to round-off
    g.step(2e-4)
    assert g.max_div() < 1e-9 * b


def test_nan_in_state_is_reported(require_gpu):
    from littleharsh_amd.solver import Channel, LHError
    s = tiny()
    good = laminar_plus_noise(s, seed=62)
    g = Channel(s)
    bad = [a.copy() for a in good]
    bad[0][5, 10] = np.nan              # one spectral coefficient of u
    g.set_state(*bad)
    assert np.isnan(g.max_div())
    with pytest.raises(LHError) as e:
        g.step_cfl(0.5)
    print(e.value, flush=True)
    assert e.value.status == capi.LH_ERR_NUMERIC
    g.set_state(*bad)
    with pytest.raises(LHError) as e:
        g.step(2e-4)
    assert e.value.status == capi.LH_ERR_NUMERIC
    g.set_state(*bad)
    with pytest.raises(LHError) as e:
        g.substep(1, 2e-4)
    assert e.value.status == capi.LH_ERR_NUMERIC
    with pytest.raises(LHError) as e:
        g.cfl_dt(0.5)
    assert e.value.status == capi.LH_ERR_NUMERIC
    # the handle still works once a finite state is set again
    g.set_state(*good)
    dt = g.step_cfl(0.5)
    assert np.isfinite(dt) and dt > 0
    assert g.max_div() < 1e-9


def test_batch_reports_the_instance_with_nan(require_gpu):
    from littleharsh_amd.solver import Batch, Channel, LHError
    s = tiny()
    st = laminar_plus_noise(s, seed=63)
    a, b = Channel(s), Channel(s)
    a.set_state(*st)
    bad = [x.copy() for x in st]
    bad[2][7, 3] = np.inf               # w
    b.set_state(*bad)
    bt = Batch([a, b])
    with pytest.raises(LHError) as e:
        bt.step_cfl(0.5)
    print(e.value, flush=True)
    assert e.value.status == capi.LH_ERR_NUMERIC
    assert "instance 1" in str(e.value)
    bt.close()
