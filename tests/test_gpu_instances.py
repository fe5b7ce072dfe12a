"""The batched multi-instance workload: channels of uneven sizes (Ngal 48, 72, 96, 144 in x
and z, nx != nz, 65 to 129 v planes, Re_tau 180 to 395, different canopies).  Single
instances against the CPU oracle at the FP64 bar, on the default and on the alternative FFT
plans, and one batch step of all 48 instances (plans mixed) against the same instances
stepped one by one."""
import numpy as np
import pytest

from littleharsh_amd.channel import random_instance
from littleharsh_amd.initial import laminar_plus_noise
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)


# seeds 0..3 and 11: (px, pz, v planes) = (144, 96, 97), (72, 96, 129), (144, 72, 65),
# (144, 48, 65), (48, 48, 129)
@pytest.mark.parametrize("seed", [0, 1, 2, 3, 11])
def test_instance_step_matches_oracle(require_gpu, seed):
    from littleharsh_amd.solver import Channel
    s = random_instance(seed)
    st = laminar_plus_noise(s, seed=70 + seed)
    g, o = Channel(s), Oracle(s)
    g.set_state(*st)
    o.set_state(*st)
    dt = g.cfl_dt(0.4)            # CFL 0.4 from the GPU, then the same fixed dt for both
    for n in range(2):
        g.step(dt)
        o.step(dt)
    errs = [rel(a, b) for a, b in zip(g.get_state(), o.get_state())]
    print(f"seed {seed} ({s.px}x{s.nyv}x{s.pz}):", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < TOL
    assert abs(g.scalars()["qx"] - o.flowrate()) < 1e-10 * abs(o.flowrate())


# the alternative FFT plans (lh_set_plans 1, 1) against the oracle: seeds 3, 5, 8, 36, 14, 0
# give (px, pz) = (144, 48), (96, 144), (96, 72), (72, 48), (48, 144), (144, 96), so every
# plan-1 length (48, 72, 96, 144) runs both as the x and as the z transform
@pytest.mark.parametrize("seed", [3, 5, 8, 36, 14, 0])
def test_alternative_plans_match_oracle(require_gpu, seed):
    from littleharsh_amd.solver import Channel
    s = random_instance(seed)
    st = laminar_plus_noise(s, seed=80 + seed)
    g, o = Channel(s), Oracle(s)
    g.set_plans(1, 1)
    assert g.get_plans() == dict(xplan=1, zplan=1, xcount=2, zcount=2)
    g.set_state(*st)
    o.set_state(*st)
    dt = g.cfl_dt(0.4)
    for n in range(2):
        g.step(dt)
        o.step(dt)
    errs = [rel(a, b) for a, b in zip(g.get_state(), o.get_state())]
    print(f"plans 1,1 seed {seed} ({s.px}x{s.nyv}x{s.pz}):", " ".join(f"{e:.2e}" for e in errs),
          flush=True)
    assert max(errs) < TOL


def test_plans_are_checked_and_tuned(require_gpu):
    from littleharsh_amd.solver import Channel, LHError
    s = random_instance(4)                      # px 96, pz 144
    st = laminar_plus_noise(s, seed=90)
    g = Channel(s)
    g.set_state(*st)
    with pytest.raises(LHError):
        g.set_plans(2, 0)
    p = g.autotune(1)
    print("tuned", p, flush=True)
    assert p["xplan"] in (0, 1) and p["zplan"] in (0, 1)
    for x, y in zip(g.get_state(), st):
        assert np.array_equal(x, y)


def test_batch_of_48_instances_matches_single_steps(require_gpu):
    from littleharsh_amd.solver import Batch, Channel
    setups = [random_instance(i) for i in range(48)]
    chans, refs = [], []
    for i, s in enumerate(setups):
        st = laminar_plus_noise(s, seed=100 + i)
        c, r = Channel(s), Channel(s)
        # every third instance on the alternative plans (the batch mixes plans)
        if i % 3 == 1:
            for x in (c, r):
                x.set_plans(1, 1)
        c.set_state(*st)
        r.set_state(*st)
        chans.append(c)
        refs.append(r)
    b = Batch(chans)
    b.step_cfl(0.5)
    for r in refs:
        r.step_cfl(0.5)
    for i, (c, r) in enumerate(zip(chans, refs)):
        for x, y in zip(c.get_state(), r.get_state()):
            assert np.array_equal(x, y), i
    print("48 instances bit-identical after one batch step", flush=True)
    b.close()
