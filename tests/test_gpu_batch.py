"""Launch parameters, the autotuner and the multi-instance batch (one graph launch per step
for channels of different sizes) give the same results as the plain per-handle steps."""
import numpy as np
import pytest

from littleharsh_amd.channel import random_instance
from littleharsh_amd.initial import laminar_plus_noise

from cases import tiny

pytestmark = pytest.mark.gpu


def test_launch_parameters_do_not_change_results(require_gpu):
    from littleharsh_amd.solver import Channel
    s = tiny()
    st = laminar_plus_noise(s, seed=41)
    a, b = Channel(s), Channel(s)
    a.set_state(*st)
    b.set_state(*st)
    b.set_launch(4, 1, 128)
    assert b.get_launch() == dict(xraw_lc=4, zasm_lc=1, col_bs=128)
    for _ in range(2):
        a.step(2e-4)
        b.step(2e-4)
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y)


def test_autotune_keeps_the_state(require_gpu):
    from littleharsh_amd.solver import Channel
    s = tiny()
    st = laminar_plus_noise(s, seed=42)
    g = Channel(s)
    g.set_state(*st)
    p = g.autotune(1)
    print("tuned", p, flush=True)
    # candidates: the modelled defaults, their halves and doubles, and a few fixed sizes
    assert p["col_bs"] in (64, 128, 256) and 1 <= p["xraw_lc"] <= 128 and 1 <= p["zasm_lc"] <= 24
    for x, y in zip(g.get_state(), st):
        assert np.array_equal(x, y)


def test_batch_matches_single_instances(require_gpu):
    from littleharsh_amd.solver import Batch, Channel
    setups = [random_instance(i) for i in (0, 1, 2)]
    states = [laminar_plus_noise(s, seed=50 + i) for i, s in enumerate(setups)]
    chans, refs = [], []
    for s, st in zip(setups, states):
        c, r = Channel(s), Channel(s)
        c.set_state(*st)
        r.set_state(*st)
        chans.append(c)
        refs.append(r)
    b = Batch(chans)
    for _ in range(2):
        b.step_cfl(0.5)
        for r in refs:
            r.step_cfl(0.5)
    for c, r in zip(chans, refs):
        for x, y in zip(c.get_state(), r.get_state()):
            assert np.array_equal(x, y)
    b.close()


def test_batch_with_a_large_member(require_gpu):
    """A batch whose members use both velocity sweeps: two small channels (4 rows per step)
    and a 256 x 33 x 256 one (33024 spectral columns, 1 row per step, px = pz = 384), so the
    column launches cover two block ranges; every member bit-identical to its single steps."""
    from littleharsh_amd.channel import make_setup
    from littleharsh_amd.solver import Batch, Channel
    big = make_setup(nx=256, nz=256, ny_total=33, dsty=4, re=180.0, alp=1.0, bet=2.0,
                     mpgx=-1.0, posth=0.2, dyq=4.0, ntilex=16, ntilez=8, dstx=6, dstz=12)
    setups = [random_instance(3), big, random_instance(11)]
    chans, refs = [], []
    for i, s in enumerate(setups):
        st = laminar_plus_noise(s, seed=60 + i)
        c, r = Channel(s), Channel(s)
        c.set_state(*st)
        r.set_state(*st)
        chans.append(c)
        refs.append(r)
    b = Batch(chans)
    for _ in range(2):
        b.step_cfl(0.5)
        for r in refs:
            r.step_cfl(0.5)
    for i, (c, r) in enumerate(zip(chans, refs)):
        for x, y in zip(c.get_state(), r.get_state()):
            assert np.array_equal(x, y), i
    b.close()


def test_batch_recaptures_when_a_member_changes(require_gpu):
    """Switching a member's forcing on (and off) after the batch graph was captured must be
    seen by the next batch step: the batch re-captures when a member's graphs were dropped."""
    from littleharsh_amd.solver import Batch, Channel
    s = tiny()
    st = laminar_plus_noise(s, seed=44)
    a, b, ra, rb = Channel(s), Channel(s), Channel(s), Channel(s)
    for c in (a, b, ra, rb):
        c.set_state(*st)
    rng = np.random.default_rng(45)
    f = [rng.standard_normal((s.ncol, n)) * 1e-2 + 0j for n in (s.nyu - 2, s.nyv - 2, s.nyu - 2)]
    bt = Batch([a, b])
    bt.step_cfl(0.5)
    ra.step_cfl(0.5)
    rb.step_cfl(0.5)
    b.set_forcing(*f)
    rb.set_forcing(*f)
    bt.step_cfl(0.5)
    ra.step_cfl(0.5)
    rb.step_cfl(0.5)
    b.set_forcing()
    rb.set_forcing()
    bt.step_cfl(0.5)
    ra.step_cfl(0.5)
    rb.step_cfl(0.5)
    for c, r in ((a, ra), (b, rb)):
        for x, y in zip(c.get_state(), r.get_state()):
            assert np.array_equal(x, y)
    # the forcing made a difference (the check above is not vacuous)
    assert not np.array_equal(a.get_state()[0], b.get_state()[0])
    bt.close()


def test_destroy_order_is_enforced(require_gpu):
    from littleharsh_amd import capi
    from littleharsh_amd.solver import Batch, Channel
    s = tiny()
    c = Channel(s)
    c.set_state(*laminar_plus_noise(s, seed=46))
    bt = Batch([c])
    lib = capi.load_library()
    assert lib.lh_destroy(c._h) == capi.LH_ERR_ARG     # still a member
    bt.step_cfl(0.5)                                    # the handle is intact
    bt.close()
    assert lib.lh_destroy(c._h) == capi.LH_OK
    c._h = None


def test_batch_rejects_a_repeated_handle(require_gpu):
    from littleharsh_amd.solver import Batch, Channel, LHError
    s = tiny()
    c = Channel(s)
    with pytest.raises(LHError):
        Batch([c, c])
