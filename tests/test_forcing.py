"""The forcing hook (lh_set_forcing): a term added to the nonlinear term Nu_new of every
substep, through which the learned closure's stress divergence enters the coarse channel.

The oracle's forcing is pinned by an identity of the RK3 update
du = C1 Lu - cRK Nu_old - gRK (Nu_new + F): in substep 2 a forced run with Nu_old = A equals
an unforced run with Nu_old = A + (gRK/cRK) F.  A wrong sign, plane offset or field mapping
of F breaks it.  The GPU is then held to the forced oracle at the FP64 bar.
"""
import numpy as np
import pytest

from littleharsh_amd.channel import rk3_tables
from littleharsh_amd.initial import laminar_plus_noise
from oracle.oracle import Oracle

from cases import tiny

TOL = 1e-10


def rel(a, b):
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)


def forcing_fields(s, seed, scale):
    """Interior planes of a random real velocity field (kx = 0 plane conjugate symmetric)."""
    u, v, w, _ = laminar_plus_noise(s, seed=seed)
    return scale * u[:, 1:-1], scale * v[:, 1:-1], scale * w[:, 1:-1]


def nu_old(s, seed):
    """A plausible Nu_old: the nonlinear term of another random state."""
    o = Oracle(s)
    o.set_state(*laminar_plus_noise(s, seed=seed))
    n1, n2, n3 = o.nonlinear(1)
    return n1[:, 1:-1], n2[:, 1:-1], n3[:, 1:-1]


def test_oracle_forcing_identity():
    s = tiny()
    st = laminar_plus_noise(s, seed=5)
    A = nu_old(s, 6)
    F = forcing_fields(s, 8, 30.0)
    rk = rk3_tables()
    g2, c2 = rk[2, 1], rk[3, 1]
    dt = 2e-4
    o1 = Oracle(s)
    o1.set_state(*st)
    o1.set_nonlinear(*A)
    o1.set_forcing(*F)
    o1.substep(2, dt)
    o2 = Oracle(s)
    o2.set_state(*st)
    o2.set_nonlinear(*[a + (g2 / c2) * f for a, f in zip(A, F)])
    o2.substep(2, dt)
    o3 = Oracle(s)
    o3.set_state(*st)
    o3.set_nonlinear(*A)
    o3.substep(2, dt)
    x1, x2, x3 = o1.get_state(), o2.get_state(), o3.get_state()
    for a, b, c in zip(x1, x2, x3):
        assert rel(a, b) < 1e-12
        assert rel(a, c) > 1e-6          # the forcing changes the result
    # Nu_new stored for the next substep carries the forcing
    n1 = o1.get_nonlinear()
    n3 = o3.get_nonlinear()
    for a, b, f in zip(n1, n3, F):
        assert rel(a - b, f) < 1e-12


def test_oracle_forcing_removed():
    s = tiny()
    st = laminar_plus_noise(s, seed=5)
    o1, o2 = Oracle(s), Oracle(s)
    o1.set_state(*st)
    o2.set_state(*st)
    o1.set_forcing(*forcing_fields(s, 8, 30.0))
    o1.set_forcing()
    o1.step(2e-4)
    o2.step(2e-4)
    for a, b in zip(o1.get_state(), o2.get_state()):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_gpu_forcing_parity(require_gpu):
    from littleharsh_amd.solver import Channel
    s = tiny()
    st = laminar_plus_noise(s, seed=5)
    F = forcing_fields(s, 8, 30.0)
    g, o = Channel(s), Oracle(s)
    g.set_state(*st)
    o.set_state(*st)
    g.set_forcing(*F)
    o.set_forcing(*F)
    dt = 2e-4
    for k in (1, 2, 3):
        g.substep(k, dt)
        o.substep(k, dt)
        errs = [rel(a, b) for a, b in zip(g.get_state(), o.get_state())]
        print(f"forced substep {k}", " ".join(f"{e:.3e}" for e in errs), flush=True)
        assert max(errs) < TOL
    # captured steps (the graph gains the add), a forcing without its v part, then none
    for n in range(2):
        g.step(dt)
        o.step(dt)
    F2 = (F[0] * 0.5, None, F[2])
    g.set_forcing(*F2)
    o.set_forcing(*F2)
    g.step(dt)
    o.step(dt)
    g.set_forcing()
    o.set_forcing()
    g.step(dt)
    o.step(dt)
    errs = [rel(a, b) for a, b in zip(g.get_state(), o.get_state())]
    print("forced steps", " ".join(f"{e:.3e}" for e in errs), flush=True)
    assert max(errs) < TOL
