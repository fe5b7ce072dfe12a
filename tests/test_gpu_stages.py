"""The stage-level hooks of the substep (lh_nonlinear, lh_solve_u, lh_solve_p) against the
oracle's matching parts of the reference substep (littleharsh.f90:166-188): the nonlinear
term, then RHS0 + nonlinear + solveU/V/U, then meanflow_ctU + solveP + v_corr, each from the
same state and the same Nu_old; and the three hooks in turn against lh_substep, which runs
the same kernels in the same order (bit-identical).  FP64, 1e-10 relative."""
import numpy as np
import pytest

from littleharsh_amd.initial import laminar_plus_noise
from oracle.oracle import Oracle

from cases import tiny

pytestmark = pytest.mark.gpu
TOL = 1e-10


def rel(a, b):
    d = np.abs(b).max()
    return np.abs(a - b).max() / (d if d > 0 else 1.0)


def nu_old(s, seed):
    rng = np.random.default_rng(seed)
    c = lambda n: 50.0 * (rng.standard_normal((s.ncol, n)) + 1j * rng.standard_normal((s.ncol, n)))
    return c(s.nyu - 2), c(s.nyv - 2), c(s.nyu - 2)


@pytest.mark.parametrize("flag", [1, 0])
@pytest.mark.parametrize("krk", [1, 2, 3])
def test_stages_match_oracle(require_gpu, krk, flag):
    from littleharsh_amd.solver import Channel
    s = tiny(flag_ctpress=flag)
    st = laminar_plus_noise(s, seed=21)
    nuo = nu_old(s, 5)
    dt = 2e-4
    g, o = Channel(s), Oracle(s)
    for x in (g, o):
        x.set_state(*st)
        x.set_nonlinear(*nuo)
    # the nonlinear stage alone
    on = Oracle(s)
    on.set_state(*st)
    on.nonlinear(krk)
    g.nonlinear(krk)
    e_nl = [rel(a, b) for a, b in zip(g.get_nonlinear(), on.get_nonlinear())]
    # the velocity stage
    g.solve_u(krk, dt)
    o.substep_vel(krk, dt)
    e_u = [rel(a, b) for a, b in zip(g.get_state(), o.get_state())]
    # the pressure stage
    g.solve_p(krk, dt)
    o.substep_pres(krk, dt)
    e_p = [rel(a, b) for a, b in zip(g.get_state(), o.get_state())]
    mg, mo = g.scalars()["mpgx"], o.scalars()["mpgx"]
    print(f"kRK {krk} flag {flag}: Nu {max(e_nl):.2e} | u v w after solve_u "
          f"{' '.join(f'{e:.2e}' for e in e_u[:3])} | u v w p after solve_p "
          f"{' '.join(f'{e:.2e}' for e in e_p)} | mpgx {mg:.15e} {mo:.15e}", flush=True)
    assert max(e_nl) < TOL
    assert max(e_u[:3]) < TOL
    assert max(e_p) < TOL
    # ctU adds (QxT - Qx0)/Qx1 to mpgx; Qx0 sums ny-1 terms in an order that differs between
    # the two, so the difference is at most (ny-1) eps QxT/|Qx1| with |Qx1| = dRK dt span (the
    # bound of the three-step constant-flow-rate parity test, for one substep)
    eps = np.finfo(float).eps
    span = s.grid.yv[s.ny - 1 + s.d] - s.grid.yv[s.d]
    dRK = s.rk.ravel()[12 + krk - 1]
    bound = 3 * (s.ny - 1) * eps * s.qxt / (dRK * dt * span) if flag == 0 else 0.0
    assert abs(mg - mo) <= bound


@pytest.mark.parametrize("krk", [1, 2, 3])
def test_stages_compose_to_the_substep(require_gpu, krk):
    from littleharsh_amd.solver import Channel
    s = tiny(flag_ctpress=0)
    st = laminar_plus_noise(s, seed=22)
    nuo = nu_old(s, 6)
    a, b = Channel(s), Channel(s)
    for x in (a, b):
        x.set_state(*st)
        x.set_nonlinear(*nuo)
    a.substep(krk, 3e-4)
    b.nonlinear(krk)
    b.solve_u(krk, 3e-4)
    b.solve_p(krk, 3e-4)
    for x, y in zip(a.get_state(), b.get_state()):
        assert np.array_equal(x, y)
    for x, y in zip(a.get_nonlinear(), b.get_nonlinear()):
        assert np.array_equal(x, y)
