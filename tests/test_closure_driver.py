"""The coarse-run driver of littleharsh_amd/closure_train.py on the CPU: the CPU oracle stands
in for the channel (a fixed time step), and a collapsing CFL time step or a non-finite CFL
maximum is recorded as a divergence instead of ending the run."""
import numpy as np

from littleharsh_amd.channel import make_setup
from littleharsh_amd.closure import Filter
from littleharsh_amd.closure_les import Coarse, coarse_setup, plane_stats
from littleharsh_amd.closure_train import run_coarse
from littleharsh_amd.initial import laminar_plus_noise
from littleharsh_amd.solver import LHError
from oracle.oracle import Oracle


def fine_setup():
    return make_setup(nx=64, nz=64, ny_total=41, dsty=6, re=180.0, alp=1.0, bet=2.0, mpgx=-1.0,
                      flag_ctpress=1, posth=0.2, dyq=4.0, ntilex=4, ntilez=2, dstx=3, dstz=5)


class OracleChannel:
    """The oracle behind the Channel methods run_coarse uses."""

    collapse_after = None                       # steps after which cfl_dt collapses
    raise_after = None                          # steps after which cfl_dt raises

    def __init__(self, s):
        self.o, self.s, self.n = Oracle(s), s, 0

    def set_state(self, *a):
        self.o.set_state(*a)

    def get_state(self):
        return self.o.get_state()

    def cfl_dt(self, cfl):
        if self.raise_after is not None and self.n >= self.raise_after:
            raise LHError("CFL maximum is zero or not finite")
        if self.collapse_after is not None and self.n >= self.collapse_after:
            return 1e-12
        return 1e-4

    def step(self, dt):
        self.o.step(dt)
        self.n += 1

    def scalars(self):
        return dict(qx=self.o.flowrate(), mpgx=-1.0, dt=1e-4, cflmax=0.5)

    def set_forcing(self, *f):
        self.o.set_forcing(*f)

    def close(self):
        pass


class CountingClosure:
    def __init__(self, ch):
        self.ch, self.applied, self.closed = ch, 0, False

    def apply(self):
        self.applied += 1

    def close(self):
        self.closed = True


def _setup(monkeypatch):
    import littleharsh_amd.closure_train as ct
    monkeypatch.setattr(ct, "Channel", OracleChannel)
    s = fine_setup()
    filt = Filter(s, 16, 16)
    c = coarse_setup(s, 16, 16, 4.0)
    co = Coarse(c, filt)
    f = filt.filtered(*laminar_plus_noise(s, seed=2)[:3])
    ic = co.initial_state(f)
    ref = plane_stats(f["ub"], f["vb"], f["wb"])
    checks = [(1.0e-3 + 3e-4 * n, f, ref, 100.0) for n in range(3)]
    return c, co, ic, checks


def test_run_coarse_records_checks_and_applies_closure(monkeypatch):
    c, co, ic, checks = _setup(monkeypatch)
    made = []

    def device(ch):
        made.append(CountingClosure(ch))
        return made[-1]

    out = run_coarse(c, co, ic, 1.0e-3, checks, None, 0.5, "t", device)
    assert [round(r["t"], 6) for r in out["check"]] == [1.0e-3, 1.3e-3, 1.6e-3]
    assert out["steps"] == 6 and made[0].applied == 6 and made[0].closed
    assert out["check"][0]["flow_rate"] is None       # no step yet: no flow rate
    assert all(np.isfinite(r["U"]) for r in out["check"])


def test_run_coarse_records_divergence(monkeypatch):
    c, co, ic, checks = _setup(monkeypatch)
    OracleChannel.collapse_after = 4
    try:
        out = run_coarse(c, co, ic, 1.0e-3, checks, None, 0.5, "t")
    finally:
        OracleChannel.collapse_after = None
    assert out["check"][-1]["diverged"] and out["check"][-1]["steps"] == 4
    OracleChannel.raise_after = 2
    try:
        out = run_coarse(c, co, ic, 1.0e-3, checks, None, 0.5, "t")
    finally:
        OracleChannel.raise_after = None
    assert out["check"][-1]["diverged"] and out["check"][-1]["steps"] == 2
