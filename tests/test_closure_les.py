"""The coarse channel's closure plumbing (littleharsh_amd/closure_les.py) on the CPU: the
stress divergence of analytic stress fields, the initial-state / sampling round trip and the
Smagorinsky stresses on the whole grid."""
import numpy as np
import pytest

from littleharsh_amd.channel import config_production
from littleharsh_amd.closure import Filter, modes_to_coarse
from littleharsh_amd.closure_les import Coarse, coarse_setup, smagorinsky_tau


@pytest.fixture(scope="module")
def co():
    s = config_production()
    c = coarse_setup(s)
    assert (c.nx, c.ny + 2 * c.d + 1, c.nz, c.px, c.dstx, c.dstz) == (96, 146, 96, 144, 2, 3)
    return Coarse(c, Filter(s))


def _points(co):
    """Coarse points x_i, z_k and the box centres the network sees."""
    f = co.filt
    x = np.arange(f.nxc) * f.dxc
    z = np.arange(f.nzc) * f.dzc
    dx = 0.5 * (f.fx - 1) * (2 * np.pi / f.s.alp) / f.s.px
    dz = 0.5 * (f.fz - 1) * (2 * np.pi / f.s.bet) / f.s.pz
    return x, z, x + dx, z + dz


def test_forcing_of_analytic_stresses(co):
    x, z, xs, zs = _points(co)
    y = co.yc
    ny, nz, nx = len(y), len(z), len(x)
    tau = np.zeros((ny, nz, nx, 6))
    tau[..., 0] = np.sin(3.0 * xs)[None, None, :]                 # uu(x)
    tau[..., 1] = (y ** 2)[:, None, None]                         # vv(y)
    tau[..., 3] = (y ** 2)[:, None, None]                         # uv(y)
    tau[..., 5] = np.cos(2.0 * zs)[None, :, None]                 # vw(z), beta = 2: one period
    f1, f2, f3 = co.forcing(tau)
    assert f1.shape == (co.c.ncol, co.c.nyu - 2) and f2.shape == (co.c.ncol, co.c.nyv - 2)
    F1 = modes_to_coarse(f1.T, co.nz, co.mx)
    F2 = modes_to_coarse(f2.T, co.nz, co.mx)
    F3 = modes_to_coarse(f3.T, co.nz, co.mx)
    # f1 = d/dx uu + d/dy uv at the points x_i (central differences of y^2: y+ + y-)
    dy_uv = np.empty(ny)
    dy_uv[1:-1] = y[2:] + y[:-2]
    dy_uv[0] = y[1] + y[0]
    dy_uv[-1] = y[-1] + y[-2]
    e1 = 3.0 * np.cos(3.0 * x)[None, None, :] + dy_uv[:, None, None]
    assert np.abs(F1 - e1).max() < 1e-9 * np.abs(e1).max()
    # f2 at the faces = d/dz vw averaged to the face + d/dy vv (exact for y^2)
    e2 = -2.0 * np.sin(2.0 * z)[None, :, None] + (y[1:] + y[:-1])[:, None, None]
    assert np.abs(F2 - e2).max() < 1e-9 * np.abs(e2).max()
    # f3 = d/dx uw + d/dz ww + d/dy vw: vw has no y dependence
    assert np.abs(F3).max() < 1e-9


def test_initial_state_sample_round_trip(co):
    x, z, xs, zs = _points(co)
    ny = len(co.yc)
    base = (np.cos(2.0 * xs)[None, :] * np.sin(4.0 * zs)[:, None] + 0.3 * np.sin(5.0 * xs)[None, :])
    f = {k: np.broadcast_to(a * base + b, (ny,) + base.shape).copy()
         for k, a, b in (("ub", 2.0, 50.0), ("vb", 0.5, 0.0), ("wb", -1.0, 0.0))}
    u, v, w, p = co.initial_state(f)
    ub, vb, wb = co.sample(u, v, w)
    mid = np.abs(co.yc) < 0.8                  # away from the walls and canopies
    for a, k in ((ub, "ub"), (vb, "vb"), (wb, "wb")):
        assert np.abs(a[mid] - f[k][mid]).max() < 1e-9 * np.abs(f[k]).max()


def test_smagorinsky_grid_matches_sample_basis(co):
    rng = np.random.default_rng(4)
    filt = co.filt
    ny = len(filt.yc)
    f = {k: rng.normal(size=(ny, filt.nzc, filt.nxc)) for k in ("ub", "vb", "wb")}
    grad = filt.gradients(f)
    t = smagorinsky_tau(grad, filt.yc, filt.dxc, filt.dzc, 1.0)
    idx = (rng.integers(0, ny, 50), rng.integers(0, filt.nzc, 50), rng.integers(0, filt.nxc, 50))
    b = filt.smagorinsky_basis(f, idx)
    assert np.allclose(t[idx], b, rtol=1e-12, atol=1e-12 * np.abs(b).max())
