"""Closure data (littleharsh_amd/closure.py) on the CPU: the inverse plane transform, the
box filter and the subfilter stresses, the stencil features."""
import numpy as np

from cases import small
from littleharsh_amd.closure import NIN, Filter, modes_to_phys
from littleharsh_amd.initial import laminar_plus_noise, phys_to_modes


def test_modes_to_phys_inverts_phys_to_modes():
    s = small()
    rng = np.random.default_rng(3)
    a = np.zeros((1, s.ncol), complex)
    a[0] = rng.normal(size=s.ncol) + 1j * rng.normal(size=s.ncol)
    ph = modes_to_phys(s, a[0])
    back = phys_to_modes(s, ph[None])[0]
    # the modes a real field can hold: kx = 0 real; the kx = nx/2 column and kz = -nz/2 row
    # are outside what the padded grid keeps exactly, so compare the others
    b = back.reshape(s.nz, s.mx)
    ref = a[0].reshape(s.nz, s.mx)
    keep = np.ones_like(b, dtype=bool)
    keep[:, s.mx - 1] = False
    keep[s.nz // 2, :] = False
    keep[:, 0] = False
    assert np.max(np.abs(b[keep] - ref[keep])) < 1e-12 * np.max(np.abs(ref))


def test_uniform_flow_has_no_subfilter_stress():
    s = small()
    f = Filter(s, nxc=s.px // 3, nzc=s.pz // 3)
    u = np.zeros((s.ncol, s.nyu), complex)
    v = np.zeros((s.ncol, s.nyv), complex)
    w = np.zeros((s.ncol, s.nyu), complex)
    u[0, :] = 2.0
    w[0, :] = -0.5
    fl = f.filtered(u, v, w)
    assert np.allclose(fl["ub"], 2.0) and np.allclose(fl["wb"], -0.5) and np.allclose(fl["vb"], 0.0)
    for k in ("uu", "vv", "ww", "uv", "uw", "vw"):
        assert np.max(np.abs(fl["tau_" + k])) < 1e-12


def test_normal_stresses_are_non_negative_and_features_are_laid_out():
    s = small()
    u, v, w, _ = laminar_plus_noise(s, seed=5)
    f = Filter(s, nxc=s.px // 3, nzc=s.pz // 3)
    fl = f.filtered(u, v, w)
    for k in ("uu", "vv", "ww"):
        assert fl["tau_" + k].min() > -1e-10 * np.abs(fl["tau_" + k]).max()
    rng = np.random.default_rng(0)
    x, y, (m, kz, ix) = f.samples(fl, 50, rng)
    assert x.shape == (50, NIN) and y.shape == (50, 6)
    g = f.gradients(fl)
    # the centre of the 3x3x3 stencil (offset 13) holds the nine gradients of the point
    assert np.allclose(x[:, 13 * 9:14 * 9], g[:, m, kz, ix].T.astype(np.float32))
    assert np.allclose(x[:, -1], fl["sf"][m, kz, ix])
    assert (x[:, -1] >= 0).all() and (x[:, -1] <= 1).all()


def test_coarse_transforms_match_solver_convention():
    """Coarse-point samples of spectral columns agree with the solver's own (3/2-padded)
    transform at the coincident points, and phys_to_modes96 inverts modes_to_coarse."""
    from littleharsh_amd.channel import make_setup
    from littleharsh_amd.closure import modes_to_coarse, modes_to_phys, phys_to_modes96
    from littleharsh_amd.initial import laminar_plus_noise
    c = make_setup(nx=32, nz=32, ny_total=41, dsty=6, re=180.0, alp=1.0, bet=2.0, mpgx=-1.0,
                   flag_ctpress=1, posth=0.2, dyq=4.0, ntilex=4, ntilez=2, dstx=3, dstz=5)
    u = laminar_plus_noise(c, seed=3)[0][:, 7]
    b = u.reshape(c.nz, c.mx).copy()
    b[c.nz // 2, :] = 0.0
    b[:, c.mx - 1] = 0.0
    b = b.ravel()
    fine = modes_to_phys(c, b)                  # (48, 48) dealiased grid
    coarse = modes_to_coarse(b, c.nz, c.mx)     # (32, 32) one point per mode
    assert coarse.shape == (32, 32)
    assert np.allclose(fine[::3, ::3], coarse[::2, ::2], atol=1e-12 * np.abs(fine).max())
    assert np.allclose(phys_to_modes96(coarse, c.nz, c.mx), b, atol=1e-12 * np.abs(b).max())


def test_standardise_is_the_float32_reciprocal_form():
    """The network's input standardisation (the form the device gathers compute): float32
    (x - mean) * (1 / scale), the reciprocal rounded to float32.  It agrees with the division
    to within one rounding and differs from it in some elements, so a switch back to
    dividing (on the host while the device multiplies) would be caught."""
    from littleharsh_amd.mlp import standardise
    rng = np.random.default_rng(21)
    x = rng.normal(size=(4000, 8)).astype(np.float32)
    mean = rng.normal(size=8).astype(np.float32)
    scale = (0.3 + rng.random(8)).astype(np.float32)
    out = standardise(x, mean, scale)
    assert out.dtype == np.float32
    r = (np.float32(1.0) / scale).astype(np.float32)
    assert np.array_equal(out, (x - mean) * r)
    div = (x - mean) / scale
    assert np.any(out != div)
    assert np.max(np.abs(out - div) / np.maximum(np.abs(div), 1e-30)) < 3e-7
