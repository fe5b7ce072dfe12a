"""Pins of the oracle parts that no reference source covers (the nonlinear term and the
plane transforms of the absent FOU3D.f90, and the immersed-boundary masking step whose
transforms live there), by independent numpy implementations and closed forms."""
import numpy as np
import pytest

from littleharsh_amd.channel import make_setup
from oracle.oracle import Oracle


def tiny(nx=16, nz=8, ib=True):
    s = make_setup(nx=nx, nz=nz, ny_total=19, dsty=3, re=180.0, posth=0.2, dyq=3.0,
                   ntilex=4, ntilez=2, dstx=2, dstz=3)
    if not ib:
        s.list_u = s.list_u[:, :0]
        s.list_v = s.list_v[:, :0]
    return s


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ---------------- independent numpy restatement with a different padding -----------------
def to_phys(s, col, P):
    """spectral (nz, mx) -> physical (Pz, Px): z inverse, x c2r, no normalisation."""
    Pz, Px = P
    h = s.nz // 2
    a = np.zeros((Pz, s.mx), complex)
    a[:h] = col[:h]
    a[Pz - h:] = col[h:]
    a = np.fft.ifft(a, axis=0) * Pz
    b = np.zeros((Pz, Px // 2 + 1), complex)
    b[:, :s.mx] = a
    return np.fft.irfft(b, n=Px, axis=1) * Px


def to_spec(s, ph, P):
    Pz, Px = P
    a = np.fft.rfft(ph, axis=1)[:, :s.mx] / Px
    a = np.fft.fft(a, axis=0) / Pz
    h = s.nz // 2
    return np.concatenate([a[:h], a[Pz - h:]], axis=0)


def numpy_nonlinear(s, u, v, w, P):
    """Nu1, Nu2, Nu3 of the divergence-form restatement with padding P = (Pz, Px)."""
    d, ny = s.d, s.ny
    plane = lambda f, r: f[:, r].reshape(s.nz, s.mx)
    U = {r: to_phys(s, plane(u, r), P) for r in range(1, s.nyu - 1)}
    U[s.nyu - 1] = to_phys(s, plane(u, s.nyu - 1), P)
    W = {r: to_phys(s, plane(w, r), P) for r in range(1, s.nyu)}
    V = {r: to_phys(s, plane(v, r), P) for r in range(0, s.nyv)}
    kx = np.tile(s.k1x, s.nz)
    kz = np.repeat(s.k1z, s.mx)
    n1 = np.zeros_like(u)
    n3 = np.zeros_like(u)
    n2 = np.zeros_like(v)
    sp = lambda ph: to_spec(s, ph, P).ravel()
    UV, VW, VV = {}, {}, {}
    for r in range(1, s.nyu - 1):            # centres (u slots) and faces (v slots) r
        vc = 0.5 * (V[r - 1] + V[r])
        VV[r] = sp(vc * vc)
        UV[r] = sp(0.5 * (U[r] + U[r + 1]) * V[r]) if r < s.nyv - 1 else 0 * VV[r]
        VW[r] = sp(V[r] * 0.5 * (W[r] + W[r + 1])) if r < s.nyv - 1 else 0 * VV[r]
    UV[0] = VW[0] = 0 * VV[1]
    for r in range(1, s.nyu - 1):
        m = s.grid.ddthetavi * s.grid.dthdyu[r]
        uu, uw, ww = sp(U[r] * U[r]), sp(U[r] * W[r]), sp(W[r] * W[r])
        n1[:, r] = 1j * kx * uu + 1j * kz * uw + m * (UV[r] - UV[r - 1])
        n3[:, r] = 1j * kx * uw + 1j * kz * ww + m * (VW[r] - VW[r - 1])
    for r in range(1, s.nyv - 1):
        m = s.grid.ddthetavi * s.grid.dthdyv[r]
        n2[:, r] = 1j * kx * UV[r] + 1j * kz * VW[r] + m * (VV[r + 1] - VV[r])
    return n1, n2, n3


def random_state(s, seed=3):
    rng = np.random.default_rng(seed)
    cr = lambda n: rng.standard_normal((s.ncol, n)) + 1j * rng.standard_normal((s.ncol, n))
    u, v, w, p = cr(s.nyu), cr(s.nyv), cr(s.nyu), cr(s.nyp)
    v[:, 0] = v[:, -1] = 0.0
    return u, v, w, p


# (nx, nz): wider in x, wider in z (the z line is then longer than the x line) and square
SHAPES = [(16, 8), (8, 16), (16, 16)]


@pytest.mark.parametrize("nx,nz", SHAPES)
def test_plane_transforms_match_numpy(nx, nz):
    """Both plane transforms of the oracle against numpy's, for every aspect of the plane."""
    s = tiny(nx=nx, nz=nz)
    o = Oracle(s)
    rng = np.random.default_rng(nx + 7 * nz)
    sp = rng.standard_normal((s.nz, s.mx)) + 1j * rng.standard_normal((s.nz, s.mx))
    sp[:, 0] = sp[:, 0].real
    sp[s.nz // 2:, 0] = np.conj(sp[1:s.nz // 2 + 1, 0][::-1])    # kx = 0 column Hermitian in z
    sp[s.nz // 2, 0] = sp[s.nz // 2, 0].real
    ph = o.four_to_phys(sp.ravel())
    assert rel(ph, to_phys(s, sp, (s.pz, s.px))) < 1e-13
    ph = rng.standard_normal((s.pz, s.px))
    assert rel(o.phys_to_four(ph).reshape(s.nz, s.mx), to_spec(s, ph, (s.pz, s.px))) < 1e-13


@pytest.mark.parametrize("nx,nz", SHAPES)
def test_nonlinear_matches_numpy_same_padding(nx, nz):
    """Same 3/2 padding, independent implementation (numpy FFTs, dict-based planes)."""
    s = tiny(nx=nx, nz=nz)
    u, v, w, p = random_state(s)
    o = Oracle(s)
    o.set_state(u, v, w, p)
    n1, n2, n3 = o.nonlinear()
    r1, r2, r3 = numpy_nonlinear(s, u, v, w, (s.pz, s.px))
    assert rel(n1[:, 1:-1], r1[:, 1:-1]) < 1e-12
    assert rel(n2[:, 1:-1], r2[:, 1:-1]) < 1e-12
    assert rel(n3[:, 1:-1], r3[:, 1:-1]) < 1e-12


@pytest.mark.parametrize("P_factor", [2.0, 2.5])
def test_nonlinear_alias_free_except_unpaired_modes(P_factor):
    """The retained modes include the unpaired kx = N/2 (start.f90:1033-1035) and kz = -N/2
    modes.  Under 3/2 padding the product of two kx = N/2 modes (kx = N) folds onto
    kx = -N/2, i.e. into the retained i = N/2 column, and the product of two (0, +N/2)
    modes (the conjugates of the retained (0, -N/2)) folds onto (0, -N/2).  Every other
    retained mode is alias-free: a wider padding gives the same values there."""
    s = tiny()
    u, v, w, p = random_state(s)
    o = Oracle(s)
    o.set_state(u, v, w, p)
    n1, n2, n3 = o.nonlinear()
    r1, r2, r3 = numpy_nonlinear(s, u, v, w, (int(P_factor * s.nz), int(P_factor * s.nx)))
    i = np.arange(s.ncol) % s.mx
    k = np.arange(s.ncol) // s.mx
    keep = (i < s.mx - 1) & ~((i == 0) & (k == s.nz // 2))
    for a, b in ((n1, r1), (n2, r2), (n3, r3)):
        assert rel(a[keep, 1:-1], b[keep, 1:-1]) < 1e-12
    assert rel(n1[~keep, 1:-1], r1[~keep, 1:-1]) > 1e-6    # the aliased modes differ


def test_nonlinear_single_mode_closed_form():
    """u = 2 Re[a e^{i alpha x}] g(y), v = w = 0: Nu1 = d/dx(u^2) has only the kx = 2 mode,
    i 2 alpha a^2 g^2, and Nu2 = Nu3 = 0."""
    s = tiny()
    rng = np.random.default_rng(5)
    g = rng.standard_normal(s.nyu)
    a = 0.3 - 0.7j
    u = np.zeros((s.ncol, s.nyu), complex)
    u[1] = a * g                                   # column (i=1, k=0)
    z = np.zeros((s.ncol, s.nyv), complex)
    o = Oracle(s)
    o.set_state(u, z, np.zeros_like(u), np.zeros((s.ncol, s.nyp), complex))
    n1, n2, n3 = o.nonlinear()
    expect = np.zeros_like(n1)
    expect[2] = 1j * 2 * s.alp * a * a * g * g    # column (i=2, k=0)
    assert np.abs(n1[:, 1:-1] - expect[:, 1:-1]).max() < 1e-12 * np.abs(expect).max()
    assert np.abs(n2).max() == 0.0 or np.abs(n2).max() < 1e-14
    assert np.abs(n3).max() < 1e-14


def test_nonlinear_vanishes_for_a_mean_profile():
    s = tiny()
    u = np.zeros((s.ncol, s.nyu), complex)
    u[0] = np.cos(np.arange(s.nyu))                 # (0,0) mode only
    z = np.zeros((s.ncol, s.nyv), complex)
    o = Oracle(s)
    o.set_state(u, z, np.zeros_like(u), np.zeros((s.ncol, s.nyp), complex))
    n1, n2, n3 = o.nonlinear()
    # u(y) has no x or z variation and v = 0: every term is a derivative of a constant or
    # a product with v, so only FFT round-off of u^2 = O(1) remains
    for n in (n1, n2, n3):
        assert np.abs(n).max() < 1e-14


@pytest.mark.parametrize("grid", ["u", "v"])
def test_immersed_boundary_masking_matches_numpy(grid):
    """immersed_boundaries_U / _V (tridLU_3D_IB.f90:224-245, 304-325): on canopy planes
    rhs -> physical, 0 at listed points, -beta*Lu at listed points of the tip planes, back to
    Fourier.  The tip planes follow the reference's tests literally: j == 0 or
    j == Ngal(4,3)-dsty+1 on the u grid (line 237), j == -1 or j == Ngal(3,3)-dsty+2 on the v
    grid (line 317), with Ngal(4,3) = ny-1+dsty and Ngal(3,3) = ny-2+dsty (y_grid_canopy.f90:
    46-52)."""
    s = tiny()
    nj = s.nyu if grid == "u" else s.nyv
    lst = s.list_u if grid == "u" else s.list_v
    rng = np.random.default_rng(9 if grid == "u" else 19)
    cr = lambda: rng.standard_normal((s.ncol, nj)) + 1j * rng.standard_normal((s.ncol, nj))
    x, lu = cr(), cr()
    beta = 1.7e-3
    o = Oracle(s)
    out = o.ib(grid == "u", beta, x, lu)
    if grid == "u":
        tips = (0, (s.ny - 1 + s.d) - s.d + 1)
    else:
        tips = (-1, (s.ny - 2 + s.d) - s.d + 2)
    P = (s.pz, s.px)
    ref = x.copy()
    ntip = 0
    for r in range(nj):
        j = r - s.d
        pts = lst[:, lst[2] == j]
        if pts.shape[1] == 0:
            continue
        ph = to_phys(s, x[:, r].reshape(s.nz, s.mx), P)
        lph = to_phys(s, lu[:, r].reshape(s.nz, s.mx), P)
        tip = j in tips
        ntip += tip
        for i, k in zip(pts[0] - 1, pts[1] - 1):
            ph[k, i] = -beta * lph[k, i] if tip else 0.0
        ref[:, r] = to_spec(s, ph, P).ravel()
    assert ntip == 2                      # both tip planes hold canopy points
    assert rel(out, ref) < 1e-13
    untouched = [r for r in range(nj) if (r - s.d) not in set(lst[2].tolist())]
    assert np.array_equal(out[:, untouched], x[:, untouched])


def test_velocity_lu_reconstructs_matrix():
    """LU_dec storage (tridLU_3D_IB.f90:601-640): a1 = lower of L, a2 = 1/diag of L,
    a3 = upper of U (unit diagonal); L U must give back LU_build's matrix."""
    s = tiny()
    o = Oracle(s)
    dt, krk = 3e-4, 2
    beta = (1.0 / 15.0) * dt / s.re
    ay, _ = o.lu_vel(True, krk, dt)
    n = s.nyu
    A = np.zeros((n, n))
    A[0, 0], A[0, 1] = 1.0, s.gw[0]
    for r in range(1, n - 1):
        w = s.grid.dyu2i[r]
        A[r, r - 1:r + 2] = [-beta * w[0], 1 - beta * w[1], -beta * w[2]]
    A[n - 1, n - 2], A[n - 1, n - 1] = s.gw[1], 1.0
    L = np.diag(1.0 / ay[:, 1]) + np.diag(ay[1:, 0], -1)
    U = np.eye(n) + np.diag(ay[:-1, 2], 1)
    assert np.abs(L @ U - A).max() < 1e-13 * np.abs(A).max()


def test_step_is_divergence_free():
    s = tiny()
    from littleharsh_amd.initial import laminar_plus_noise
    o = Oracle(s)
    o.set_state(*laminar_plus_noise(s, seed=4))
    o.step(2e-4)
    assert o.max_div() < 1e-10


@pytest.mark.parametrize("nx,nz", SHAPES)
def test_cfl_maximum_matches_numpy(nx, nz):
    """max over centres and collocation points of |u|/dx + |v_c| dthdyu + |w|/dz, v_c the
    average of the faces below and above (the CFL restatement; dtc/dti are declared at
    declaration.f90:51 but set by the absent nonlinear routine)."""
    s = tiny(nx=nx, nz=nz)
    u, v, w, p = random_state(s, seed=12)
    o = Oracle(s)
    o.set_state(u, v, w, p)
    o.nonlinear(1)
    P = (s.pz, s.px)
    dxi = s.px * s.alp / (2 * np.pi)
    dzi = s.pz * s.bet / (2 * np.pi)
    m = 0.0
    for r in range(1, s.nyu - 1):
        U = to_phys(s, u[:, r].reshape(s.nz, s.mx), P)
        W = to_phys(s, w[:, r].reshape(s.nz, s.mx), P)
        V = 0.5 * (to_phys(s, v[:, r - 1].reshape(s.nz, s.mx), P)
                   + to_phys(s, v[:, r].reshape(s.nz, s.mx), P))
        m = max(m, float((np.abs(U) * dxi + np.abs(V) * s.grid.dthdyu[r] + np.abs(W) * dzi).max()))
    assert abs(o.scalars()["cflmax"] - m) < 1e-12 * m
