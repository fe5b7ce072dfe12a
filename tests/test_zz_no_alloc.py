"""lh_max_div_noalloc (the divergence monitor that allocates nothing) against lh_max_div
through the C-ABI: two calls in a row on two channels, the second with fewer spectral
columns than the first and a divergence a thousand times smaller, so that a reduction
reaching past the entries the second call wrote would return the first call's maximum.
A NaN in the state comes back as NaN, as from lh_max_div."""
import ctypes

import numpy as np
import pytest

from littleharsh_amd import capi

from cases import small, tiny

pytestmark = pytest.mark.gpu


def random_state(s, seed, scale):
    rng = np.random.default_rng(seed)
    cr = lambda n: scale * (rng.standard_normal((s.ncol, n)) + 1j * rng.standard_normal((s.ncol, n)))
    return cr(s.nyu), cr(s.nyv), cr(s.nyu), cr(s.nyp)


def noalloc_max_div(lib, ch, out):
    import torch
    out.fill_(-1.0)
    torch.cuda.synchronize()
    r = lib.lh_max_div_noalloc(ch._h, ctypes.c_void_p(out.data_ptr()))
    assert r == capi.LH_OK, lib.lh_last_error(ch._h).decode()
    torch.cuda.synchronize()             # the call returns without waiting for its kernels
    return float(out.cpu()[0])


def test_noalloc_max_div_matches_max_div(require_gpu):
    import torch
    from littleharsh_amd.solver import Channel
    lib = capi.load_library()
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    got = []
    for s, seed, scale in ((small(), 5, 1e3), (tiny(), 6, 1.0)):
        print(f"case nx={s.nx} nz={s.nz} nyv={s.nyv} ncol={s.ncol}", flush=True)
        ch = Channel(s)
        ch.set_state(*random_state(s, seed, scale))
        ref = ch.max_div()
        val = noalloc_max_div(lib, ch, out)
        print(f"  lh_max_div {ref:.17e}  lh_max_div_noalloc {val:.17e}", flush=True)
        assert ref > 0.0
        assert abs(val - ref) <= 1e-13 * ref
        got.append(val)
        ch.close()
    assert got[0] > 100.0 * got[1]


def test_noalloc_max_div_keeps_nan(require_gpu):
    import torch
    from littleharsh_amd.solver import Channel
    lib = capi.load_library()
    s = tiny()
    u, v, w, p = random_state(s, 7, 1.0)
    u[3, 5] = np.nan
    ch = Channel(s)
    ch.set_state(u, v, w, p)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    assert np.isnan(ch.max_div())
    assert np.isnan(noalloc_max_div(lib, ch, out))
    ch.close()
