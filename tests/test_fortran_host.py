"""The Fortran host (littleharsh_amd/fortran/lh_demo.f90, through the lh_amd iso_c_binding
module) drives the same C-ABI as the Python host and gets the same answers."""
import os
import subprocess

import numpy as np
import pytest

from littleharsh_amd.fortran_io import read_result, write_case
from littleharsh_amd.initial import laminar_plus_noise

from cases import tiny
from conftest import gpu_available

DEMO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "littleharsh_amd", "fortran", "lh_demo")


def test_fortran_host_reports_errors_like_the_reference(tmp_path):
    """Without a device lh_create fails; the host prints lh_last_error and stops with a
    non-zero code (the reference's write + stop convention)."""
    if gpu_available():
        pytest.skip("a GPU is visible: the failure path is exercised on CPU-only hosts")
    s = tiny()
    case = tmp_path / "case.bin"
    write_case(str(case), s, laminar_plus_noise(s, seed=3), 1, 2e-4)
    r = subprocess.run([DEMO, str(case), str(tmp_path / "res.bin")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode != 0
    assert "lh error in lh_create" in r.stdout


@pytest.mark.gpu
def test_fortran_host_matches_python_host(require_gpu, tmp_path):
    from littleharsh_amd.solver import Channel
    s = tiny()
    st = laminar_plus_noise(s, seed=3)
    dt, nsteps = 2e-4, 2
    case, res = tmp_path / "case.bin", tmp_path / "res.bin"
    write_case(str(case), s, st, nsteps, dt)
    r = subprocess.run([DEMO, str(case), str(res)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    fstate, fdiv, fmpgx = read_result(str(res), s)
    g = Channel(s)
    g.set_state(*st)
    for _ in range(nsteps):
        for k in (1, 2, 3):
            g.substep(k, dt)
    for a, b in zip(fstate, g.get_state()):
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()
    assert fdiv < 1e-9
