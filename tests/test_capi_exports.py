"""The HIP library loads on a machine without a GPU and exports every entry point that
include/littleharsh_amd.h declares (no compute calls are made)."""
import ctypes
import os
import re

from littleharsh_amd import capi

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                   "littleharsh_amd.h")


def declared():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lh_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_python_binding_list():
    assert declared() == sorted(capi.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_closure_apply_refuses_a_failed_create():
    """A closure whose creation failed (here: no channel) is handed back for its error
    message; applying it is refused before any device work, and destroying it is safe."""
    lib = capi.load_library()
    c = ctypes.c_void_p()
    st = lib.lh_closure_create_smagorinsky(None, None, 1.0, 1.0, 0.0, 0.0, 0.01, ctypes.byref(c))
    assert st != 0 and c.value
    assert b"null argument" in lib.lh_closure_last_error(c)
    assert lib.lh_closure_apply(c) != 0
    assert b"not created successfully" in lib.lh_closure_last_error(c)
    assert lib.lh_closure_destroy(c) == 0
