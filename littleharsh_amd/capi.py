"""ctypes mirror of ``include/littleharsh_amd.h`` and the loader of the HIP library.

The product path is the shared library ``littleharsh_amd/liblh_amd.so`` built from
``littleharsh_amd/csrc`` for gfx950.  Loading fails loudly when it is missing: there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .channel import ChannelSetup

LH_OK = 0
LH_ERR_ARG = 1
LH_ERR_HIP = 2
LH_ERR_NUMERIC = 3
LH_PREC_FP64 = 0
LH_PREC_MIXED = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class LHConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("nz", ctypes.c_int), ("px", ctypes.c_int), ("pz", ctypes.c_int),
        ("ny", ctypes.c_int), ("dsty", ctypes.c_int),
        ("re", ctypes.c_double), ("alp", ctypes.c_double), ("bet", ctypes.c_double),
        ("flag_ctpress", ctypes.c_int),
        ("mpgx", ctypes.c_double), ("qxt", ctypes.c_double), ("ddthetavi", ctypes.c_double),
        ("yu", _dp), ("yv", _dp), ("dthdyu", _dp), ("dthdyv", _dp), ("dyu2i", _dp),
        ("dyv2i", _dp),
        ("gridweighting", _dp), ("rk", _dp),
        ("k1x", _dp), ("k2x", _dp), ("k1z", _dp), ("k2z", _dp),
        ("nlist_u", ctypes.c_int), ("nlist_v", ctypes.c_int),
        ("list_u", _ip), ("list_v", _ip),
        ("tip_u", ctypes.c_int * 2), ("tip_v", ctypes.c_int * 2),
        ("ntilex", ctypes.c_int), ("ntilez", ctypes.c_int),
        ("band_v", ctypes.c_int * 2),
        ("precision", ctypes.c_int),
    ]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def make_config(s: ChannelSetup, precision: int = LH_PREC_FP64):
    """Return (LHConfig, keepalive) for a set-up.  ``keepalive`` holds the arrays the
    struct points into; keep it referenced until lh_create returns."""
    g = s.grid
    keep = dict(
        yu=np.ascontiguousarray(g.yu, dtype=np.float64),
        yv=np.ascontiguousarray(g.yv, dtype=np.float64),
        dthdyu=np.ascontiguousarray(g.dthdyu, dtype=np.float64),
        dthdyv=np.ascontiguousarray(g.dthdyv, dtype=np.float64),
        dyu2i=np.ascontiguousarray(g.dyu2i, dtype=np.float64),     # (nyu, 3) == dyu2i(3, j)
        dyv2i=np.ascontiguousarray(g.dyv2i, dtype=np.float64),
        gw=np.ascontiguousarray(s.gw, dtype=np.float64),
        rk=np.ascontiguousarray(s.rk.ravel(), dtype=np.float64),
        k1x=np.ascontiguousarray(s.k1x), k2x=np.ascontiguousarray(s.k2x),
        k1z=np.ascontiguousarray(s.k1z), k2z=np.ascontiguousarray(s.k2z),
        lu=np.ascontiguousarray(s.list_u.T, dtype=np.int32),          # list_ib(3, n)
        lv=np.ascontiguousarray(s.list_v.T, dtype=np.int32),
    )
    c = LHConfig()
    c.nx, c.nz, c.px, c.pz, c.ny, c.dsty = s.nx, s.nz, s.px, s.pz, s.ny, s.d
    c.re, c.alp, c.bet = s.re, s.alp, s.bet
    c.flag_ctpress, c.mpgx, c.qxt, c.ddthetavi = s.flag_ctpress, s.mpgx, s.qxt, g.ddthetavi
    c.yu, c.yv = _dptr(keep["yu"]), _dptr(keep["yv"])
    c.dthdyu, c.dthdyv = _dptr(keep["dthdyu"]), _dptr(keep["dthdyv"])
    c.dyu2i, c.dyv2i = _dptr(keep["dyu2i"]), _dptr(keep["dyv2i"])
    c.gridweighting, c.rk = _dptr(keep["gw"]), _dptr(keep["rk"])
    c.k1x, c.k2x, c.k1z, c.k2z = (_dptr(keep[k]) for k in ("k1x", "k2x", "k1z", "k2z"))
    c.nlist_u, c.nlist_v = s.list_u.shape[1], s.list_v.shape[1]
    c.list_u = keep["lu"].ctypes.data_as(_ip)
    c.list_v = keep["lv"].ctypes.data_as(_ip)
    c.tip_u[0], c.tip_u[1] = s.tips["u"]
    c.tip_v[0], c.tip_v[1] = s.tips["v"]
    c.ntilex, c.ntilez = s.ntilex, s.ntilez
    c.band_v[0], c.band_v[1] = g.band_v
    c.precision = precision
    return c, keep


_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblh_amd.so")
_lib = None


def load_library():
    """Load the HIP library; raises if it has not been built (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library {LIB_PATH} is missing: run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.lh_create.argtypes = [ctypes.POINTER(LHConfig), ctypes.POINTER(vp)]
    lib.lh_destroy.argtypes = [vp]
    lib.lh_last_error.argtypes = [vp]
    lib.lh_last_error.restype = ctypes.c_char_p
    lib.lh_set_state.argtypes = [vp, _dp, _dp, _dp, _dp]
    lib.lh_get_state.argtypes = [vp, _dp, _dp, _dp, _dp]
    lib.lh_set_nonlinear.argtypes = [vp, _dp, _dp, _dp]
    lib.lh_set_forcing.argtypes = [vp, _dp, _dp, _dp]
    lib.lh_get_forcing.argtypes = [vp, _dp, _dp, _dp]
    lib.lh_get_nonlinear.argtypes = [vp, _dp, _dp, _dp]
    lib.lh_substep.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    lib.lh_nonlinear.argtypes = [vp, ctypes.c_int]
    lib.lh_solve_u.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    lib.lh_solve_p.argtypes = [vp, ctypes.c_int, ctypes.c_double]
    lib.lh_step.argtypes = [vp, ctypes.c_double]
    lib.lh_step_cfl.argtypes = [vp, ctypes.c_double, _dp]
    lib.lh_cfl_dt.argtypes = [vp, ctypes.c_double, _dp]
    lib.lh_get_scalars.argtypes = [vp, _dp, _dp, _dp, _dp]
    lib.lh_max_div.argtypes = [vp, _dp]
    lib.lh_max_div_noalloc.argtypes = [vp, vp]
    lib.lh_stats_accumulate.argtypes = [vp]
    lib.lh_stats_read.argtypes = [vp, _dp, _ip, ctypes.c_int]
    lib.lh_cond_stats_size.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    lib.lh_cond_stats_read.argtypes = [vp, _dp, ctypes.c_longlong, ctypes.c_int]
    lib.lh_spectra_accumulate.argtypes = [vp]
    lib.lh_spectra_read.argtypes = [vp, _dp, ctypes.c_int]
    lib.lh_set_launch.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.lh_get_launch.argtypes = [vp, _ip, _ip, _ip]
    lib.lh_set_plans.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.lh_get_plans.argtypes = [vp, _ip, _ip, _ip, _ip]
    lib.lh_autotune.argtypes = [vp, ctypes.c_int]
    lib.lh_batch_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp)]
    lib.lh_batch_step_cfl.argtypes = [vp, ctypes.c_double]
    lib.lh_batch_last_error.argtypes = [vp]
    lib.lh_batch_last_error.restype = ctypes.c_char_p
    lib.lh_batch_destroy.argtypes = [vp]
    lib.lh_timing_enable.argtypes = [vp, ctypes.c_int]
    lib.lh_timing_reset.argtypes = [vp]
    lib.lh_timing_read.argtypes = [vp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_longlong), _dp]
    _fp = ctypes.POINTER(ctypes.c_float)
    lib.lh_mlp_create.argtypes = [ctypes.c_int] * 5 + [ctypes.c_ulonglong, ctypes.POINTER(vp)]
    lib.lh_mlp_destroy.argtypes = [vp]
    lib.lh_mlp_last_error.argtypes = [vp]
    lib.lh_mlp_last_error.restype = ctypes.c_char_p
    lib.lh_mlp_param_count.argtypes = [vp]
    lib.lh_mlp_param_count.restype = ctypes.c_longlong
    lib.lh_mlp_get_params.argtypes = [vp, _fp]
    lib.lh_mlp_set_params.argtypes = [vp, _fp]
    lib.lh_mlp_set_data.argtypes = [vp, _fp, _fp, ctypes.c_longlong]
    lib.lh_mlp_loss_grad.argtypes = [vp, _fp, _fp, _fp, _fp]
    lib.lh_mlp_train.argtypes = [vp, ctypes.c_int, ctypes.c_float, ctypes.c_ulonglong, _fp]
    lib.lh_mlp_set_fused.argtypes = [vp, ctypes.c_int]
    lib.lh_mlp_predict.argtypes = [vp, _fp, ctypes.c_longlong, _fp]
    lib.lh_mlp_predict_grid.argtypes = [vp, _fp, ctypes.c_int, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        _fp, _fp, _fp]
    lib.lh_closure_create.argtypes = [vp, vp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, _fp, _fp, _fp, _fp, _fp, ctypes.POINTER(vp)]
    lib.lh_closure_create_smagorinsky.argtypes = [vp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_double, ctypes.POINTER(vp)]
    lib.lh_closure_apply.argtypes = [vp]
    lib.lh_closure_destroy.argtypes = [vp]
    lib.lh_closure_set_mean_only.argtypes = [vp, ctypes.c_int]
    lib.lh_closure_last_error.argtypes = [vp]
    lib.lh_closure_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


EXPORTED = [
    "lh_create", "lh_destroy", "lh_last_error", "lh_set_state", "lh_get_state",
    "lh_set_nonlinear", "lh_get_nonlinear", "lh_set_forcing", "lh_get_forcing", "lh_substep", "lh_nonlinear", "lh_solve_u",
    "lh_solve_p", "lh_step", "lh_step_cfl",
    "lh_cfl_dt", "lh_get_scalars", "lh_max_div", "lh_max_div_noalloc", "lh_stats_accumulate", "lh_stats_read",
    "lh_cond_stats_size", "lh_cond_stats_read", "lh_spectra_accumulate", "lh_spectra_read",
    "lh_timing_enable", "lh_timing_reset", "lh_timing_read",
    "lh_set_launch", "lh_get_launch", "lh_set_plans", "lh_get_plans", "lh_autotune",
    "lh_batch_create", "lh_batch_step_cfl", "lh_batch_last_error", "lh_batch_destroy",
    "lh_mlp_create", "lh_mlp_destroy", "lh_mlp_last_error", "lh_mlp_param_count",
    "lh_mlp_get_params", "lh_mlp_set_params", "lh_mlp_set_data", "lh_mlp_loss_grad",
    "lh_mlp_train", "lh_mlp_set_fused", "lh_mlp_predict", "lh_mlp_predict_grid", "lh_closure_create", "lh_closure_create_smagorinsky", "lh_closure_apply",
    "lh_closure_set_mean_only", "lh_closure_destroy", "lh_closure_last_error",
]

KERNEL_NAMES = ["zinv", "xpass_nl", "zfwd", "assemble", "ib_z", "ib_x", "vel", "pressure", "misc",
                "rhs", "step"]
