"""Mixed precision against FP64 on the same workload: run both from one initial state, take
statistics and spectra every ``every`` steps, and report the errors of the mean profile
U(y), the rms profiles u', v', w', the shear stress uv and the 1D spectra.

    python -m littleharsh_amd.precision_report --steps 1000 --every 10

Prints one JSON line.  Limits (from the mixed-precision target): mean <= 0.5 %, rms and
the other profiles <= 2 %.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import capi
from .channel import config_production, config_small
from .initial import laminar_plus_modes
from .solver import Channel


def profiles(s, st, n):
    """Mean and rms profiles from the per-plane sums (rows in write_stats order)."""
    npts = float(s.px * s.pz * n)
    U, U2 = st[0] / npts, st[1] / npts
    V, V2 = st[2, :s.nyv] / npts, st[3, :s.nyv] / npts
    W, W2 = st[4] / npts, st[5] / npts
    UV = st[8] / npts
    out = {
        "U": U,
        "urms": np.sqrt(np.maximum(U2 - U * U, 0.0)),
        "vrms": np.sqrt(np.maximum(V2 - V * V, 0.0)),
        "wrms": np.sqrt(np.maximum(W2 - W * W, 0.0)),
    }
    Vc = np.zeros_like(U)
    Vc[1:-1] = 0.5 * (V[:-1] + V[1:])
    out["uv"] = UV - U * Vc
    return out


def spectra_1d(s, sp, n):
    """Streamwise and spanwise 1D spectra of u, v, w summed over y (from sp(j, column))."""
    mx, nz = s.mx, s.nz
    res = {}
    for f, name in ((0, "u"), (1, "v"), (2, "w")):
        a = sp[f].reshape(nz, mx, s.nyu).sum(axis=2) / n
        res[f"Ex_{name}"] = a.sum(axis=0)
        res[f"Ez_{name}"] = a.sum(axis=1)
    return res


def rel_err(a, b, floor=1e-3):
    """max |a - b| / max |b| over the entries that carry at least `floor` of the maximum."""
    sc = np.abs(b).max()
    if sc == 0:
        return 0.0
    m = np.abs(b) >= floor * sc
    return float(np.abs(a - b)[m].max() / sc)


def run(s, precision, state, steps, every, cfl):
    ch = Channel(s, precision=precision)
    ch.set_state(*state)
    t0 = time.perf_counter()
    for k in range(steps):
        ch.step_cfl(cfl)
        if (k + 1) % every == 0:
            ch.stats_accumulate()
            ch.spectra_accumulate()
    wall = time.perf_counter() - t0
    st, n = ch.stats()
    sp = ch.spectra()
    return st, n, sp, wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--workload", choices=["production", "small"], default="production")
    args = ap.parse_args()
    s = config_production() if args.workload == "production" else config_small()
    state = laminar_plus_modes(s, seed=1234)
    res = {}
    for prec, tag in ((capi.LH_PREC_FP64, "fp64"), (capi.LH_PREC_MIXED, "mixed")):
        print(f"# {tag}: {args.steps} steps", file=sys.stderr, flush=True)
        res[tag] = run(s, prec, state, args.steps, args.every, args.cfl)
    (st64, n64, sp64, w64), (st32, n32, sp32, w32) = res["fp64"], res["mixed"]
    p64, p32 = profiles(s, st64, n64), profiles(s, st32, n32)
    e64, e32 = spectra_1d(s, sp64, n64), spectra_1d(s, sp32, n32)
    out = {
        "workload": f"{s.nx}x{s.nyv}x{s.nz}", "steps": args.steps, "samples": n64,
        "ms_per_step": {"fp64": w64 / args.steps * 1e3, "mixed": w32 / args.steps * 1e3},
        "speedup": w64 / w32,
        "errors": {k: rel_err(p32[k], p64[k]) for k in p64},
        "spectra_errors": {k: rel_err(e32[k], e64[k]) for k in e64},
        "limits": {"U": 0.005, "other": 0.02},
    }
    errs = list(out["errors"].values()) + list(out["spectra_errors"].values())
    out["within_limits"] = out["errors"]["U"] <= 0.005 and max(errs) <= 0.02
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
