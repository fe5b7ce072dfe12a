"""Benchmark: grid-point RK3 substeps/s of the FP64 channel DNS on one MI355X.

Workload: Re_tau = 550 channel 2pi x 2 x pi, Nx = Nz = 384 (Ngal 576), 289 v planes,
48 x 24 canopy posts per wall, laminar profile + random |k|^-5/3 modes (seed 1234),
CFL 0.5 time step chosen on the device.  Inputs are resident in HBM before the timed
region; the timed region is K full RK3 steps (3 substeps each) after W warm-up steps, each
step one replay of the captured graph.  `roofline` is the whole substep against the
compulsory bytes B_alg = (21 + 9 f_IB) S per substep (S one complex FP64 spectral field),
timed by HIP events around the graph replays on the solver's stream; `roofline_kernel` is
the dominant kernel on its own inputs and outputs, from a separate kernel-by-kernel pass.

Prints one JSON line (rank 0).  With --gpus N > 1 it exits: the solver runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-point RK3 substeps/s at 384x289x384 FP64; HBM roofline fraction"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(setup, state, dt, threads):
    """The CPU oracle (a C restatement of the reference algorithm, OpenMP) timed on one
    full RK3 step of the same workload on this host's cores."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle.oracle import Oracle
    o = Oracle(setup)
    o.set_state(*state)
    t0 = time.perf_counter()
    o.step(dt)
    t1 = time.perf_counter()
    del o
    return setup.grid_points * 3 / (t1 - t0), t1 - t0


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)
    except Exception:
        return None


def calibrate_copy(nbytes=1 << 30, reps=3):
    """Device-to-device copies of a known size through the HIP runtime's blit kernel: the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    for _ in range(reps):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3):   # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibration": "device copy", "bytes_read_per_copy": nbytes,
                      "bytes_written_per_copy": nbytes, "copies": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--workload", choices=["production", "small"], default="production")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 1 GiB device copies (counter calibration for profiles)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: the solver runs on exactly one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    from littleharsh_amd import capi
    from littleharsh_amd.channel import config_production, config_small
    from littleharsh_amd.initial import laminar_plus_modes
    from littleharsh_amd.solver import Channel

    s = config_production() if args.workload == "production" else config_small()
    prec = capi.LH_PREC_FP64 if args.precision == "fp64" else capi.LH_PREC_MIXED
    state = laminar_plus_modes(s, seed=1234)
    ch = Channel(s, precision=prec)
    ch.set_state(*state)
    print(f"# set-up done: {s.nx}x{s.nyv}x{s.nz}, ncol={s.ncol}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        ch.step_cfl(args.cfl)
    # timed region: K CFL steps replayed from the captured graph, with one HIP event pair
    # around each replay on the solver's stream
    ch.timing(2)
    ch.timing_reset()
    t0 = time.perf_counter()
    dts = []
    for _ in range(args.steps):
        dts.append(ch.step_cfl(args.cfl))   # synchronous at return
    t1 = time.perf_counter()
    step_ev = ch.timing_read()["step"]
    ch.timing(0)
    wall = t1 - t0
    G = s.grid_points
    substeps = 3 * args.steps
    value = G * substeps / wall
    # per-kernel split, outside the timed region: two steps issued kernel by kernel with an
    # event pair around every launch
    ch.timing(1)
    ch.timing_reset()
    for _ in range(2):
        ch.step_cfl(args.cfl)
    tk = ch.timing_read()
    ch.timing(0)
    # roofline of the whole substep: compulsory bytes B_alg = (21 + 9 f_IB) S per substep
    # (nonlinear stage 13 S, column stage 8 S, canopy stage 9 S on the canopy planes), S one
    # complex FP64 spectral field, f_IB the fraction of planes holding canopy points
    S = s.ncol * s.nyv * 16.0
    ib_planes = len(set(s.list_u[2].tolist())) if s.list_u.shape[1] else 0
    f_ib = ib_planes / s.nyu
    b_alg = (21.0 + 9.0 * f_ib) * S
    ev_substep_s = step_ev["ms"] / 1e3 / max(step_ev["launches"], 1) / 3.0
    ach = b_alg / ev_substep_s / 1e9
    tr = load_traffic(f"traffic_{args.workload}_{args.precision}.json")
    roof = {"bound": "hbm", "kernel": "RK3 substep (captured graph)", "achieved": round(ach, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
            "traffic": (tr or {}).get("bytes_per_substep"),
            "bytes_per_launch": b_alg, "f_ib": round(f_ib, 4), "avg_launch_ms": ev_substep_s * 1e3}
    # the dominant kernel on its own algorithmic bytes (inputs read once, outputs written once)
    kern = {k: v for k, v in tk.items() if k != "step"}
    dname, dk = max(kern.items(), key=lambda kv: kv[1]["ms"])
    avg_s = dk["ms"] / 1e3 / max(dk["launches"], 1)
    kach = dk["bytes_per_launch"] / avg_s / 1e9
    ktr = (tr or {}).get("kernels", {}).get(dname)
    out = {
        "metric": METRIC, "value": value, "unit": "grid-point substeps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"channel Re_tau={s.re:g} {s.nx}x{s.nyv}x{s.nz} canopy "
                               f"{s.ntilex}x{s.ntilez} posts, CFL {args.cfl}",
                   "nx": s.nx, "ny": s.nyv, "nz": s.nz, "ngal": [s.px, s.pz],
                   "precision": args.precision, "dt_last": dts[-1]},
        "roofline": roof,
        "roofline_kernel": {"bound": "hbm", "kernel": dname, "achieved": round(kach, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(kach / HBM_PEAK_GBS, 4),
                            "bytes_per_launch": dk["bytes_per_launch"], "avg_launch_ms": avg_s * 1e3,
                            "traffic": ktr.get("bytes_per_launch") if ktr else None},
        "kernels_ms_per_step": {k: round(v["ms"] / 2, 3) for k, v in kern.items()},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        del ch
        threads = min(len(os.sched_getaffinity(0)), 16)
        print(f"# cpu baseline: oracle, one RK3 step on {threads} threads", file=sys.stderr, flush=True)
        v, secs = cpu_baseline(s, state, dts[-1], threads)
        out["cpu_baseline"] = {"value": v, "unit": "grid-point substeps/s", "cores": threads,
                               "kind": "port",
                               "sample": f"one full RK3 step (3 substeps) of the same workload, "
                                         f"{secs:.1f} s"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
