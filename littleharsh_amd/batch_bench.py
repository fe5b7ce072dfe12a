"""The batched multi-instance workload: 48 independent channels of different sizes and
canopies (random_instance(0..47)), each autotuned, advanced together by one graph launch
per RK3 step (lh_batch_step_cfl).  Reports the aggregate grid-point substeps/s next to
stepping the instances one after another, and the FFT plans and launch parameters the
autotuner chose.

    python -m littleharsh_amd.batch_bench --instances 48 --steps 20
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from .channel import random_instance
from .initial import laminar_plus_noise
from .solver import Batch, Channel


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=48)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    setups = [random_instance(i) for i in range(args.instances)]
    chans = []
    tuned = []
    for i, s in enumerate(setups):
        c = Channel(s)
        c.set_state(*laminar_plus_noise(s, seed=1000 + i))
        tuned.append(c.get_launch() if args.no_autotune else c.autotune(2))
        chans.append(c)
    print(f"# {len(chans)} instances ready", file=sys.stderr, flush=True)
    G = sum(s.grid_points for s in setups)
    # one after another (each instance replays its own graph)
    for _ in range(args.warmup):
        for c in chans:
            c.step_cfl(args.cfl)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for c in chans:
            c.step_cfl(args.cfl)
    seq = time.perf_counter() - t0
    # all together: one graph launch per step
    b = Batch(chans)
    for _ in range(args.warmup):
        b.step_cfl(args.cfl)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.step_cfl(args.cfl)
    bat = time.perf_counter() - t0
    b.close()
    # (x length, its plan, z length, its plan) -> instances
    plan_counts = {}
    for s, c in zip(setups, chans):
        p = c.get_plans()
        key = (s.px, p["xplan"], s.pz, p["zplan"])
        plan_counts[key] = plan_counts.get(key, 0) + 1
    out = {
        "workload": f"{len(setups)} channels, nx,nz in {{32,48,64,96}}, ny in {{65,97,129}}",
        "grid_points": G, "steps": args.steps,
        "batched": {"value": G * 3 * args.steps / bat, "ms_per_step": bat / args.steps * 1e3},
        "sequential": {"value": G * 3 * args.steps / seq, "ms_per_step": seq / args.steps * 1e3},
        "unit": "grid-point substeps/s",
        "speedup": seq / bat,
        "plans": {f"{px}:{xp},{pz}:{zp}": n for (px, xp, pz, zp), n in sorted(plan_counts.items())},
        "tuned": tuned,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
