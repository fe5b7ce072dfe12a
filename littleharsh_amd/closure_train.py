"""Learned canopy closure end to end: GPU DNS snapshots -> filtered data set -> MLP trained on
the GPU (BF16 matrix cores) -> a-priori scores on later, held-out snapshots.

    python -m littleharsh_amd.closure_train --spinup 400 --snaps 10 --every 10 --steps 20000

The DNS is the 384x289x384 canopy channel of the benchmark (laminar profile plus
|k|^-5/3 modes, CFL 0.5).  Snapshots are taken every ``every`` steps after ``spinup``
steps; the first ``train_frac`` of them train the network, the rest, taken after ``gap``
further steps, test it.  The network
(244 -> 64 x 4 -> 6, tanh) predicts the six subfilter stresses from standardised features;
the baselines are no closure (tau = 0), the mean stress of the training data and the
Smagorinsky model with its constant fitted by least squares on the training data.

Prints progress lines and one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .channel import config_production, config_small
from .closure import NIN, Filter, scores
from .initial import laminar_plus_modes
from .mlp import MLP
from .solver import Channel


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--spinup", type=int, default=400)
    ap.add_argument("--snaps", type=int, default=10)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--train-frac", type=float, default=0.8)
    ap.add_argument("--gap", type=int, default=100, help="extra steps between the training and "
                    "the test snapshots")
    ap.add_argument("--per-snap", type=int, default=150000)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--config", choices=("production", "small"), default="production")
    ap.add_argument("--nxc", type=int, default=96)
    ap.add_argument("--nzc", type=int, default=96)
    args = ap.parse_args()

    s = config_production() if args.config == "production" else config_small()
    filt = Filter(s, args.nxc, args.nzc)
    ch = Channel(s)
    ch.set_state(*laminar_plus_modes(s, seed=1234))
    t0 = time.perf_counter()
    tsim = 0.0

    def advance(nsteps: int) -> float:
        t = 0.0
        for _ in range(nsteps):
            t += ch.step_cfl(args.cfl)
        sc = ch.scalars()
        if not all(np.isfinite(list(sc.values()))):
            raise RuntimeError(f"DNS state not finite: {sc}")
        return t

    for k in range(0, args.spinup, 100):
        tsim += advance(min(100, args.spinup - k))
        sc = ch.scalars()
        log(f"# spin-up step {min(k + 100, args.spinup)}: t {tsim:.5f} dt {sc['dt']:.3e} "
            f"flow rate {sc['qx']:.4f} div {ch.max_div():.2e}")
    ntr = max(1, int(round(args.train_frac * args.snaps)))
    fields, times = [], []
    for n in range(args.snaps):
        if n == ntr:
            tsim += advance(args.gap)
        tsim += advance(args.every)
        u, v, w, _ = ch.get_state()
        fields.append(filt.filtered(u, v, w))
        times.append(tsim)
        log(f"# snapshot {n + 1}/{args.snaps} at t {tsim:.5f} filtered")
    t_dns = time.perf_counter() - t0
    del ch

    rng = np.random.default_rng(2024)
    xs, ys, bs = [], [], []
    for f in fields[:ntr]:
        x, y, idx = filt.samples(f, args.per_snap, rng)
        xs.append(x)
        ys.append(y)
        bs.append(filt.smagorinsky_basis(f, idx))
    xtr, ytr, btr = np.concatenate(xs), np.concatenate(ys), np.concatenate(bs)
    xs, ys, bs = [], [], []
    for f in fields[ntr:]:
        x, y, idx = filt.samples(f, args.per_snap, rng)
        xs.append(x)
        ys.append(y)
        bs.append(filt.smagorinsky_basis(f, idx))
    xte, yte, bte = np.concatenate(xs), np.concatenate(ys), np.concatenate(bs)
    log(f"# data: {len(xtr)} training and {len(xte)} test points")

    # standardisation (training statistics)
    xm, xsd = xtr.mean(axis=0), xtr.std(axis=0)
    xsd[xsd < 1e-12] = 1.0
    ym, ysd = ytr.mean(axis=0), ytr.std(axis=0)
    ysd[ysd < 1e-30] = 1.0
    net = MLP(NIN, 64, 4, 6, batch=args.batch, seed=7)
    net.set_data((xtr - xm) / xsd, (ytr - ym) / ysd)
    t1 = time.perf_counter()
    done, chunk = 0, 1000
    while done < args.steps:
        k = min(chunk, args.steps - done)
        loss = net.train(k, lr=args.lr, seed=11)
        done += k
        log(f"# training step {done}: batch loss {loss:.4e}")
    t_train = time.perf_counter() - t1

    pred = net.predict((xte - xm) / xsd) * ysd + ym
    # Smagorinsky: tau = mean + c * (-2 Delta^2 |S| S), c = C_s^2 by least squares
    dtr = (ytr - ym).ravel().astype(np.float64)
    btv = btr.ravel()
    c = float(np.dot(btv, dtr) / max(np.dot(btv, btv), 1e-300))
    smag = ym + c * bte
    res = {
        "workload": f"{s.nx}x{s.ny}x{s.nz} DNS -> {args.nxc}x{len(filt.centres)}x{args.nzc} "
                    f"box-filtered ({filt.fx}x3x{filt.fz} collocation points)",
        "snapshots": {"train": ntr, "test": args.snaps - ntr, "every": args.every, "spinup": args.spinup,
                      "gap": args.gap, "times": times},
        "points": {"train": int(len(xtr)), "test": int(len(xte))},
        "network": {"inputs": NIN, "hidden": [64] * 4, "outputs": 6, "params": net.nparams,
                    "batch": args.batch, "steps": args.steps, "lr": args.lr},
        "train_seconds": t_train,
        "dns_and_filter_seconds": t_dns,
        "smagorinsky_cs2": c,
        "scores": {
            "mlp": scores(pred, yte),
            "smagorinsky": scores(smag, yte),
            "mean": scores(np.broadcast_to(ym, yte.shape), yte),
            "zero": scores(np.zeros_like(yte), yte),
        },
    }
    nw = sum(a * b for a, b in [(64, NIN), (64, 64), (64, 64), (64, 64), (6, 64)])
    flops = 6.0 * args.batch * nw * args.steps          # forward 2, backward 4 per weight per row
    res["train_tflops"] = flops / t_train / 1e12
    res["train_samples_per_s"] = args.batch * args.steps / t_train
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
