"""Timing of the learned-closure kernels on one GPU: the training step (batch 4096, Adam), the
closure's coupling into the 96 x 145 x 96 coarse channel (lh_closure_apply: sampling,
gradients, the network on all 1.33M coarse points, stress divergence) and the coarse step
itself.  Synthetic inputs: a seeded random coarse state and a network at its initial
weights (the kernels' cost does not depend on the values).

    python -m littleharsh_amd.closure_bench --train-steps 2000 --applies 50

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from .channel import config_production
from .closure import NIN, Filter
from .closure_les import Coarse, DeviceClosure, coarse_setup
from .initial import laminar_plus_noise
from .mlp import MLP
from .solver import Channel

BF16_PEAK = 2.5e15


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--applies", type=int, default=50)
    ap.add_argument("--batch", type=int, default=4096)
    args = ap.parse_args()
    rng = np.random.default_rng(1)
    nw = 64 * NIN + 3 * 64 * 64 + 6 * 64                  # weights of 244 -> 64 x 4 -> 6

    # training step
    net = MLP(NIN, 64, 4, 6, batch=args.batch, seed=3)
    x = rng.normal(size=(200000, NIN)).astype(np.float32)
    y = rng.normal(size=(200000, 6)).astype(np.float32)
    net.set_data(x, y)
    net.train(50, lr=1e-4, seed=1)                        # capture and warm up
    t0 = time.perf_counter()
    net.train(args.train_steps, lr=1e-4, seed=1)
    t_train = (time.perf_counter() - t0) / args.train_steps
    train_flops = 6.0 * args.batch * nw

    # coarse channel and its closure
    s = config_production()
    co = Coarse(coarse_setup(s), Filter(s))
    c = co.c
    u, v, w, p = laminar_plus_noise(c, seed=2)
    grad = co.gradients(*co.sample(u, v, w))
    gs = grad.reshape(9, -1).std(axis=1)
    xm = np.zeros(NIN, np.float32)
    xs = np.concatenate([np.tile(gs, 27), [1.0]]).astype(np.float32)
    sf = np.zeros((len(co.yc), c.nz, 2 * (c.mx - 1)), np.float32)
    ym = np.zeros(6, np.float32)
    ys = np.full(6, 1e-3, np.float32)
    ch = Channel(c)
    ch.set_state(u, v, w, p)
    dt = 0.5 * ch.cfl_dt(0.5)
    for _ in range(3):
        ch.step(dt)
    t0 = time.perf_counter()
    for _ in range(args.applies):
        ch.step(dt)
    t_step = (time.perf_counter() - t0) / args.applies

    def time_apply(cl):
        for _ in range(3):
            cl.apply()
        ch.scalars()                                      # synchronises the stream
        a = time.perf_counter()
        for _ in range(args.applies):
            cl.apply()
        ch.scalars()
        return (time.perf_counter() - a) / args.applies

    smag = DeviceClosure.smagorinsky(ch, co, 0.01)
    t_smag = time_apply(smag)
    smag.close()
    dc = DeviceClosure(ch, net, co, sf, xm, xs, ym, ys)
    t_apply = time_apply(dc)
    npt = len(co.yc) * c.nz * 2 * (c.mx - 1)
    infer_flops = 2.0 * npt * nw
    dc.close()
    ch.close()
    net.close()
    out = {
        "train": {"batch": args.batch, "ms_per_step": t_train * 1e3, "tflops": train_flops / t_train / 1e12,
                  "frac_of_bf16_peak": train_flops / t_train / BF16_PEAK},
        "coarse_step_ms": t_step * 1e3,
        "closure_apply_ms": {"learned": t_apply * 1e3, "smagorinsky": t_smag * 1e3},
        "inference": {"points": npt, "tflops_incl_coupling": infer_flops / t_apply / 1e12},
        "coarse": f"{c.nx}x{c.ny + 2 * c.d}x{c.nz} (Ngal {c.px})",
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
