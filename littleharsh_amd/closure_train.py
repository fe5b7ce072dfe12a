"""Learned canopy closure end to end: GPU DNS snapshots -> filtered data set -> MLP trained on
the GPU (BF16 matrix cores) -> a-priori scores on later, held-out snapshots -> a-posteriori
runs of the coarse channel with the closure (all of its forcing, and its plane-mean part
alone), with a fitted Smagorinsky closure and without any closure, compared with the
filtered DNS.

    python -m littleharsh_amd.closure_train --spinup 1500 --snaps 30 --every 10 --steps 20000

The DNS is the 384x289x384 canopy channel of the benchmark (laminar profile plus
|k|^-5/3 modes, CFL 0.5).  After ``spinup`` steps, ``snaps`` snapshots ``every`` steps apart
train the network.  After ``gap`` further steps the check window starts: its first snapshot
is the initial state of the coarse runs, and ``check_snaps`` more follow ``check_every``
steps apart.  The check snapshots are also the a-priori test set.  The network
(244 -> 64 x 4 -> 6, tanh) predicts the six subfilter stresses from standardised features.
The a-priori baselines are no closure (tau = 0), the mean stress of the training data and
the Smagorinsky model with its constant fitted by least squares on the training data.

Prints progress lines on stderr and one JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .channel import config_production, config_small
from .closure import NIN, Filter, scores
from .closure_les import Coarse, DeviceClosure, coarse_setup, compare, plane_stats
from .initial import laminar_plus_modes
from .mlp import MLP, standardise
from .solver import Channel, LHError


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def dataset(filt, fields, per_snap, rng):
    xs, ys, bs = [], [], []
    for f in fields:
        grad = filt.gradients(f)
        x, y, idx = filt.samples(f, per_snap, rng, grad)
        xs.append(x)
        ys.append(y)
        bs.append(filt.smagorinsky_basis(f, idx, grad))
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(bs)


def run_coarse(c, co, ic, t0, checks, closure, cfl, tag, device=None, per_substep=False):
    """Advance the coarse channel from ic at time t0 to every check time, with no closure,
    a host closure (closure(u, v, w) -> forcing) or a device closure (device(ch) -> an object
    whose apply() sets the forcing on the GPU); returns the statistics errors per check time
    and the cost."""
    ch = Channel(c)
    ch.set_state(*ic)
    dev = device(ch) if device is not None else None
    t, nstep, t_cl = t0, 0, 0.0
    rows = []
    dt_first = None
    diverged = False
    w0 = time.perf_counter()
    for tj, dns, ref, qdns in checks:
        while t < tj - 1e-12 * tj:
            if closure is not None:
                a = time.perf_counter()
                u, v, w, _ = ch.get_state()
                ch.set_forcing(*closure(u, v, w))
                t_cl += time.perf_counter() - a
            if dev is not None and not per_substep:
                dev.apply()
            try:
                dcfl = ch.cfl_dt(cfl)
            except LHError:                      # CFL maximum not finite
                dcfl = 0.0
            dt_first = dt_first or dcfl
            if not dcfl > 1e-3 * dt_first:
                diverged = True
                break
            dt = min(dcfl, tj - t)
            if dev is not None and per_substep:      # the closure re-evaluated before each substep
                for krk in (1, 2, 3):
                    dev.apply()
                    ch.substep(krk, dt)
            else:
                ch.step(dt)
            t += dt
            nstep += 1
            if nstep % 25 == 0:
                log(f"# {tag}: step {nstep} t {t:.5f} dt {dt:.3e}")
        sc = ch.scalars()
        u, v, w, _ = ch.get_state()
        if diverged or not (np.isfinite(sc["qx"]) and np.all(np.isfinite(u))):
            log(f"# {tag}: diverged at t {t:.5f} after {nstep} steps (CFL time step collapsed)")
            rows.append({"t": t, "steps": nstep, "diverged": True})
            break
        ub, vb, wb = co.sample(u, v, w)
        st = plane_stats(ub, vb, wb)
        err = compare(st, ref, np.abs(co.yc) > 0.85)
        fl = dns["ub"] - ref["U"][:, None, None]
        err["u_field"] = float(np.linalg.norm(ub - dns["ub"]) / np.linalg.norm(fl))
        err["t"] = t
        err["steps"] = nstep
        # the flow rate is evaluated at the end of a step: none before the first
        err["flow_rate"] = sc["qx"] if nstep else None
        err["flow_rate_err"] = abs(sc["qx"] - qdns) / abs(qdns) if nstep else None
        rows.append(err)
        last_u = st["U"]
        log(f"# {tag}: t {t:.5f} after {nstep} steps: U {err['U']:.3e} urms {err['urms']:.3e} "
            f"uv {err['uv']:.3e} u-field {err['u_field']:.3e} flow rate {err['flow_rate_err'] or 0.0:.2e}")
    wall = time.perf_counter() - w0
    prof = [float(x) for x in last_u] if rows and not rows[-1].get("diverged") else None
    if dev is not None:
        dev.close()
    ch.close()
    span = t - t0
    return {"check": rows, "steps": nstep, "wall_s": wall, "closure_s": t_cl, "sim_time": span,
            "s_per_unit_time": wall / span if span > 0 else None, "U_last": prof}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--spinup", type=int, default=1500)
    ap.add_argument("--snaps", type=int, default=30)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--gap", type=int, default=50, help="steps between the training snapshots "
                    "and the check window")
    ap.add_argument("--check-snaps", type=int, default=5)
    ap.add_argument("--check-every", type=int, default=250)
    ap.add_argument("--per-snap", type=int, default=100000)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--les-cfl", type=float, default=0.5, help="CFL number of the coarse runs")
    ap.add_argument("--per-substep", action="store_true", help="evaluate the closures before every RK3 "
                    "substep instead of once per step")
    ap.add_argument("--config", choices=("production", "small"), default="production")
    ap.add_argument("--nxc", type=int, default=96)
    ap.add_argument("--nzc", type=int, default=96)
    args = ap.parse_args()

    s = config_production() if args.config == "production" else config_small()
    dyq = 10.0 if args.config == "production" else 6.0
    filt = Filter(s, args.nxc, args.nzc)
    ch = Channel(s)
    ch.set_state(*laminar_plus_modes(s, seed=1234))
    t0 = time.perf_counter()
    tsim = 0.0

    def advance(nsteps: int):
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

    def snapshot():
        u, v, w, _ = ch.get_state()
        return filt.filtered(u, v, w)

    train_f, train_t = [], []
    for n in range(args.snaps):
        tsim += advance(args.every)
        train_f.append(snapshot())
        train_t.append(tsim)
        log(f"# training snapshot {n + 1}/{args.snaps} at t {tsim:.5f}")
    tsim += advance(args.gap)
    check_f, check_t, check_q = [snapshot()], [tsim], [ch.scalars()["qx"]]
    dns_steps, dns_wall, dns_span = 0, 0.0, 0.0
    for n in range(args.check_snaps):
        a = time.perf_counter()
        dt = advance(args.check_every)
        dns_wall += time.perf_counter() - a
        dns_steps += args.check_every
        dns_span += dt
        tsim += dt
        check_f.append(snapshot())
        check_t.append(tsim)
        check_q.append(ch.scalars()["qx"])
        log(f"# check snapshot {n + 1}/{args.check_snaps} at t {tsim:.5f}")
    t_dns = time.perf_counter() - t0
    ch.close()

    # ---- data, training
    rng = np.random.default_rng(2024)
    xtr, ytr, btr = dataset(filt, train_f, args.per_snap, rng)
    xte, yte, bte = dataset(filt, check_f, args.per_snap, rng)
    del train_f
    log(f"# data: {len(xtr)} training and {len(xte)} test points")
    xm, xsd = xtr.mean(axis=0), xtr.std(axis=0)
    xsd[xsd < 1e-12] = 1.0
    ym, ysd = ytr.mean(axis=0), ytr.std(axis=0)
    ysd[ysd < 1e-30] = 1.0
    net = MLP(NIN, 64, 4, 6, batch=args.batch, seed=7)
    net.set_data(standardise(xtr, xm, xsd), (ytr - ym) / ysd)   # as the coupling does
    t1 = time.perf_counter()
    done, chunk = 0, 1000
    while done < args.steps:
        k = min(chunk, args.steps - done)
        loss = net.train(k, lr=args.lr, seed=11)
        done += k
        log(f"# training step {done}: batch loss {loss:.4e}")
    t_train = time.perf_counter() - t1

    pred = net.predict(standardise(xte, xm, xsd)) * ysd + ym
    dtr = (ytr - ym).ravel().astype(np.float64)
    btv = btr.ravel()
    cs2 = float(np.dot(btv, dtr) / max(np.dot(btv, btv), 1e-300))
    smag = ym + cs2 * bte
    nw = sum(a * b for a, b in [(64, NIN), (64, 64), (64, 64), (64, 64), (6, 64)])
    flops = 6.0 * args.batch * nw * args.steps          # forward 2, backward 4 per weight per row
    res = {
        "workload": f"{s.nx}x{s.ny + 2 * s.d}x{s.nz} DNS -> {args.nxc}x{len(filt.centres)}x{args.nzc} "
                    f"box-filtered ({filt.fx}x3x{filt.fz} collocation points)",
        "snapshots": {"train": args.snaps, "every": args.every, "spinup": args.spinup, "gap": args.gap,
                      "train_times": train_t, "check_times": check_t},
        "points": {"train": int(len(xtr)), "test": int(len(xte))},
        "network": {"inputs": NIN, "hidden": [64] * 4, "outputs": 6, "params": net.nparams,
                    "batch": args.batch, "steps": args.steps, "lr": args.lr, "last_batch_loss": loss},
        "train_seconds": t_train,
        "train_tflops": flops / t_train / 1e12,
        "train_frac_of_bf16_peak": flops / t_train / 2.5e15,
        "train_samples_per_s": args.batch * args.steps / t_train,
        "dns_and_filter_seconds": t_dns,
        "smagorinsky_cs2": cs2,
        "apriori": {
            "mlp": scores(pred, yte),
            "smagorinsky": scores(smag, yte),
            "mean": scores(np.broadcast_to(ym, yte.shape), yte),
            "zero": scores(np.zeros_like(yte), yte),
        },
    }
    del xtr, ytr, btr, xte, yte, bte

    # ---- a posteriori: the coarse channel from the first check snapshot
    c = coarse_setup(s, args.nxc, args.nzc, dyq)
    co = Coarse(c, filt)
    ic = co.initial_state(check_f[0])
    sf = check_f[0]["sf"].astype(np.float32)
    # the first entry (no steps) measures how well the coarse initial state represents the DNS
    checks = [(tj, f, plane_stats(f["ub"], f["vb"], f["wb"]), q) for tj, f, q in zip(check_t, check_f, check_q)]

    def learned(ch):
        return DeviceClosure(ch, net, co, sf, xm, xsd, ym, ysd)

    def learned_mean(ch):
        return DeviceClosure(ch, net, co, sf, xm, xsd, ym, ysd).set_mean_only()

    def smagorinsky(ch):
        return DeviceClosure.smagorinsky(ch, co, cs2)

    runs = {}
    for tag, dev in (("no_closure", None), ("smagorinsky", smagorinsky), ("learned", learned),
                     ("learned_mean", learned_mean)):
        runs[tag] = run_coarse(c, co, ic, check_t[0], checks, None, args.les_cfl, tag, dev, args.per_substep)
    dns_cost = dns_wall / dns_span
    for r in runs.values():
        r["speedup_vs_dns"] = dns_cost / r["s_per_unit_time"] if r["s_per_unit_time"] else None
    res["aposteriori"] = {
        "coarse": f"{c.nx}x{c.ny + 2 * c.d}x{c.nz} (Ngal {c.px}), posts {c.dstx}x{c.dstz} of "
                  f"{c.px // c.ntilex}x{c.pz // c.ntilez} points",
        "dns": {"steps": dns_steps, "wall_s": dns_wall, "sim_time": dns_span, "s_per_unit_time": dns_cost,
                "flow_rate": check_q, "U_last": [float(x) for x in checks[-1][2]["U"]],
                "yc": [float(x) for x in co.yc]},
        "runs": runs,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
