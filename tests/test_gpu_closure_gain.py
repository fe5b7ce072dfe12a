"""What the trained closure gains over the method without it, on the small channel (64x97x64,
Re_tau 180, 8x4 posts): GPU DNS snapshots box-filtered to 32x48x32, the network trained on the
GPU, scored on later snapshots it has not seen.  Without a closure the subfilter stresses are
zero (relative error 1); the network must do clearly better than that, than the mean stress
of the training data and than a Smagorinsky model with a least-squares constant."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_trained_closure_beats_no_closure(require_gpu):
    from littleharsh_amd.channel import config_small
    from littleharsh_amd.closure import NIN, Filter, scores
    from littleharsh_amd.initial import laminar_plus_modes
    from littleharsh_amd.mlp import MLP
    from littleharsh_amd.solver import Channel
    s = config_small()
    filt = Filter(s, 32, 32)
    ch = Channel(s)
    ch.set_state(*laminar_plus_modes(s, seed=1234))
    for _ in range(300):
        ch.step_cfl(0.5)
    fields = []
    for n in range(10):
        for _ in range(10 if n != 8 else 40):          # a gap before the two test snapshots
            ch.step_cfl(0.5)
        fields.append(filt.filtered(*ch.get_state()[:3]))
    ch.close()
    rng = np.random.default_rng(5)

    def data(fs):
        xs, ys, bs = [], [], []
        for f in fs:
            g = filt.gradients(f)
            x, y, idx = filt.samples(f, None, rng, g)
            xs.append(x)
            ys.append(y)
            bs.append(filt.smagorinsky_basis(f, idx, g))
        return np.concatenate(xs), np.concatenate(ys), np.concatenate(bs)

    xtr, ytr, btr = data(fields[:8])
    xte, yte, bte = data(fields[8:])
    xm, xs = xtr.mean(axis=0), xtr.std(axis=0)
    xs[xs < 1e-12] = 1.0
    ym, ys = ytr.mean(axis=0), ytr.std(axis=0)
    ys[ys < 1e-30] = 1.0
    net = MLP(NIN, 64, 4, 6, batch=1024, seed=7)
    net.set_data((xtr - xm) / xs, (ytr - ym) / ys)
    net.train(3000, lr=1e-3, seed=3)
    pred = net.predict((xte - xm) / xs) * ys + ym
    dtr = (ytr - ym).ravel().astype(np.float64)
    cs2 = float(np.dot(btr.ravel(), dtr) / np.dot(btr.ravel(), btr.ravel()))
    sc = {"mlp": scores(pred, yte), "smagorinsky": scores(ym + cs2 * bte, yte),
          "mean": scores(np.broadcast_to(ym, yte.shape), yte), "zero": scores(np.zeros_like(yte), yte)}
    mean_err = {k: float(np.mean([v["rel_err"] for v in sc[k].values()])) for k in sc}
    mean_cor = {k: float(np.mean([v["corr"] for v in sc[k].values()])) for k in sc}
    print("mean relative error", {k: round(v, 3) for k, v in mean_err.items()},
          "mean correlation", {k: round(v, 3) for k, v in mean_cor.items()}, flush=True)
    assert mean_err["zero"] == 1.0
    assert mean_err["mlp"] < 0.85
    assert mean_err["mlp"] < mean_err["mean"] - 0.1
    assert mean_err["mlp"] < mean_err["smagorinsky"] - 0.1
    assert mean_cor["mlp"] > 0.5
