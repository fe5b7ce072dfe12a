"""The learned-closure MLP on the GPU (BF16 matrix cores, FP32 accumulation) against the CPU
oracle (oracle/mlp.py with the same bfloat16 rounding of the matrix-product operands)."""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu

NIN, NHID, NLAY, NOUT = 244, 64, 4, 6


# one fixed target function for every data set drawn below
_W = np.random.default_rng(99).normal(size=(NIN, NOUT)) / np.sqrt(NIN)


def _data(rng, n):
    x = rng.normal(size=(n, NIN)).astype(np.float32)
    y = (np.tanh(x @ _W) + 0.3 * np.sin(2.0 * (x[:, :NOUT]))).astype(np.float32)
    return x, y


def test_gradient_matches_oracle(require_gpu):
    from littleharsh_amd.mlp import MLP
    rng = np.random.default_rng(11)
    net = MLP(NIN, NHID, NLAY, NOUT, batch=256, seed=3)
    x, y = _data(rng, 256)
    loss, g = net.loss_grad(x, y)
    p = net.get_params().astype(np.float64)
    shp = omlp.shapes(NIN, NHID, NLAY, NOUT)
    lo, go = omlp.loss_grad(p, x.astype(np.float64), y.astype(np.float64), shp, bf16_ops=True)
    print(f"loss gpu {loss:.6e} oracle {lo:.6e}; grad rel err "
          f"{np.linalg.norm(g - go) / np.linalg.norm(go):.3e}", flush=True)
    # FP32 accumulation against float64, and one-ulp bfloat16 flips where the two sit on a
    # rounding boundary
    assert abs(loss - lo) <= 1e-4 * lo
    assert np.linalg.norm(g - go) <= 2e-3 * np.linalg.norm(go)
    # every layer's weight and bias gradients, not just the largest ones
    o = 0
    for n, k in shp:
        for sz in (n * k, n):
            seg = slice(o, o + sz)
            assert np.linalg.norm(g[seg] - go[seg]) <= 5e-3 * np.linalg.norm(go[seg]) + 1e-9
            o += sz


def test_params_round_trip_and_predict(require_gpu):
    from littleharsh_amd.mlp import MLP
    rng = np.random.default_rng(12)
    net = MLP(NIN, NHID, NLAY, NOUT, batch=64, seed=4)
    p = (rng.normal(size=net.nparams) * 0.05).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    x, _ = _data(rng, 100)                    # not a multiple of the batch
    yp = net.predict(x)
    yo = omlp.predict(omlp.bf16(p), omlp.bf16(x), omlp.shapes(NIN, NHID, NLAY, NOUT))
    assert np.max(np.abs(yp - yo)) <= 2e-2 * np.max(np.abs(yo))


def test_training_lowers_the_loss(require_gpu):
    from littleharsh_amd.mlp import MLP
    rng = np.random.default_rng(13)
    net = MLP(NIN, NHID, NLAY, NOUT, batch=512, seed=5)
    x, y = _data(rng, 20000)
    net.set_data(x, y)
    xe, ye = _data(rng, 512)
    l0, _ = net.loss_grad(xe, ye)
    net.train(600, lr=2e-3, seed=9)
    l1, _ = net.loss_grad(xe, ye)
    print(f"held-out loss {l0:.4e} -> {l1:.4e}", flush=True)
    assert l1 < 0.3 * l0
