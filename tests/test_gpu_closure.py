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


# (hidden width, batch): the products step through 32 rows or columns at a time; width 48
# and batch 272 (odd multiples of 16) also take the last 16-wide step of every product
@pytest.mark.parametrize("nhid,batch", [(NHID, 256), (48, 272)])
def test_gradient_matches_oracle(require_gpu, nhid, batch):
    from littleharsh_amd.mlp import MLP
    rng = np.random.default_rng(11)
    net = MLP(NIN, nhid, NLAY, NOUT, batch=batch, seed=3)
    x, y = _data(rng, batch)
    loss, g = net.loss_grad(x, y)
    p = net.get_params().astype(np.float64)
    shp = omlp.shapes(NIN, nhid, NLAY, NOUT)
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


def _mix64(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _batch_rows(seed, step, B, n):
    """Rows the device draws for training step `step` (SplitMix64 finaliser, counter based)."""
    r = np.arange(B, dtype=np.uint64)
    inner = _mix64(np.uint64(step) * np.uint64(0x100000001B3) + r)
    return (_mix64(np.uint64(seed) ^ inner) % np.uint64(n)).astype(np.int64)


def test_train_steps_match_oracle_adam(require_gpu):
    """Adam steps of the captured training graph (batch sampling, step counter, bias
    correction, re-capture on a new learning rate) against the oracle's gradients and a
    float64 Adam."""
    from littleharsh_amd.mlp import MLP
    rng = np.random.default_rng(15)
    B, n = 64, 1000
    net = MLP(NIN, NHID, NLAY, NOUT, batch=B, seed=6)
    p = net.get_params().astype(np.float64)
    net.set_params(p.astype(np.float32))              # step counter and moments reset
    x, y = _data(rng, n)
    net.set_data(x, y)
    shp = omlp.shapes(NIN, NHID, NLAY, NOUT)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    b1, b2, eps = 0.9, 0.999, 1e-8
    for step, lr in enumerate((1e-3, 1e-3, 5e-4)):
        rows = _batch_rows(21, step, B, n)
        _, g = omlp.loss_grad(p, x[rows].astype(np.float64), y[rows].astype(np.float64), shp,
                              bf16_ops=True)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        t = step + 1
        dp = -lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
        net.train(1, lr=lr, seed=21)
        pg = net.get_params().astype(np.float64)
        e = np.linalg.norm((pg - p) - dp) / np.linalg.norm(dp)
        print(f"Adam step {t}: update rel err {e:.3e}", flush=True)
        # sign(g) decides each element's first update: elements whose gradient sits at the
        # rounding level may flip, so the bound is on the norm
        assert e < 0.05
        p = p + dp


def test_fused_training_is_bit_identical_to_the_per_layer_kernels(require_gpu):
    """The closure's shape trains by a fused pass (forward, loss terms and input gradients per
    16 batch rows; every layer's weight gradients in one launch).  The per-layer kernels give
    the same loss, gradients and parameters bit for bit: batch 272 (17 row tiles), 40 steps
    (both captured graphs: 2 x 16 steps and 8 single steps)."""
    from littleharsh_amd.mlp import MLP
    rng = np.random.default_rng(18)
    x, y = _data(rng, 3000)
    out = []
    for fused in (True, False):
        net = MLP(NIN, NHID, NLAY, NOUT, batch=272, seed=10)
        net.set_fused(fused)
        loss, g = net.loss_grad(x[:272], y[:272])
        net.set_data(x, y)
        last = net.train(40, lr=1e-3, seed=4)
        out.append((loss, g, last, net.get_params()))
        net.close()
    (la, ga, ta, pa), (lb, gb, tb, pb) = out
    print(f"fused / per-layer: loss {la:.6e} {lb:.6e}, last batch loss {ta:.6e} {tb:.6e}", flush=True)
    assert la == lb and np.array_equal(ga, gb)
    assert ta == tb and np.array_equal(pa, pb)


@pytest.mark.parametrize("ny,nz,nx", [(5, 6, 7), (4, 5, 48)])
def test_predict_grid_matches_stencil_rows(require_gpu, ny, nz, nx):
    """The device-side stencil gather gives the inputs the training rows were built from.
    nx = 7 runs the general fused kernel; nx = 48 (a multiple of 16) the one with the closure
    network's widths compiled in, whose tiles have an interior and a boundary path."""
    from littleharsh_amd.closure import stencil_features
    from littleharsh_amd.mlp import MLP, standardise
    rng = np.random.default_rng(16)
    net = MLP(NIN, NHID, NLAY, NOUT, batch=64, seed=8)
    g = rng.normal(size=(9, ny, nz, nx)).astype(np.float32)
    sf = rng.random((ny, nz, nx)).astype(np.float32)
    mean = (0.1 * rng.normal(size=NIN)).astype(np.float32)
    scale = (1.0 + rng.random(NIN)).astype(np.float32)
    m, kz, ix = (a.ravel() for a in np.meshgrid(np.arange(ny), np.arange(nz), np.arange(nx), indexing="ij"))
    ref = net.predict(standardise(stencil_features(g, sf, m, kz, ix), mean, scale))
    out = net.predict_grid(g, sf, mean, scale).reshape(-1, NOUT)
    same = np.mean(np.all(out == ref, axis=1))
    print(f"rows identical: {same:.3f}; max diff {np.abs(out - ref).max():.3e}", flush=True)
    assert same == 1.0                          # the same operands, k order and instructions
    assert np.abs(out - ref).max() <= 2e-2 * np.abs(ref).max()


def test_predict_grid_unfused_shape(require_gpu):
    """A network wider than the fused kernel takes (80 hidden) runs the pack + layer-by-layer
    path and gives the same rows."""
    from littleharsh_amd.closure import stencil_features
    from littleharsh_amd.mlp import MLP
    rng = np.random.default_rng(17)
    net = MLP(NIN, 80, 2, NOUT, batch=64, seed=9)
    ny, nz, nx = 4, 5, 6
    g = rng.normal(size=(9, ny, nz, nx)).astype(np.float32)
    sf = rng.random((ny, nz, nx)).astype(np.float32)
    mean = np.zeros(NIN, np.float32)
    scale = np.ones(NIN, np.float32)
    m, kz, ix = (a.ravel() for a in np.meshgrid(np.arange(ny), np.arange(nz), np.arange(nx), indexing="ij"))
    ref = net.predict(stencil_features(g, sf, m, kz, ix))
    out = net.predict_grid(g, sf, mean, scale).reshape(-1, NOUT)
    assert np.mean(np.all(out == ref, axis=1)) > 0.95
    assert np.abs(out - ref).max() <= 2e-2 * np.abs(ref).max()
