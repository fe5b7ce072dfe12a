"""CPU checks of the closure oracle (oracle/mlp.py): exact backpropagation against central
finite differences, and the bfloat16 rounding against hand-worked values."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is halfway between bfloat16 neighbours 1 and 1 + 2^-7: ties go to even (1)
    assert mlp.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties go to even (1 + 2^-6)
    assert mlp.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    # above the halfway point rounds up
    assert mlp.bf16(np.array([1.0 + 2.0 ** -8 + 2.0 ** -12]))[0] == 1.0 + 2.0 ** -7
    # exactly representable values are unchanged, signs kept
    v = np.array([0.0, -2.5, 3.0, 2.0 ** -20, -(1 + 2.0 ** -7)])
    assert np.array_equal(mlp.bf16(v), v)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(5)
    shp = mlp.shapes(nin=5, nhid=8, nlayers=2, nout=3)
    npar = sum(n * k + n for n, k in shp)
    params = rng.normal(scale=0.5, size=npar)
    x = rng.normal(size=(16, 5))
    y = rng.normal(size=(16, 3))
    _, g = mlp.loss_grad(params, x, y, shp)
    eps = 1e-6
    fd = np.empty(npar)
    for i in range(npar):
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd[i] = (mlp.loss_grad(p1, x, y, shp)[0] - mlp.loss_grad(p2, x, y, shp)[0]) / (2 * eps)
    assert np.max(np.abs(g - fd)) <= 1e-7 * max(1.0, np.max(np.abs(fd)))


def test_predict_matches_the_loss_forward_pass():
    rng = np.random.default_rng(6)
    shp = mlp.shapes(nin=4, nhid=6, nlayers=3, nout=2)
    npar = sum(n * k + n for n, k in shp)
    params = rng.normal(scale=0.5, size=npar)
    x = rng.normal(size=(10, 4))
    yp = mlp.predict(params, x, shp)
    # the loss against the prediction itself is zero and so is its gradient
    loss, g = mlp.loss_grad(params, x, yp, shp)
    assert loss < 1e-28 and np.max(np.abs(g)) < 1e-14
