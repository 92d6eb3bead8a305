"""The FP64 network oracle (oracle/mlp.py) against what the mathematics
fixes: its gradient against central finite differences of its own loss in
every parameter block, and GELU against its closed-form limits.  A wrong
transpose, a dropped bias gradient or a wrong derivative of GELU fails here."""
import numpy as np

from oracle import mlp


def _params(rng):
    p = rng.normal(0.0, 0.2, mlp.NPARAM)
    w1, b1, w2, b2, w3, b3, w4, b4 = mlp.unpack(p)
    w1[mlp.NIN:, :] = 0.0
    w4[:, mlp.NOUT:] = 0.0
    b4[mlp.NOUT:] = 0.0
    return mlp.pack(w1, b1, w2, b2, w3, b3, w4, b4)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = _params(rng)
    x = np.zeros((12, mlp.INP))
    x[:, :mlp.NIN] = rng.normal(size=(12, mlp.NIN))
    t = rng.normal(size=(12, mlp.NOUT))
    loss, g = mlp.loss_grad(p, x, t)
    blocks = [(mlp.OW1, mlp.OB1), (mlp.OB1, mlp.OW2), (mlp.OW2, mlp.OB2), (mlp.OB2, mlp.OW3),
              (mlp.OW3, mlp.OB3), (mlp.OB3, mlp.OW4), (mlp.OW4, mlp.OB4), (mlp.OB4, mlp.NPARAM)]
    h = 1e-6
    for lo, hi in blocks:
        idx = rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False)
        for i in idx:
            e = np.zeros_like(p)
            e[i] = h
            fd = (mlp.loss_grad(p + e, x, t)[0] - mlp.loss_grad(p - e, x, t)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (lo, i, fd, g[i])
    # the padded inputs and outputs get no gradient
    assert np.all(g[mlp.OW1 + mlp.NIN * mlp.H:mlp.OB1] == 0.0)
    w4g = g[mlp.OW4:mlp.OB4].reshape(mlp.H, mlp.OUTP)
    assert np.all(w4g[:, mlp.NOUT:] == 0.0)


def test_gelu_limits_and_derivative():
    z = np.linspace(-8.0, 8.0, 161)
    assert mlp.gelu(np.array([0.0]))[0] == 0.0
    assert np.allclose(mlp.gelu(np.array([8.0, 20.0])), [8.0, 20.0], rtol=1e-12)
    assert np.allclose(mlp.gelu(np.array([-8.0, -20.0])), [0.0, 0.0], atol=1e-12)
    h = 1e-6
    fd = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    assert np.allclose(fd, mlp.dgelu(z), atol=1e-8)


def test_bf16_rounding_is_nearest_even():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float32)
    r = mlp.bf16(a)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 2 ** -8 * 3.0e-3
