"""The FP64 MLP oracle against mathematics: finite differences and closed forms."""
import numpy as np

from oracle import mlp_oracle as mo


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(mo.N_PARAMS)
    for k, (o, n) in mo.OFF.items():
        fan = mo.SHAPES[k][0] if k in mo.SHAPES else 1
        p[o:o + n] = rng.uniform(-1, 1, n) * (np.sqrt(6.0 / fan) if k in mo.SHAPES else 0.1) * scale
    return p


def test_layout_size():
    assert mo.N_PARAMS == 200196
    assert mo.OFF["b5"][0] + 4 == mo.N_PARAMS


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    p = _params(1)
    x = rng.normal(size=(8, 6))
    t = rng.normal(size=(8, 4))
    _, g = mo.loss_grad(p, x, t)
    # a few entries of every tensor
    for k, (o, n) in mo.OFF.items():
        for i in rng.choice(n, size=4, replace=False):
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[o + i] += h
            pm[o + i] -= h
            fd = (mo.loss_grad(pp, x, t)[0] - mo.loss_grad(pm, x, t)[0]) / (2 * h)
            assert abs(fd - g[o + i]) <= 1e-6 * max(1.0, abs(fd)), (k, i, fd, g[o + i])


def test_linear_limit():
    """With W2..W4 = 0 the hidden layers output silu(b): y is affine in nothing but b5, W5."""
    p = _params(2)
    for k in ("W2", "W3", "W4"):
        o, n = mo.OFF[k]
        p[o:o + n] = 0
    x = np.random.default_rng(0).normal(size=(5, 6))
    P = mo.unpack(p)
    h = P["b4"] / (1 + np.exp(-P["b4"]))
    y = mo.forward(p, x)
    np.testing.assert_allclose(y, np.tile(h @ P["W5"] + P["b5"], (5, 1)), rtol=1e-13, atol=1e-13)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], np.float32)
    r = mo.bf16(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1
    np.testing.assert_array_equal(r[:2], [1.0, 1.0])
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8
