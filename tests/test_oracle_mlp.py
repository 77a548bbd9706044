"""Pins the CPU reference of the candidate scorer network: analytic
gradients against central finite differences in float64, the bfloat16
rounding against its bit-level definition, and Adam against its closed
form on a quadratic."""
import numpy as np

from oracle import mlp


def p64(seed):
    p = mlp.init(seed)
    rng = np.random.default_rng(seed + 1)
    for k in p:
        p[k] = (p[k] + 0.05 * rng.standard_normal(p[k].shape)).astype(np.float64)
    return p


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((24, 32))
    y = (rng.random(24) < 0.4).astype(np.float64)
    p = p64(5)
    g = mlp.grads(p, x, y)
    eps = 1e-6
    for name, shape in mlp.SHAPES:
        idx = [tuple(rng.integers(0, d) for d in shape) for _ in range(6)]
        for ix in idx:
            q = {k: v.copy() for k, v in p.items()}
            q[name][ix] += eps
            lp = mlp.loss(q, x, y)
            q[name][ix] -= 2 * eps
            lm = mlp.loss(q, x, y)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[name][ix]) <= 1e-6 + 1e-4 * abs(fd), (name, ix, fd, g[name][ix])


def test_gradient_sign_mistake_is_caught():
    """A planted mistake (dropping the ReLU mask of the first layer) must
    fail the finite-difference comparison."""
    rng = np.random.default_rng(4)
    x = rng.standard_normal((16, 32))
    y = (rng.random(16) < 0.5).astype(np.float64)
    p = p64(6)
    s, (xx, z1, h1, z2, h2) = mlp.forward(p, x)
    ds = (1 / (1 + np.exp(-s)) - y) / len(s)
    dz2 = np.outer(ds, p["w3"]) * (z2 > 0)
    wrong = xx.T @ (dz2 @ p["W2"].T)  # no (z1 > 0) mask
    right = mlp.grads(p, x, y)["W1"]
    assert not np.allclose(wrong, right, rtol=1e-3, atol=1e-8)


def test_bf16_rounding_definition():
    vals = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-30, 0.0], np.float32)
    r = mlp.bf16(vals)
    u = r.view(np.uint32)
    assert np.all(u & 0xFFFF == 0)
    # nearest with ties to even: 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> 1
    assert r[1] == 1.0
    assert abs(r[2] - 1.0078125) < 1e-9
    assert np.all(np.abs(r - vals) <= np.abs(vals) * 2.0 ** -8)


def test_adam_matches_closed_form_first_step():
    p = {n: np.ones(s, np.float32) for n, s in mlp.SHAPES}
    g = {n: np.full(s, 0.5, np.float32) for n, s in mlp.SHAPES}
    st = {}
    mlp.adam(p, g, st, lr=0.01)
    # first step: m_hat = g, v_hat = g^2, so the update is lr * sign(g)
    for n, _ in mlp.SHAPES:
        assert np.allclose(p[n], 1.0 - 0.01, atol=1e-6)


def test_flatten_roundtrip():
    p = mlp.init(1)
    v = mlp.flatten(p)
    assert v.size == mlp.N_PARAMS == 32 * 128 + 128 + 128 * 128 + 128 + 128 + 1
    q = mlp.unflatten(v)
    for n, _ in mlp.SHAPES:
        assert np.array_equal(p[n], q[n])
