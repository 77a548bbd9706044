"""The scorer network's GPU kernels (bfloat16 MFMA forward, backward and
Adam) against the CPU float32 reference in oracle/mlp.py."""
import numpy as np
import pytest

from curc_amd import scorer
from oracle import mlp

pytestmark = pytest.mark.gpu


def data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 32)).astype(np.float32)
    y = (rng.random(n) < 0.3).astype(np.float32)
    return x, y


def test_forward_matches_reference(gpu_ctx):
    p = scorer.init_params(1)
    net = scorer.Mlp(gpu_ctx, p)
    for n in (1, 31, 32, 33, 1000, 4099):
        x, _ = data(n, n)
        print(f"mlp forward n={n}", flush=True)
        got = net.forward(x)
        ref, _ = mlp.forward(mlp.unflatten(p), x, bf=True)
        # bf16 operands, float32 accumulation: compare on the logit scale
        tol = 2e-2 * (np.abs(ref).max() + 1)
        assert np.max(np.abs(got - ref)) < tol, (n, np.max(np.abs(got - ref)))
    net.close()


def test_gradients_match_reference(gpu_ctx):
    p = scorer.init_params(2)
    net = scorer.Mlp(gpu_ctx, p)
    for n in (64, 777, 5000):
        x, y = data(100 + n, n)
        print(f"mlp gradients n={n}", flush=True)
        g = net.gradients(x, y)
        ref = mlp.flatten(mlp.grads(mlp.unflatten(p), x, y, bf=True))
        gp = mlp.unflatten(g)
        rp = mlp.unflatten(ref)
        for name, _ in mlp.SHAPES:
            a, b = gp[name], rp[name]
            scale = np.abs(b).max() + 1e-6
            assert np.max(np.abs(a - b)) <= 3e-2 * scale, (n, name, np.max(np.abs(a - b)), scale)
    net.close()


def test_training_reduces_loss_like_reference(gpu_ctx):
    """A learnable rule (label = x0 + x1 > 0.5): Adam steps on the GPU lower
    the loss as the float32 reference's steps do."""
    rng = np.random.default_rng(9)
    x = rng.standard_normal((8192, 32)).astype(np.float32)
    y = (x[:, 0] + x[:, 1] > 0.5).astype(np.float32)
    p0 = scorer.init_params(3)
    net = scorer.Mlp(gpu_ctx, p0)
    dx, dy = scorer.DeviceArray(gpu_ctx, x), scorer.DeviceArray(gpu_ctx, y)
    first = net.train_step_dev(dx.ptr, dy.ptr, len(x), 3e-3, want_loss=True)
    for _ in range(150):
        last = net.train_step_dev(dx.ptr, dy.ptr, len(x), 3e-3, want_loss=True)
    ref = mlp.unflatten(p0)
    st = {}
    ref_first = mlp.loss(ref, x, y)
    for _ in range(151):
        mlp.adam(ref, mlp.grads(ref, x, y), st, lr=3e-3)
    ref_last = mlp.loss(ref, x, y)
    assert abs(first - ref_first) < 0.02 * (1 + ref_first)
    assert last < 0.5 * first
    assert last < ref_last + 0.05
    acc = np.mean((net.forward(x) > 0) == (y > 0.5))
    assert acc > 0.95
    dx.free()
    dy.free()
    net.close()
