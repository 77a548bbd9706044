"""CPU FP32 reference of the candidate scorer network -- TEST INFRASTRUCTURE
ONLY (see oracle.h): only tests/ load it.

Network: 32 -> 128 -> 128 -> 1, ReLU after the two hidden layers, a logit
output trained with binary cross-entropy on sigmoid(logit), Adam updates.
Parameters are kept as a dict of float32 arrays:
    W1 [32, 128], b1 [128], W2 [128, 128], b2 [128], w3 [128], b3 [1]
and flattened in that order (the layout of curc_mlp_set_params).

The GPU kernels compute the matrix products in bfloat16 with float32
accumulation; `bf16=True` rounds weights and activations the same way so
the two can be compared tightly.
"""
import numpy as np

SHAPES = (("W1", (32, 128)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)), ("w3", (128,)), ("b3", (1,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(x):
    """Round to bfloat16 (nearest even) and back to float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def init(seed=0):
    """He-uniform initialisation; biases zero."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in SHAPES:
        if name.startswith("b"):
            p[name] = np.zeros(shape, np.float32)
        else:
            fan_in = shape[0]
            lim = np.sqrt(6.0 / fan_in)
            p[name] = rng.uniform(-lim, lim, shape).astype(np.float32)
    return p


def flatten(p):
    return np.concatenate([p[n].reshape(-1) for n, _ in SHAPES]).astype(np.float32)


def unflatten(v):
    v = np.asarray(v, np.float32)
    p, o = {}, 0
    for n, s in SHAPES:
        k = int(np.prod(s))
        p[n] = v[o:o + k].reshape(s).copy()
        o += k
    return p


def forward(p, x, bf=False):
    """Returns (logit [n], cache)."""
    r = bf16 if bf else (lambda a: a)
    x = r(np.asarray(x, p["W1"].dtype))
    z1 = x @ r(p["W1"]) + p["b1"]
    h1 = np.maximum(z1, 0)
    z2 = r(h1) @ r(p["W2"]) + p["b2"]
    h2 = np.maximum(z2, 0)
    s = h2 @ p["w3"] + p["b3"][0]
    return s, (x, z1, h1, z2, h2)


def loss(p, x, y):
    """Mean binary cross-entropy of sigmoid(logit) against labels in {0, 1}."""
    s, _ = forward(p, x)
    y = np.asarray(y, s.dtype)
    return float(np.mean(np.logaddexp(0.0, s) - y * s))


def grads(p, x, y, bf=False):
    """Analytic gradients of loss() (mean over the batch)."""
    s, (xx, z1, h1, z2, h2) = forward(p, x, bf)
    r = bf16 if bf else (lambda a: a)
    n = len(s)
    y = np.asarray(y, s.dtype)
    ds = (1.0 / (1.0 + np.exp(-s)) - y) / n
    g = {"w3": h2.T @ ds, "b3": np.array([ds.sum()], s.dtype)}
    dz2 = np.outer(ds, p["w3"]) * (z2 > 0)
    g["W2"] = r(h1).T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ r(p["W2"]).T) * (z1 > 0)
    g["W1"] = xx.T @ dz1
    g["b1"] = dz1.sum(0)
    return {k: np.asarray(v, s.dtype) for k, v in g.items()}


def adam(p, g, state, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in place; state holds m, v and the step count t."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for n, _ in SHAPES:
        m = state.setdefault("m_" + n, np.zeros_like(p[n]))
        v = state.setdefault("v_" + n, np.zeros_like(p[n]))
        m[:] = b1 * m + (1 - b1) * g[n]
        v[:] = b2 * v + (1 - b2) * g[n] * g[n]
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        p[n] -= (lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)
    return p
