"""The learned candidate scorer: host side of the curc_mlp_* entry points.

Network 32 -> 128 -> 128 -> 1 (ReLU), bfloat16 MFMA products with float32
accumulation, trained by the library's own kernels with Adam on
binary cross-entropy.  Parameters are a flat float32 vector in the order
W1[32][128], b1[128], W2[128][128], b2[128], w3[128], b3.
"""
import ctypes

import numpy as np

from . import CurcError, lib

N_PARAMS = 20865
N_FEATURES = 32


def init_params(seed=0):
    """He-uniform weights, zero biases, as a flat float32 vector."""
    rng = np.random.default_rng(seed)
    parts = []
    for shape, fan_in in (((32, 128), 32), ((128,), 0), ((128, 128), 128), ((128,), 0), ((128,), 128), ((1,), 0)):
        if fan_in:
            lim = np.sqrt(6.0 / fan_in)
            parts.append(rng.uniform(-lim, lim, shape).astype(np.float32).reshape(-1))
        else:
            parts.append(np.zeros(shape, np.float32).reshape(-1))
    v = np.concatenate(parts)
    assert v.size == N_PARAMS
    return v


class DeviceArray:
    """A float32 host array copied to device memory of a context."""

    def __init__(self, ctx, arr):
        self.ctx = ctx
        self.host = np.ascontiguousarray(arr, dtype=np.float32)
        self.ptr = ctx.dev_alloc(max(4, self.host.nbytes))
        if self.host.size:
            ctx.h2d(self.ptr, self.host)

    def free(self):
        if self.ptr:
            self.ctx.dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Mlp:
    def __init__(self, ctx, params=None):
        self.ctx = ctx
        h = lib().curc_mlp_create(ctx._h)
        if not h:
            raise CurcError("curc_mlp_create failed")
        self._h = ctypes.c_void_p(h)
        ctx._dependents.append(self)
        if params is not None:
            self.set_params(params)

    def _check(self, rc, what):
        if rc != 0:
            msg = lib().curc_ctx_last_error(self.ctx._h)
            raise CurcError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(lib().curc_mlp_set_params(self._h, p.ctypes.data, p.size), "curc_mlp_set_params")

    def get_params(self):
        p = np.zeros(N_PARAMS, np.float32)
        self._check(lib().curc_mlp_get_params(self._h, p.ctypes.data, N_PARAMS), "curc_mlp_get_params")
        return p

    def forward(self, x):
        """Logits of host features x [n, 32]."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, N_FEATURES)
        dx = DeviceArray(self.ctx, x)
        out = DeviceArray(self.ctx, np.zeros(len(x), np.float32))
        self._check(lib().curc_mlp_forward(self._h, ctypes.c_void_p(dx.ptr), len(x), ctypes.c_void_p(out.ptr)),
                    "curc_mlp_forward")
        res = np.zeros(len(x), np.float32)
        self.ctx.d2h(res, out.ptr)
        dx.free()
        out.free()
        return res

    def forward_dev(self, dx_ptr, n, dout_ptr):
        """Logits of n device feature rows (float32 [n, 32]) into device memory."""
        self._check(lib().curc_mlp_forward(self._h, ctypes.c_void_p(dx_ptr), n, ctypes.c_void_p(dout_ptr)),
                    "curc_mlp_forward")

    def gradients(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, N_FEATURES)
        dx, dy = DeviceArray(self.ctx, x), DeviceArray(self.ctx, np.asarray(y, np.float32))
        g = np.zeros(N_PARAMS, np.float32)
        self._check(lib().curc_mlp_gradients(self._h, ctypes.c_void_p(dx.ptr), ctypes.c_void_p(dy.ptr), len(x),
                                             g.ctypes.data), "curc_mlp_gradients")
        dx.free()
        dy.free()
        return g

    def train_step_dev(self, dx_ptr, dy_ptr, n, lr, want_loss=False):
        """One Adam step on a batch already in device memory."""
        loss = ctypes.c_float(0)
        self._check(lib().curc_mlp_train_step(self._h, ctypes.c_void_p(dx_ptr), ctypes.c_void_p(dy_ptr), n,
                                              ctypes.c_float(lr), ctypes.byref(loss) if want_loss else None),
                    "curc_mlp_train_step")
        return float(loss.value) if want_loss else None

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().curc_mlp_destroy(self._h)
            self._h = None
            try:
                self.ctx._dependents.remove(self)
            except ValueError:
                pass

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
