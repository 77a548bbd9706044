"""The 2-bit pack entry point (curc_pack) against the oracle's pack.

Reads hold upper and lower case bases, 'N' and 'n' and other symbols (which
pack as 0, and only an uppercase 'N' flags a read), at read lengths around
word boundaries, from device buffers starting at every byte offset modulo 4
(the kernel reads dwords and must not read outside the buffer).
"""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ALPHABET = np.frombuffer(b"ACGTacgtNnRX.", dtype=np.uint8)
P = np.array([0.2, 0.2, 0.2, 0.2, 0.03, 0.03, 0.03, 0.03, 0.01, 0.01, 0.01, 0.01, 0.02])


@pytest.mark.parametrize("L", [1, 31, 32, 33, 75, 150, 250])
def test_pack_entry_point(gpu_ctx, L):
    rng = np.random.default_rng(L)
    n = 1001
    reads = ALPHABET[rng.choice(len(ALPHABET), size=(n, L), p=P / P.sum())]
    want = orc.pack(reads)
    want_n = (reads == ord("N")).any(axis=1).astype(np.uint8)
    W = (L + 31) // 32
    for shift in range(4):
        print(f"L={L} shift={shift}", flush=True)
        buf = gpu_ctx.dev_alloc(n * L + 8)
        d_packed = gpu_ctx.dev_alloc(n * W * 8)
        d_isn = gpu_ctx.dev_alloc(n + 8)
        try:
            gpu_ctx.h2d(buf + shift, reads)
            gpu_ctx.pack(buf + shift, n, L, d_packed, d_isn + shift)
            got = np.empty((n, W), dtype=np.uint64)
            got_n = np.empty(n, dtype=np.uint8)
            gpu_ctx.d2h(got, d_packed)
            gpu_ctx.d2h(got_n, d_isn + shift)
            assert np.array_equal(got, want), f"L={L} shift={shift}: packed words differ"
            assert np.array_equal(got_n, want_n), f"L={L} shift={shift}: N flags differ"
        finally:
            gpu_ctx.dev_free(buf)
            gpu_ctx.dev_free(d_packed)
            gpu_ctx.dev_free(d_isn)
