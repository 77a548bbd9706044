"""curc_fastq_pack_static (the entry that allocates nothing, its scratch a
fixed device array) against curc_fastq_pack: same record count, packed
words, N flags and sequence starts, on two calls in a row, the second with
a smaller input (the fixed scratch keeps the first call's contents)."""
import numpy as np
import pytest

import curc_amd as ca
from test_gpu_fastq import records, workload

pytestmark = pytest.mark.gpu


def run(ctx, fn, fq, L, max_reads):
    W = (L + 31) // 32
    d_fq = ctx.dev_alloc(len(fq))
    d_p = ctx.dev_alloc(max_reads * W * 8)
    d_n = ctx.dev_alloc(max_reads + 4)
    d_s = ctx.dev_alloc(max_reads * 8)
    try:
        ctx.h2d(d_fq, fq)
        n = fn(d_fq, len(fq), L, max_reads, d_p, d_n, d_s)
        packed = np.empty(n * W, dtype=np.uint64)
        isn = np.empty(n, dtype=np.uint8)
        starts = np.empty(n, dtype=np.uint64)
        if n:
            ctx.d2h(packed, d_p)
            ctx.d2h(isn, d_n)
            ctx.d2h(starts, d_s)
        return n, packed, isn, starts
    finally:
        for p in (d_fq, d_p, d_n, d_s):
            ctx.dev_free(p)


def test_static_entry_matches_allocating_entry():
    ctx = ca.Context(0)
    try:
        L = 150
        for k, n in enumerate((60_000, 7_000)):
            reads = workload(101 + k, n, L, n_rate=0.003)
            fq = np.frombuffer(records(reads), dtype=np.uint8)
            print(f"call {k}: {n} reads, {len(fq)} bytes", flush=True)
            want = run(ctx, ctx.fastq_pack, fq, L, n)
            got = run(ctx, ctx.fastq_pack_static, fq, L, n)
            assert got[0] == want[0] == n
            for a, b in zip(got[1:], want[1:]):
                assert np.array_equal(a, b)
        # a malformed record is still an error, and the scratch is reusable after it
        bad = records(reads).replace(reads[5].tobytes(), reads[5].tobytes()[:-2], 1)
        with pytest.raises(ca.CurcError, match="read length"):
            run(ctx, ctx.fastq_pack_static, np.frombuffer(bad, dtype=np.uint8), L, n)
        got = run(ctx, ctx.fastq_pack_static, fq, L, n)
        assert got[0] == n and np.array_equal(got[3], want[3])
        # beyond the fixed maximum: refused before any device work
        d = ctx.dev_alloc(256)
        try:
            with pytest.raises(ca.CurcError, match=r"\(-4\)"):
                ctx.fastq_pack_static(d, 16_000_000_001, L, 1, d, None, d)
        finally:
            ctx.dev_free(d)
    finally:
        ctx.close()
