"""Regression vectors under tests/golden/ (written by make_fixtures.py).

The reference holds no golden vectors for this path, so these are the
oracle's own outputs on three small seeded blocks (150 bp; 75 bp with half
the reads reverse-complemented; 100 bp from a genome with diverged repeat
families).  The CPU test pins the oracle to them, so any change to its
results shows up here; the GPU test checks the HIP path against the same
arrays without running the oracle at all.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("sorted_ids", "next", "prev", "offset", "hq_pg", "hq_rec", "map_rec", "map_mis", "unmapped", "unmapped_N",
          "lq_pg", "lq_rec", "nlq_pg", "nlq_rec")


def load():
    z = np.load(os.path.join(HERE, "blocks.npz"))
    cases = sorted({k.split("/")[0] for k in z.files})
    return {c: {k.split("/")[1]: z[k] for k in z.files if k.startswith(c + "/")} for c in cases}


CASES = load()


def flat(v):
    return v.view(np.uint8).reshape(-1) if v.dtype.names else v


def check(blk, want, case):
    for f in FIELDS:
        a, b = flat(blk.get(f)), want[f]
        assert a.shape == b.shape, f"{case}/{f}: length {a.shape} vs fixture {b.shape}"
        if not np.array_equal(a, b):
            i = int(np.flatnonzero(a != b)[0])
            raise AssertionError(f"{case}/{f}: first difference at {i}: got {a[i]} fixture {b[i]}")


def test_fixture_cases_present():
    assert set(CASES) == {"l150", "l75_rc", "l100_repeats"}
    for c, d in CASES.items():
        assert set(FIELDS) <= set(d), c
        n = d["reads"].shape[0]
        ids = d["sorted_ids"]
        # reads without N, each once, in dedup order
        assert len(np.unique(ids)) == len(ids) and 0 < len(ids) <= n and ids.max() < n


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_matches_fixture(case):
    want = CASES[case]
    ob = orc.OracleBlock(want["reads"], stop="all")
    check(ob, want, case)
    ob.free()


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_matches_fixture(gpu_ctx, case):
    want = CASES[case]
    print(f"golden case {case}: {want['reads'].shape}", flush=True)
    g = gpu_ctx.run_block(want["reads"], keep_stages=True)
    try:
        check(g, want, case)
    finally:
        g.free()
