"""Offset loops replayed as HIP graphs (contexts created with
CURC_GRAPHS=1): a block run again and again (first eagerly, then captured,
then replayed), with other blocks and host work in between, gives the same
results every time, equal to the oracle and to a context without graphs."""
import os

import numpy as np
import pytest

import curc_amd as ca
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("hq_pg", "hq_rec", "map_rec", "map_mis", "unmapped", "unmapped_N", "lq_pg", "lq_rec", "nlq_pg",
          "nlq_rec")


def reads_of(seed, n, L, genome):
    cfg = ca.synth_config(seed=seed, genome_len=genome, gc=0.45, n_reads=n, L=L, sub_first=0.003, sub_last=0.01,
                          n_rate=0.001)
    return ca.synth_reads(cfg)[1]


class graphs_env:
    def __init__(self, value):
        self.value = value

    def __enter__(self):
        self.old = os.environ.get("CURC_GRAPHS")
        os.environ["CURC_GRAPHS"] = self.value

    def __exit__(self, *exc):
        if self.old is None:
            del os.environ["CURC_GRAPHS"]
        else:
            os.environ["CURC_GRAPHS"] = self.old


def results(b):
    out = {f: b.get(f) for f in FIELDS}
    b.free()
    return out


def test_replays_match_eager_runs_and_oracle():
    a = reads_of(901, 12_000, 150, 80_000)
    other = reads_of(902, 7_000, 100, 50_000)
    ref = orc.OracleBlock(a)
    with graphs_env("1"), ca.Context(0) as ctx:
        ctx.reset_stats()
        runs = []
        for k in range(5):
            print(f"graph replay test: run {k}", flush=True)
            runs.append(results(ctx.run_block(a)))
            # other work on the same context between replays
            results(ctx.run_block(other))
            ctx.device_sync()
        st = {s["name"]: s["items"] for s in ctx.stats()}
    assert st.get("graph_captures", 0) > 0 and st.get("graph_replays", 0) > 0, st
    for k, r in enumerate(runs):
        for f in FIELDS:
            assert np.array_equal(r[f], ref.get(f)), f"run {k}: {f} differs from the oracle"


def test_graphs_off_gives_same_results():
    a = reads_of(903, 9_000, 150, 60_000)
    with graphs_env("0"), ca.Context(0) as ctx:
        plain = [results(ctx.run_block(a)) for _ in range(3)]
        st = {s["name"]: s["items"] for s in ctx.stats()}
    assert st.get("graph_replays", 0) == 0 and st.get("graph_captures", 0) == 0
    with graphs_env("1"), ca.Context(0) as ctx:
        graphed = [results(ctx.run_block(a)) for _ in range(3)]
    for p, g in zip(plain, graphed):
        for f in FIELDS:
            assert np.array_equal(p[f], g[f]), f
