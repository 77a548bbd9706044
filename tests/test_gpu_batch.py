"""Batched runs (curc_batch_run): many inputs of uneven read counts and
lengths on concurrent streams give, input by input, the same results as
single runs and the oracle."""
import numpy as np
import pytest

import curc_amd as ca
from curc_amd import scorer, train_scorer
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("hq_pg", "hq_rec", "map_rec", "map_mis", "unmapped", "unmapped_N", "lq_pg", "lq_rec", "nlq_pg",
          "nlq_rec")


def inputs(seed, k):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        L = int(rng.choice([75, 100, 150, 250]))
        n = int(np.exp(rng.uniform(np.log(300), np.log(20_000))))
        cfg = ca.synth_config(seed=seed * 100 + i, genome_len=max(3_000, n * L // 20), gc=0.45, n_reads=n, L=L,
                              sub_first=0.003, sub_last=0.01, n_rate=0.001)
        out.append(ca.synth_reads(cfg)[1])
    return out


def test_batch_matches_single_runs_and_oracle(gpu_ctx):
    ins = inputs(5, 10)
    print("batch of", [r.shape for r in ins], flush=True)
    blocks = gpu_ctx.batch_run(ins, workers=4)
    for i, (r, b) in enumerate(zip(ins, blocks)):
        s = gpu_ctx.run_block(r)
        for f in FIELDS:
            assert np.array_equal(b.get(f), s.get(f)), f"input {i}: {f} differs from the single run"
        if i < 3:
            o = orc.OracleBlock(r)
            for f in FIELDS:
                assert np.array_equal(b.get(f), o.get(f)), f"input {i}: {f} differs from the oracle"
        s.free()
        b.free()


@pytest.mark.parametrize("workers", [1, 2, 3])
def test_batch_worker_counts_agree(gpu_ctx, workers):
    ins = inputs(6, 7)
    print(f"batch with {workers} workers", flush=True)
    ref = gpu_ctx.batch_run(ins, workers=4, streams=True, refmatch=True)
    got = gpu_ctx.batch_run(ins, workers=workers, streams=True, refmatch=True)
    for i, (a, b) in enumerate(zip(ref, got)):
        for f in FIELDS + ("s_strand", "s_read_off", "s_mis_count", "s_mis_base", "rm_hq_text", "rm_hq_len",
                           "rm_lq_off"):
            assert np.array_equal(a.get(f), b.get(f)), f"input {i}: {f} differs"
        a.free()
        b.free()


def test_batch_empty_and_single(gpu_ctx):
    assert gpu_ctx.batch_run([]) == []
    r = inputs(7, 1)[0]
    (b,) = gpu_ctx.batch_run([r])
    s = gpu_ctx.run_block(r)
    assert np.array_equal(b.get("hq_pg"), s.get("hq_pg"))


def test_batch_learned_matches_single_learned_runs(gpu_ctx):
    """Worker streams map under the calling context's scorer."""
    ins = inputs(8, 6)
    params, thr = train_scorer.load()
    net = scorer.Mlp(gpu_ctx, params)
    gpu_ctx.set_scorer(net, thr)
    try:
        print("learned batch of", [r.shape for r in ins], flush=True)
        blocks = gpu_ctx.batch_run(ins, workers=4, learned=True)
        for i, (r, b) in enumerate(zip(ins, blocks)):
            s = gpu_ctx.run_block(r, learned=True)
            for f in FIELDS:
                assert np.array_equal(b.get(f), s.get(f)), f"input {i}: {f} differs from the single learned run"
            s.free()
            b.free()
    finally:
        gpu_ctx.set_scorer(None)
        net.close()
