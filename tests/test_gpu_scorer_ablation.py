"""What the trained scorer gains over mapping without it, on a workload whose
seed is disjoint from the training seeds: far fewer k-mer probes and
verified candidates, nearly every read still mapped, and a lossless
archive no more than 1% larger.  A network with random weights, probing as
little, maps fewer reads and writes a larger archive."""
import os
import tempfile

import numpy as np
import pytest

import curc_amd as ca
from curc_amd import archive, scorer, train_scorer

pytestmark = pytest.mark.gpu


def run(ctx, reads, learned):
    ctx.reset_stats()
    ctx.set_timing(True)
    b = ctx.run_block(reads, learned=learned, streams=True, refmatch=True)
    ctx.set_timing(False)
    st = {s["name"]: s for s in ctx.stats()}
    work = {k: st[k]["items"] for k in ("map_probes", "map_cands") if k in st}
    mis = b.get("map_mis").astype(np.int64)
    bs = archive.BlockStreams.from_block(b, len(reads), reads.shape[1], False, False)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "a.curc")
        archive.write_archive(p, [bs], False, False, reads.shape[1])
        size = os.path.getsize(p)
        _, _, L, blocks = archive.read_archive(p)
        dec = archive.decode_block(blocks[0], L, False, False)
    assert sorted(map(bytes, dec)) == sorted(map(bytes, reads))
    b.free()
    return dict(probes=work["map_probes"], cands=work["map_cands"], mapped=len(mis), mis=int(mis.sum()),
                size=size)


def test_trained_scorer_gains(gpu_ctx):
    cfg = ca.config_scale(seed=31, n_reads=600_000, genome_len=3_000_000)
    _, reads = ca.synth_reads(cfg)
    print("ablation: plain", flush=True)
    # the method the scorer stops early: the exhaustive loop of each strand
    # pass (the one-launch plain kernel also ends passes early, by a rule
    # that never changes a result)
    before = gpu_ctx.get_tuning()
    gpu_ctx.set_tuning(map_both=0)
    plain = run(gpu_ctx, reads, False)
    gpu_ctx.set_tuning(**before)
    params, thr = train_scorer.load()
    net = scorer.Mlp(gpu_ctx, params)
    gpu_ctx.set_scorer(net, thr)
    print("ablation: trained scorer", flush=True)
    got = run(gpu_ctx, reads, True)
    print(plain, got, flush=True)
    assert got["probes"] < 0.6 * plain["probes"]
    assert got["cands"] < 0.7 * plain["cands"]
    assert got["mapped"] >= 0.99 * plain["mapped"]
    assert got["size"] <= 1.01 * plain["size"]
    # random weights, threshold set to stop at least as much probing
    rnd = scorer.Mlp(gpu_ctx, scorer.init_params(11))
    x = np.random.default_rng(0).random((4096, 32)).astype(np.float32)
    lr = rnd.forward(x)
    best = None
    for q in (0.5, 0.7, 0.85):
        gpu_ctx.set_scorer(rnd, float(np.quantile(lr, q)))
        print(f"ablation: random network, quantile {q}", flush=True)
        r = run(gpu_ctx, reads, True)
        if r["probes"] <= got["probes"]:
            best = r
            break
    gpu_ctx.set_scorer(None)
    net.close()
    rnd.close()
    print("random network:", best, flush=True)
    if best is not None:
        assert got["mapped"] > best["mapped"] and got["size"] < best["size"]


@pytest.mark.parametrize("L,sub,genome", [(75, 0.01, 2_000_000), (250, 0.006, 6_000_000), (100, 0.003, 30_000_000)])
def test_trained_scorer_other_read_lengths(gpu_ctx, L, sub, genome):
    """Read lengths, uniform error rates and coverages of the many-input
    workload: the scorer keeps nearly every mapping and the archive size, and
    never probes more than the exhaustive loop.  (At 0.8x coverage almost no
    read maps, and a read that will never map cannot be told early from one
    that maps late, so there the savings are small.)"""
    cfg = ca.synth_config(seed=57 + L, genome_len=genome, gc=0.45, n_reads=250_000, L=L, sub_first=sub,
                          sub_last=sub, n_rate=0.0002)
    _, reads = ca.synth_reads(cfg)
    print(f"ablation: L={L}, subst {sub}, genome {genome}", flush=True)
    before = gpu_ctx.get_tuning()
    gpu_ctx.set_tuning(map_both=0)  # the exhaustive loop, as above
    plain = run(gpu_ctx, reads, False)
    gpu_ctx.set_tuning(**before)
    params, thr = train_scorer.load()
    net = scorer.Mlp(gpu_ctx, params)
    gpu_ctx.set_scorer(net, thr)
    try:
        got = run(gpu_ctx, reads, True)
    finally:
        gpu_ctx.set_scorer(None)
        net.close()
    print(plain, got, flush=True)
    assert got["probes"] <= plain["probes"]
    assert got["mapped"] >= 0.99 * plain["mapped"]
    assert got["size"] <= 1.01 * plain["size"]
