"""Mapping in probe ranges with the scorer's checkpoints: without stops it
equals the exhaustive mapping; with stops every record stays exact and the
archive round trip stays lossless."""
import os
import tempfile

import numpy as np
import pytest

import curc_amd as ca
from curc_amd import archive, scorer

pytestmark = pytest.mark.gpu

MAP_FIELDS = ("map_rec", "map_mis", "unmapped", "unmapped_N", "lq_pg", "lq_rec", "nlq_pg", "nlq_rec")
COMP = np.zeros(256, np.uint8)
for a, b in zip(b"ACGTN", b"TGCAN"):
    COMP[a] = b


def reads_cfg(seed, L=150, n=20_000, g=100_000):
    cfg = ca.synth_config(seed=seed, genome_len=g, gc=0.45, n_reads=n, L=L, sub_first=0.002, sub_last=0.015,
                          n_rate=0.001)
    return ca.synth_reads(cfg)[1]


def recount(blk, reads):
    """Every map record's mismatch count, recounted against the HQ text."""
    pg = np.frombuffer(blk.get("hq_pg").tobytes(), np.uint8)
    rec = blk.get("map_rec")
    mis = blk.get("map_mis")
    nids = blk.get("N_ids") if len(blk.get("N_ids")) else None
    L = reads.shape[1]
    for r, m in zip(rec, mis):
        w = int(r["word"])
        isn, strand, pos = w & 1, (w >> 1) & 1, w >> 2
        rid = int(r["read_id"])
        rd = reads[nids[rid]] if isn else reads[rid]
        if strand:
            rd = COMP[rd[::-1]]
        ref = pg[pos:pos + L]
        cnt = int(np.count_nonzero(ref != rd))
        assert cnt == int(m) and cnt <= L // 5, (rid, cnt, int(m))


@pytest.mark.parametrize("L", [100, 150])
def test_ranges_without_stops_equal_exhaustive(gpu_ctx, L):
    r = reads_cfg(81 + L, L=L)
    print(f"trace vs plain L={L}", flush=True)
    plain = gpu_ctx.run_block(r)
    gpu_ctx.set_trace(every=4, max_records=1 << 20)
    tr = gpu_ctx.run_block(r, trace=True)
    for f in MAP_FIELDS:
        assert np.array_equal(plain.get(f), tr.get(f)), f
    x, y = tr.get("trace_x"), tr.get("trace_y")
    assert len(x) == len(y) > 100
    assert np.all(np.isfinite(x)) and set(np.unique(y)) <= {0.0, 1.0}
    assert 0 < y.mean() < 1


def test_never_stopping_scorer_equals_exhaustive(gpu_ctx):
    r = reads_cfg(90)
    print("learned with threshold -inf", flush=True)
    net = scorer.Mlp(gpu_ctx, scorer.init_params(4))
    gpu_ctx.set_scorer(net, threshold=-1e30)
    a = gpu_ctx.run_block(r)
    b = gpu_ctx.run_block(r, learned=True)
    for f in MAP_FIELDS:
        assert np.array_equal(a.get(f), b.get(f)), f
    gpu_ctx.set_scorer(None)
    net.close()


@pytest.mark.parametrize("threshold", [1e30, 0.0])
def test_stopping_keeps_records_exact_and_lossless(gpu_ctx, threshold):
    r = reads_cfg(91)
    print(f"learned with threshold {threshold}", flush=True)
    net = scorer.Mlp(gpu_ctx, scorer.init_params(5))
    gpu_ctx.set_scorer(net, threshold=threshold)
    plain = gpu_ctx.run_block(r, keep_stages=True)
    blk = gpu_ctx.run_block(r, keep_stages=True, learned=True, streams=True, refmatch=True)
    recount(blk, r)
    # a read maps only if the exhaustive loop maps it, and no better unless
    # both settle at the target (<= 1): a pass stopped early can leave the
    # other strand's pass a candidate that the exhaustive loop never needed
    key = lambda b: [int(x) | (int(w) & 1) << 32 for x, w in zip(b.get("map_rec")["read_id"], b.get("map_rec")["word"])]
    pm = dict(zip(key(plain), plain.get("map_mis").tolist()))
    for k, m in zip(key(blk), blk.get("map_mis").tolist()):
        assert k in pm and (m >= pm[k] or m <= 1)
    bs = archive.BlockStreams.from_block(blk, len(r), r.shape[1], False, False)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "l.curc")
        archive.write_archive(p, [bs], False, False, r.shape[1])
        _, _, L, blocks = archive.read_archive(p)
        dec = archive.decode_block(blocks[0], L, False, False)
    assert sorted(map(bytes, dec)) == sorted(map(bytes, r))
    gpu_ctx.set_scorer(None)
    net.close()


def test_ranges_on_the_compact_index():
    """A block large enough for the bucket-inline index (6 Mbp genome, 1.2M
    reads): the probe ranges run on the compact two-strand index (tuning
    map_both 2) or on one strand's 64-byte slots (map_both 0).  Without
    stops both equal the plain mapping and walk the same probes and
    candidates as its exhaustive per-strand form (map_both 0); with a
    stopping scorer both stop the same reads, since the features do not
    depend on the index layout."""
    cfg = ca.config_scale(seed=83, n_reads=1_200_000, genome_len=6_000_000)
    r = ca.synth_reads(cfg)[1]

    def work(ctx, **kw):
        ctx.reset_stats()
        ctx.set_timing(True)
        b = ctx.run_block(r, **kw)
        ctx.set_timing(False)
        st = {s["name"]: s for s in ctx.stats()}
        return b, {k: st[k]["items"] for k in ("map_probes", "map_cands")}

    with ca.Context(0) as ctx:
        ctx.set_trace(every=64, max_records=1 << 20)
        print("compact index: plain", flush=True)
        plain = work(ctx)[0]
        ctx.set_tuning(map_both=0)
        wp = work(ctx)[1]
        runs = {}
        for mb in (2, 0):
            ctx.set_tuning(map_both=mb)
            print(f"compact index: trace, map_both {mb}", flush=True)
            runs[mb] = work(ctx, trace=True)
        for mb, (b, w) in runs.items():
            for f in MAP_FIELDS:
                assert np.array_equal(plain.get(f), b.get(f)), (mb, f)
            assert w == wp, (mb, w, wp)
        # trace records come in the order their reads claimed them
        rows = lambda b: (lambda x: x[np.lexsort(x.T[::-1])])(b.get("trace_x"))
        assert np.array_equal(rows(runs[2][0]), rows(runs[0][0]))
        net = scorer.Mlp(ctx, scorer.init_params(5))
        ctx.set_scorer(net, threshold=0.0)
        got = {}
        for mb in (2, 0):
            ctx.set_tuning(map_both=mb)
            print(f"compact index: learned, map_both {mb}", flush=True)
            got[mb] = work(ctx, learned=True)
        for f in MAP_FIELDS:
            assert np.array_equal(got[2][0].get(f), got[0][0].get(f)), f
        assert got[2][1] == got[0][1]
        assert got[2][1]["map_probes"] <= wp["map_probes"]
        recount(got[2][0], r)
        ctx.set_scorer(None)
        net.close()
