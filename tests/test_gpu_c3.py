"""The repeat-heavy paired-end configuration at its stated size: a 500 Mbp
genome, 45% of it copies of 300 bp - 6 kbp repeat elements at 5-15%
divergence, 2 x 100 bp pairs (fragment ~N(400, 50)), 20x coverage = 100M
reads, order preserved, cut by block_ratio 0.25 at the reference's 50 MB
chunk granularity into 4 uneven blocks that run in one batch from FASTQ
bytes in pinned host memory.

Every block: the counts of placed reads add up; every read is rebuilt from
the pseudogenomes, records and mismatch streams (GPU read rebuild) exactly
as the input, in order; map records carry at most L/5 mismatches and the
streams' recount against the pseudogenome equals them.  Block 0 (25M
reads; a prefix hash table of more than 2^24 slots, with the skewed buckets
of the repeat families) also equals the OpenMP oracle."""
import time

import numpy as np
import pytest

import curc_amd as ca
from curc_amd import archive, fastq, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("hq_pg", "hq_rec", "map_rec", "map_mis", "unmapped", "unmapped_N", "lq_pg", "lq_rec", "nlq_pg",
          "nlq_rec")


def log(msg, t0):
    print(f"[{time.perf_counter() - t0:6.1f} s] {msg}", flush=True)


def gids(rec, Nids):
    ids = rec["read_id"].astype(np.int64)
    isn = (rec["word"] & 1).astype(bool)
    ids[isn] = Nids[ids[isn]]
    return ids


@pytest.mark.timeout(1200)
def test_repeat_heavy_pairs_at_full_size(gpu_ctx):
    t0 = time.perf_counter()
    cfg = workloads.config_repeat_pe()
    _, reads = ca.synth_reads(cfg)
    L = 100
    log(f"generated {len(reads)} reads", t0)
    f1, f2 = workloads.pinned_fastq(gpu_ctx, reads, True)
    ranges = fastq.block_ranges(f1.array, f2.array, 0.25)
    rec_len = ca.synth_fastq_record_len(L, 1)
    parts = workloads.block_reads(reads, ranges, rec_len, True)
    sizes = [len(p) for p in parts]
    log(f"FASTQ 2 x {f1.nbytes / 1e9:.2f} GB; blocks {sizes}", t0)
    assert len(ranges) == 4 and len(set(sizes)) > 1 and sum(sizes) == len(reads)
    jobs = [(f1.array[s:e], f2.array[s:e]) for s, e in ranges]
    blocks = gpu_ctx.fastq_batch(jobs, L, workers=4, streams=True, preserve_order=True, paired=True)
    log("batch done", t0)
    for i, (b, p) in enumerate(zip(blocks, parts)):
        n = b.n_reads
        assert n == len(p)
        placed = sum(len(b.get(f)) for f in ("hq_rec", "map_rec", "lq_rec", "nlq_rec"))
        assert placed == n, f"block {i}: {placed} records for {n} reads"
        mis = b.get("map_mis")
        assert int(mis.max(initial=0)) <= L // 5
        Nids = b.get("N_ids").astype(np.int64)
        by_gid = np.full(n, -1, dtype=np.int64)
        by_gid[gids(b.get("hq_rec"), Nids)] = 0
        by_gid[gids(b.get("map_rec"), Nids)] = mis
        srt = b.get("s_sorted_rec")
        assert np.array_equal(b.get("s_mis_count").astype(np.int64), by_gid[gids(srt, Nids)]), \
            f"block {i}: recount differs from the mapping counts"
        bs = archive.BlockStreams.from_block(b, n, L, True, True)
        got = gpu_ctx.decode(bs, L, True, True)
        assert np.array_equal(got, p), f"block {i}: rebuilt reads differ from the input"
        log(f"block {i}: {n} reads, {len(mis)} mapped, rebuilt exactly", t0)
    # block 0 against the OpenMP oracle, from the same FASTQ bytes
    s, e = ranges[0]
    o_reads = orc.parse_fastq(f1.array[s:e], L, f2.array[s:e], kind="omp")
    assert np.array_equal(o_reads, parts[0])
    ob = orc.OracleBlock(o_reads, stop="all", kind="omp")
    log("oracle block 0 done", t0)
    sorted_ids = ob.get("sorted_ids")
    n_prefix = int((ob.get("prev_dedup")[sorted_ids] == 0xFFFFFFFF).sum())
    assert n_prefix > 1 << 24, f"{n_prefix} prefixes: the hash table would have 2^24 slots"
    for f in FIELDS:
        assert np.array_equal(blocks[0].get(f), ob.get(f)), f"block 0: {f} differs from the oracle"
    ob.free()
    for b in blocks:
        b.free()
    f1.free()
    f2.free()
