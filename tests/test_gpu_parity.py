"""GPU path vs the CPU oracle, field by field, on seeded synthetic blocks.

Every comparison is exact (integer/byte work).  Intermediate link arrays are
compared stage by stage, so a failure names the first stage that diverges.
"""
import os

import numpy as np
import pytest

import curc_amd as ca
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

STAGE_FIELDS = [
    ("nonN_ids", "N_ids", "sorted_ids", "next_dedup", "prev_dedup"),
    ("next_overlap", "prev_overlap", "offset_overlap"),
    ("next_prune", "prev_prune", "hq_ids", "unmatched_ids"),
    ("next", "prev", "offset"),
    ("hq_pg", "hq_rec", "hq_pg2"),
    ("map_rec", "map_mis", "unmapped", "unmapped_N"),
    ("lq_next", "lq_prev", "lq_offset", "lq_pg", "lq_rec"),
    ("nlq_pg", "nlq_rec"),
]


def compare(gpu_blk, orc_blk, fields):
    for f in fields:
        a, b = gpu_blk.get(f), orc_blk.get(f)
        assert a.shape == b.shape, f"{f}: length {a.shape} vs oracle {b.shape}"
        if not np.array_equal(a, b):
            bad = np.flatnonzero(a != b)
            i = int(bad[0])
            raise AssertionError(f"{f}: {len(bad)} entries differ; first at {i}: gpu {a[i]} oracle {b[i]}")


def run_both(ctx, reads):
    print(f"case: {reads.shape[0]} reads x {reads.shape[1]} bp", flush=True)
    g = ctx.run_block(reads, keep_stages=True)
    o = orc.OracleBlock(reads, stop="all")
    for group in STAGE_FIELDS:
        compare(g, o, group)
    packed = g.get("packed")
    assert np.array_equal(packed, o.get("packed")), "packed reads differ"
    return g, o


def test_small_block_l150(gpu_ctx):
    cfg = ca.synth_config(seed=101, genome_len=50_000, gc=0.5, n_reads=8_000, L=150,
                          sub_first=0.003, sub_last=0.008, n_rate=0.0005)
    run_both(gpu_ctx, ca.synth_reads(cfg)[1])


def test_config_small(gpu_ctx):
    """The 1 Mbp / 133,333-read / 150 bp workload."""
    run_both(gpu_ctx, ca.synth_reads(ca.config_small())[1])


@pytest.mark.parametrize("L", [20, 24, 33, 40, 64, 100, 150])
def test_dedup_order_of_reads_sharing_their_first_bases(gpu_ctx, L):
    """The dedup sort orders reads by their first 16 bases and then the runs
    of reads that share them by everything after: reads here come from 6
    prefixes of 16 bases with tails from 40 variants (point changes of one
    tail spread over every word, so ties reach every word), many exact
    duplicates, and N-reads.  L = 20 and 24 end inside the first word, and
    L = 33 ends one base into the second."""
    rng = np.random.default_rng(900 + L)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    heads = acgt[rng.integers(0, 4, (6, 16))]
    base = acgt[rng.integers(0, 4, L - 16)]
    tails = np.repeat(base[None], 40, 0)
    for t in range(1, 40):
        j = rng.integers(0, L - 16, 1 + t % 3)
        tails[t, j] = acgt[rng.integers(0, 4, len(j))]
    n = 20_000
    reads = np.concatenate([heads[rng.integers(0, 6, n)], tails[rng.integers(0, 40, n)]], axis=1)
    nr = rng.choice(n, 300, replace=False)
    reads[nr, rng.integers(0, L, 300)] = ord("N")
    g = gpu_ctx.run_block(reads, stop_after=ca.STAGE_DEDUP, keep_stages=True)
    o = orc.OracleBlock(reads, stop="dedup")
    compare(g, o, STAGE_FIELDS[0])
    g.free()


@pytest.mark.parametrize("L", [75, 100, 250, 300])
def test_read_lengths(gpu_ctx, L):
    cfg = ca.synth_config(seed=200 + L, genome_len=40_000, gc=0.48, n_reads=6_000, L=L,
                          sub_first=0.002, sub_last=0.01, n_rate=0.0003)
    run_both(gpu_ctx, ca.synth_reads(cfg)[1])


@pytest.mark.parametrize("form", [2, 3])
@pytest.mark.parametrize("L", [60, 75, 150, 250])
def test_overlap_filter_verdicts(L, form):
    """Overlap probe forms 2 and 3: minimizer-blocked k-mer filter, verdicts
    for all offsets computed in one pass; form 3 turns them into one list per
    offset of the suffixes that pass.  Same links as the oracle."""
    cfg = ca.synth_config(seed=600 + L, genome_len=40_000, gc=0.46, repeat_frac=0.2, rep_min=200, rep_max=900,
                          rep_families=4, div_min=0.0, div_max=0.02, n_reads=7_000, L=L,
                          sub_first=0.002, sub_last=0.01, n_rate=0.0003)
    with ca.Context(0) as ctx:
        ctx.set_tuning(ov_queue=form)
        run_both(ctx, ca.synth_reads(cfg)[1])


def test_overlap_forms_count_the_same_work():
    """Overlap probe forms 2 and 3 search the same suffixes at every offset
    (verdict passed, not linked yet): their work counts (probes and bucket
    searches, in the stage's algorithmic bytes) are equal."""
    cfg = ca.synth_config(seed=611, genome_len=300_000, gc=0.45, n_reads=60_000, L=150, sub_first=0.002,
                          sub_last=0.01, n_rate=0.0003)
    reads = ca.synth_reads(cfg)[1]
    got = {}
    with ca.Context(0) as ctx:
        for form in (2, 3):
            ctx.set_tuning(ov_queue=form)
            ctx.reset_stats()
            ctx.set_timing(True)
            ctx.run_block(reads).free()
            ctx.set_timing(False)
            st = {s["name"]: s for s in ctx.stats()}
            got[form] = (st["overlap"]["items"], st["overlap"]["alg_bytes"])
    print(got, flush=True)
    assert got[2] == got[3] and got[2][1] > 32 * got[2][0]


@pytest.mark.parametrize("case", ["plain", "repeats"])
def test_mapping_both_strands_one_launch(case):
    """Tuning map_both=1: both strand passes in one launch on the two-strand index of
    128-byte lines (taken by blocks whose pseudogenome fills bucket-inline slots,
    so 6 and 9 Mbp genomes at 30x here); map_both=2 (the default): the same on the
    compact index of 32-byte slots, whose larger buckets keep their positions
    in the overflow array (the repeat case's index has about 37,500 buckets of
    more than 7 positions, up to 24).  Same results as the oracle (OpenMP build)
    and as the forward pass followed by the reverse-complement pass
    (map_both=0, one launch per strand); both one-launch forms end a strand
    pass early once no unseen window can beat it (k_map_both), so they do the
    same, smaller work as each other."""
    if case == "plain":
        cfg = ca.config_scale(seed=81, n_reads=1_200_000, genome_len=6_000_000)
    else:
        cfg = ca.synth_config(seed=82, genome_len=9_000_000, gc=0.45, repeat_frac=0.3, rep_min=300, rep_max=3000,
                              rep_families=12, div_min=0.0, div_max=0.03, n_reads=1_800_000, L=150,
                              sub_first=0.002, sub_last=0.02, n_rate=0.001)
    reads = ca.synth_reads(cfg)[1]
    print(f"case {case}: {reads.shape}", flush=True)
    names = ("map_active", "map_probes", "map_cands", "map_first_probe", "map_open_probes", "map_open_cands")

    def work(ctx):
        ctx.reset_stats()
        ctx.set_timing(True)
        blk = ctx.run_block(reads)
        ctx.set_timing(False)
        st = {s["name"]: s for s in ctx.stats()}
        return blk, {k: st[k]["items"] for k in names}, st

    with ca.Context(0) as ctx:
        ctx.set_tuning(map_both=0)
        b0, separate, st0 = work(ctx)
        ctx.set_tuning(map_both=1)
        b1, together, st1 = work(ctx)
        ctx.set_tuning(map_both=2)
        b2, compact, st2 = work(ctx)
        assert st1["mapping"]["launches"] == 1 and st0["mapping"]["launches"] == 2, "the combined path did not run"
        assert st2["mapping"]["launches"] == 1
        assert together == compact
        print("work per strand launches", separate, "one launch", compact, flush=True)
        assert compact["map_probes"] < separate["map_probes"] and compact["map_cands"] <= separate["map_cands"]
        o = orc.OracleBlock(reads, stop="all", kind="omp")
        fields = ("hq_pg", "hq_rec", "map_rec", "map_mis", "unmapped", "unmapped_N", "lq_pg", "lq_rec", "nlq_pg",
                  "nlq_rec")
        compare(b0, o, fields)
        compare(b1, o, fields)
        compare(b2, o, fields)
        b0.free()
        b1.free()
        b2.free()


def test_high_coverage_duplicates(gpu_ctx):
    """Tiny genome at ~1500x: many exact duplicates and contended claims."""
    cfg = ca.synth_config(seed=7, genome_len=2_000, gc=0.5, n_reads=20_000, L=100,
                          sub_first=0.0, sub_last=0.002, n_rate=0.0)
    run_both(gpu_ctx, ca.synth_reads(cfg)[1])


def test_repeats(gpu_ctx):
    cfg = ca.synth_config(seed=9, genome_len=200_000, gc=0.45, repeat_frac=0.45, rep_min=300,
                          rep_max=3000, rep_families=8, div_min=0.0, div_max=0.02,
                          n_reads=30_000, L=100, sub_first=0.002, sub_last=0.004, n_rate=0.0002)
    run_both(gpu_ctx, ca.synth_reads(cfg)[1])


def test_periodic_sequence(gpu_ctx):
    """Reads from a short-period sequence: self-overlaps, cycles, self-avoidance."""
    rng = np.random.default_rng(3)
    unit = np.frombuffer(b"ACGTTGCAAC", dtype=np.uint8)
    genome = np.tile(unit, 200)
    n, L = 600, 60
    starts = rng.integers(0, len(genome) - L, size=n)
    reads = np.stack([genome[s:s + L] for s in starts]).copy()
    flip = rng.random((n, L)) < 0.01
    reads[flip] = np.frombuffer(b"G", dtype=np.uint8)[0]
    run_both(gpu_ctx, reads)


def test_n_heavy(gpu_ctx):
    cfg = ca.synth_config(seed=31, genome_len=30_000, gc=0.5, n_reads=5_000, L=120,
                          sub_first=0.004, sub_last=0.004, n_rate=0.01)
    run_both(gpu_ctx, ca.synth_reads(cfg)[1])


def test_tiny_blocks(gpu_ctx):
    for n in (1, 2, 3, 17):
        cfg = ca.synth_config(seed=40 + n, genome_len=1_000, n_reads=n, L=90, n_rate=0.0)
        run_both(gpu_ctx, ca.synth_reads(cfg)[1])
    # all reads identical
    reads = np.tile(np.frombuffer(b"ACGT" * 25, dtype=np.uint8), (50, 1))
    run_both(gpu_ctx, reads)


def test_determinism(gpu_ctx):
    cfg = ca.synth_config(seed=55, genome_len=30_000, n_reads=6_000, L=150)
    reads = ca.synth_reads(cfg)[1]
    a = gpu_ctx.run_block(reads)
    b = gpu_ctx.run_block(reads)
    for f in ("next", "prev", "offset", "hq_pg", "map_rec", "map_mis", "lq_pg", "lq_rec"):
        assert np.array_equal(a.get(f), b.get(f)), f


def test_csr_hash_indexes():
    """The CSR prefix and mapping indexes (taken when bucket slots would pass
    16 GiB; forced here by CURC_INDEX=csr) give the same results."""
    cfg = ca.synth_config(seed=303, genome_len=60_000, gc=0.45, n_reads=9_000, L=150,
                          sub_first=0.002, sub_last=0.012, n_rate=0.0005)
    reads = ca.synth_reads(cfg)[1]

    def work(ctx):
        # mapping work counters: the slot index runs the register-resident
        # kernel (one launch per strand, map_both=0), the CSR index the
        # generic one; both walk the same probes and verify the same
        # candidates
        ctx.set_tuning(map_both=0)
        ctx.reset_stats()
        ctx.set_timing(True)
        ctx.run_block(reads).free()
        ctx.set_timing(False)
        st = {s["name"]: s["items"] for s in ctx.stats()}
        return {k: st[k] for k in ("map_active", "map_probes", "map_cands", "map_first_probe", "map_open_probes",
                                   "map_open_cands")}

    with ca.Context(0) as ctx:
        fast = work(ctx)
    old = os.environ.get("CURC_INDEX")
    os.environ["CURC_INDEX"] = "csr"
    try:
        with ca.Context(0) as ctx:
            run_both(ctx, reads)
            assert work(ctx) == fast
    finally:
        if old is None:
            del os.environ["CURC_INDEX"]
        else:
            os.environ["CURC_INDEX"] = old
