"""Exact all-offset matching (compress.cu:1003-1174 for HQ and LQ reads,
:1176-1347 for N-reads) with its lists kept in place (tuning fm_static=1:
linked suffixes skipped by flag, linked prefixes by skip pointer, suffixes
screened by the prefix filter at offsets that leave 32 or more bases) and
with both lists compacted after every offset (fm_static=0), against the
oracle.  The blocks are large enough for HQ and LQ lists above the
short-list bound (block ranges, filter verdicts); the repeat-heavy one gives
groups of many suffixes with one overlap.  Exact comparison."""
import numpy as np
import pytest

import curc_amd as ca
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("next", "prev", "offset", "hq_pg", "hq_rec", "lq_next", "lq_prev", "lq_offset", "lq_pg", "lq_rec",
          "nlq_pg", "nlq_rec")


def case_reads(case):
    if case == "plain":
        cfg = ca.config_scale(seed=61, n_reads=300_000, genome_len=1_500_000)
    elif case == "noisy":
        # many unmapped reads: LQ lists above the short-list bound
        cfg = ca.synth_config(seed=62, genome_len=800_000, gc=0.45, n_reads=200_000, L=100, sub_first=0.01,
                              sub_last=0.05, n_rate=0.002)
    else:
        cfg = ca.synth_config(seed=63, genome_len=1_000_000, gc=0.4, repeat_frac=0.45, rep_min=300, rep_max=3000,
                              rep_families=6, div_min=0.0, div_max=0.01, n_reads=250_000, L=150, sub_first=0.001,
                              sub_last=0.01, n_rate=0.0005)
    return ca.synth_reads(cfg)[1]


@pytest.mark.timeout(600)
@pytest.mark.parametrize("case", ["plain", "noisy", "repeats"])
def test_exact_matching_lists_in_place(case):
    reads = case_reads(case)
    print(f"case {case}: {reads.shape[0]} reads x {reads.shape[1]} bp", flush=True)
    o = orc.OracleBlock(reads, stop="all", kind="omp")
    with ca.Context(0) as ctx:
        ctx.set_timing(True)
        for mode in (1, 0):
            ctx.set_tuning(fm_static=mode)
            ctx.reset_stats()
            b = ctx.run_block(reads, keep_stages=True)
            for f in FIELDS:
                a, r = b.get(f), o.get(f)
                assert a.shape == r.shape, f"fm_static={mode}: {f}: length {a.shape} vs oracle {r.shape}"
                if not np.array_equal(a, r):
                    bad = np.flatnonzero(a != r)
                    raise AssertionError(f"fm_static={mode}: {f}: {len(bad)} entries differ, first at {bad[0]}")
            st = {s["name"]: s for s in ctx.stats()}
            print(f"  fm_static={mode}: full_match {st['full_match']['ms']:.2f} ms "
                  f"({st['full_match']['launches']} launches), lq {st['lq']['ms']:.2f} ms", flush=True)
            # the loop over lists above the short-list bound counts its bytes
            if st["full_match"]["launches"] > 1:
                assert st["full_match"]["alg_bytes"] > 0
            b.free()
        n_hq = int(np.count_nonzero(o.get("next") != 0xFFFFFFFF))
        assert n_hq > 0
