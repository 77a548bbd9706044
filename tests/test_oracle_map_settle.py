"""The early end of a strand pass in the mapping (k_map_both; the oracle's
map_settled, off unless a test turns it on): once every class of probes
(t mod 5) has had more than 3 (B - 1) probes of complete buckets, no window
the pass has not seen can have fewer than B mismatches, so the rest of the
pass cannot change its result.  Here the oracle's sequential loop runs with
and without the rule on blocks chosen to stress it (many reads with 2-4
mismatches, full buckets from repeats, N-reads, short and long reads), and
every mapping result and everything built from it must be equal; the rule
must actually fire.  A weaker rule (a mismatch taken to spoil 1 or 2
probes of a class instead of 3) must change some result, so the cases can
tell.  CPU only."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import curc_amd as ca  # noqa: E402
from oracle import oracle as orc  # noqa: E402

FIELDS = ("map_rec", "map_mis", "unmapped", "unmapped_N", "lq_pg", "lq_rec", "nlq_pg", "nlq_rec")

CASES = {
    "errors_150": dict(seed=501, genome_len=120_000, gc=0.45, n_reads=24_000, L=150, sub_first=0.004,
                       sub_last=0.03, n_rate=0.0005),
    "repeats_100": dict(seed=502, genome_len=150_000, gc=0.45, repeat_frac=0.5, rep_min=200, rep_max=2000,
                        rep_families=6, div_min=0.0, div_max=0.02, n_reads=25_000, L=100, sub_first=0.005,
                        sub_last=0.02, n_rate=0.0005),
    "n_heavy_120": dict(seed=503, genome_len=80_000, gc=0.5, n_reads=12_000, L=120, sub_first=0.005,
                        sub_last=0.02, n_rate=0.01),
    "long_250": dict(seed=504, genome_len=200_000, gc=0.42, n_reads=16_000, L=250, sub_first=0.003,
                     sub_last=0.02, n_rate=0.0005),
    "short_75": dict(seed=505, genome_len=60_000, gc=0.47, n_reads=16_000, L=75, sub_first=0.01,
                     sub_last=0.03, n_rate=0.0005),
    "dense_2k": dict(seed=506, genome_len=3_000, gc=0.5, n_reads=6_000, L=100, sub_first=0.01, sub_last=0.04,
                     n_rate=0.002),
}


def run(reads, kill, both=False, tie=False):
    orc.lib().orc_set_map_settle(kill)
    orc.lib().orc_set_map_both(int(both), int(tie))
    try:
        b = orc.OracleBlock(reads, stop="all")
        out = {f: b.get(f).copy() for f in FIELDS}
        return out, int(orc.lib().orc_map_settled_count())
    finally:
        orc.lib().orc_set_map_settle(0)
        orc.lib().orc_set_map_both(0, 0)


def differs(a, b):
    return [f for f in FIELDS if a[f].shape != b[f].shape or not np.array_equal(a[f], b[f])]


@pytest.fixture(scope="module")
def blocks():
    out = {}
    for name, kw in CASES.items():
        reads = ca.synth_reads(ca.synth_config(**kw))[1]
        out[name] = (reads, run(reads, 0)[0])
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_rule_keeps_every_result(blocks, name):
    reads, plain = blocks[name]
    got, fired = run(reads, 3)
    print(name, "passes ended early:", fired, flush=True)
    assert not differs(plain, got), differs(plain, got)
    if reads.shape[1] >= 100:
        assert fired > 0


def test_weaker_rule_is_caught(blocks):
    caught = {k: [] for k in (1, 2)}
    for name, (reads, plain) in blocks.items():
        for kill in (1, 2):
            got, _ = run(reads, kill)
            if differs(plain, got):
                caught[kill].append(name)
    print(caught, flush=True)
    assert caught[1] and caught[2]


@pytest.mark.parametrize("name", list(CASES))
def test_interleaved_passes_with_the_rule_keep_every_result(blocks, name):
    """Both strand passes interleaved probe by probe (as k_map_both runs
    them), without and with the early end of each pass; the forward pass
    then also ends once no unseen forward window can reach the reverse
    best (B = min(forward best, reverse best + 1))."""
    reads, plain = blocks[name]
    got, _ = run(reads, 0, both=True)
    assert not differs(plain, got), differs(plain, got)
    got, fired = run(reads, 3, both=True)
    print(name, "passes ended early (interleaved):", fired, flush=True)
    assert not differs(plain, got), differs(plain, got)
    assert fired > 0


def test_interleaved_mutants_are_caught(blocks):
    """Without the + 1 (a forward window equal to the reverse best would win
    the tie and is skipped) or with a weaker rule, some result changes."""
    caught = {"tie": [], "kill2": []}
    for name, (reads, plain) in blocks.items():
        if differs(plain, run(reads, 3, both=True, tie=True)[0]):
            caught["tie"].append(name)
        if differs(plain, run(reads, 2, both=True)[0]):
            caught["kill2"].append(name)
    print(caught, flush=True)
    assert caught["tie"] and caught["kill2"]
