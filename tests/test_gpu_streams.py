"""GPU record sort and stream builders vs the oracle's restatement of
compress.cu:1730-2044, for the order-preserving and non-preserving layouts,
single- and paired-end."""
import numpy as np
import pytest

import curc_amd as ca
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

STREAMS = ("s_sorted_rec", "s_strand", "s_mis_count", "s_mis_base", "s_mis_table")


def check(ctx, reads, preserve, paired):
    L = reads.shape[1]
    max_mm = L // 5
    print(f"streams: {reads.shape[0]} reads, preserve={preserve}, paired={paired}", flush=True)
    g = ctx.run_block(reads, streams=True, preserve_order=preserve, paired=paired)
    o = orc.OracleBlock(reads, stop="all")
    o.streams(preserve, paired)
    fields = list(STREAMS)
    if preserve:
        fields.append("s_id_pos")
    else:
        fields += ["s_read_off", "s_lq_off", "s_nlq_off"]
        if paired:
            fields.append("s_pe_order")
    for f in fields:
        a, b = g.get(f), o.get(f)
        assert a.shape == b.shape and np.array_equal(a, b), f"{f} differs"
    for m in range(1, max_mm + 1):
        a, b = g.mis_off(m), o.mis_off(m)
        assert np.array_equal(a, b), f"mismatch offsets for m={m} differ"


@pytest.mark.parametrize("preserve", [False, True])
def test_streams_single_end(gpu_ctx, preserve):
    cfg = ca.synth_config(seed=61, genome_len=80_000, gc=0.45, n_reads=12_000, L=150,
                          sub_first=0.002, sub_last=0.02, n_rate=0.001)
    check(gpu_ctx, ca.synth_reads(cfg)[1], preserve, False)


@pytest.mark.parametrize("preserve", [False, True])
def test_streams_paired_end(gpu_ctx, preserve):
    cfg = ca.synth_config(seed=62, genome_len=60_000, gc=0.5, n_reads=10_000, L=100, paired=1,
                          sub_first=0.003, sub_last=0.01, n_rate=0.001)
    check(gpu_ctx, ca.synth_reads(cfg)[1], preserve, True)


def test_streams_config_small(gpu_ctx):
    check(gpu_ctx, ca.synth_reads(ca.config_small())[1], False, False)


def noisy_mix(seed=63):
    """Clean reads build the pseudogenome; noisy reads from the same genome
    map with many mismatches, exercising every offset-stream width."""
    clean = ca.synth_config(seed=seed, genome_len=30_000, n_reads=5_000, L=120, sub_first=0.001,
                            sub_last=0.002, n_rate=0.0005)
    genome, a = ca.synth_reads(clean)
    noisy = ca.synth_config(seed=seed + 1, genome_len=30_000, n_reads=2_000, L=120, sub_first=0.03,
                            sub_last=0.1, n_rate=0.003)
    _, b = ca.synth_reads(noisy, genome)
    return np.concatenate([a, b])


@pytest.mark.parametrize("preserve", [False, True])
def test_streams_high_error(gpu_ctx, preserve):
    check(gpu_ctx, noisy_mix(), preserve, False)
