"""GPU read reconstruction (curc_decode_block, decompress.cu:909-1005 and
472-577) against the decoder restated in test_oracle_streams.py, the host
decoder of the archive and the input reads: order not preserved and
preserved, single-end and paired, from blocks and from archives read back."""
import os
import tempfile

import numpy as np
import pytest

import curc_amd as ca
from curc_amd import archive
from test_oracle_streams import noisy_mix, rebuild

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("preserve", [False, True])
def test_decode_matches_restated_decoder(gpu_ctx, preserve):
    reads = noisy_mix()
    L = reads.shape[1]
    blk = gpu_ctx.run_block(reads, streams=True, preserve_order=preserve)
    bs = archive.BlockStreams.from_block(blk, len(reads), L, preserve, False)
    got = gpu_ctx.decode(bs, L, preserve, False)
    want = rebuild(blk, reads, preserve)
    assert max(m for m in range(1, L // 5 + 1) if len(bs.mis_off[m])) >= 8
    if preserve:
        assert np.array_equal(got, reads)
    else:
        assert [g.tobytes().decode() for g in got] == [s for _, s in want]
    blk.free()


def archive_round_trip(ctx, reads, preserve, paired):
    L = reads.shape[1]
    blk = ctx.run_block(reads, streams=True, preserve_order=preserve, paired=paired, refmatch=True)
    bs = archive.BlockStreams.from_block(blk, len(reads), L, preserve, paired)
    blk.free()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "x.curc")
        archive.write_archive(p, [bs], preserve, paired, L)
        _, _, _, blocks = archive.read_archive(p)
    return blocks[0]


@pytest.mark.parametrize("preserve,paired", [(False, False), (True, False), (False, True), (True, True)])
def test_decode_archive_round_trip(gpu_ctx, preserve, paired):
    cfg = ca.synth_config(seed=77, genome_len=150_000, gc=0.45, n_reads=24_000, L=100, paired=int(paired),
                          rc_prob=0.5, sub_first=0.003, sub_last=0.01, n_rate=0.0005)
    reads = ca.synth_reads(cfg)[1]
    print(f"archive round trip preserve={preserve} paired={paired}", flush=True)
    bs = archive_round_trip(gpu_ctx, reads, preserve, paired)
    got = gpu_ctx.decode(bs, 100, preserve, paired)
    host = archive.decode_block(bs, 100, preserve, paired)
    assert np.array_equal(got, host)
    if preserve:
        assert np.array_equal(got, reads)
    else:
        assert sorted(map(bytes, got)) == sorted(map(bytes, reads))
        if paired:
            # pairs stay together: rows 2k, 2k+1 are the two mates of one input pair
            pairs_in = sorted(zip(map(bytes, reads[0::2]), map(bytes, reads[1::2])))
            pairs_out = sorted(zip(map(bytes, got[0::2]), map(bytes, got[1::2])))
            assert pairs_in == pairs_out


def test_decode_rejects_corrupt_streams(gpu_ctx):
    reads = noisy_mix()
    L = reads.shape[1]
    blk = gpu_ctx.run_block(reads, streams=True)
    bs = archive.BlockStreams.from_block(blk, len(reads), L, False, False)
    blk.free()
    m = max(k for k in bs.mis_off if len(bs.mis_off[k]))
    good = bs.mis_off[m]
    bs.mis_off[m] = good[:-1]
    with pytest.raises(ca.CurcError, match="corrupt"):
        gpu_ctx.decode(bs, L, False, False)
    bs.mis_off[m] = good
    bs.read_off = bs.read_off.copy()
    bs.read_off[-1] = 60000
    with pytest.raises(ca.CurcError, match="corrupt"):
        gpu_ctx.decode(bs, L, False, False)
    bs.mis_base = bs.mis_base[:-1]
    with pytest.raises(ca.CurcError, match="mismatch symbols"):
        gpu_ctx.decode(bs, L, False, False)
