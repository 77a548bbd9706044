"""The repeat-heavy paired-end configuration at small scale: a genome made
45% of diverged repeat copies (300 bp - 6 kbp, 5-15% divergence), 2 x 100 bp
pairs with fragment length ~N(400, 50), order preserved, cut by
block_ratio 0.25 into uneven blocks that run together in one batch.  Every
block equals the oracle, and the archive decodes to the input in order."""
import os
import tempfile

import numpy as np
import pytest

import curc_amd as ca
from curc_amd import archive, cli, fastq
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("hq_pg", "hq_rec", "map_rec", "map_mis", "unmapped", "unmapped_N", "lq_pg", "lq_rec", "nlq_pg",
          "nlq_rec")


def test_repeat_heavy_pairs_in_uneven_blocks():
    cfg = ca.synth_config(seed=61, genome_len=150_000, gc=0.45, repeat_frac=0.45, rep_min=300, rep_max=6000,
                          div_min=0.05, div_max=0.15, rep_families=12, n_reads=30_000, L=100, paired=1,
                          frag_mean=400, frag_sd=50, sub_first=0.005, sub_last=0.005, n_rate=0.0002)
    _, reads = ca.synth_reads(cfg)
    with tempfile.TemporaryDirectory() as d:
        f1, f2 = os.path.join(d, "r_1.fastq"), os.path.join(d, "r_2.fastq")
        fastq.write_fastq(f1, reads[0::2])
        fastq.write_fastq(f2, reads[1::2])
        # 400 kB read chunks: the cut rule yields uneven blocks of whole chunks
        chunk = 400_000
        parts = list(fastq.read_blocks(f1, f2, 0.25, chunk))
        sizes = [len(p) for p in parts]
        print("blocks", sizes, flush=True)
        assert len(parts) >= 3 and len(set(sizes)) > 1
        assert np.array_equal(np.concatenate(parts), reads)
        with ca.Context(0) as ctx:
            blocks = ctx.batch_run(parts, streams=True, preserve_order=True, paired=True, refmatch=True)
            for i, (p, b) in enumerate(zip(parts, blocks)):
                if i < 2:
                    o = orc.OracleBlock(p)
                    for f in FIELDS:
                        assert np.array_equal(b.get(f), o.get(f)), f"block {i}: {f} differs from the oracle"
                single = ctx.run_block(p, streams=True, preserve_order=True, paired=True, refmatch=True)
                for f in FIELDS + ("s_id_pos", "rm_hq_text", "rm_hq_off", "rm_hq_len"):
                    assert np.array_equal(b.get(f), single.get(f)), f"block {i}: {f} differs from a single run"
                single.free()
                b.free()
        out = cli.compress([f1, f2], "c3", block_ratio=0.25, preserve_order=True, working_dir=d, log=print,
                           io_chunk=chunk)
        dec = cli.decompress(out, working_dir=d)
        got1 = np.array([list(l.strip()) for l in open(dec[0], "rb")], dtype=np.uint8)
        got2 = np.array([list(l.strip()) for l in open(dec[1], "rb")], dtype=np.uint8)
        assert np.array_equal(got1, reads[0::2]) and np.array_equal(got2, reads[1::2])
        _, _, _, blks = archive.read_archive(out)
        assert len(blks) == len(parts)
