"""Seeded synthetic inputs of the headline configurations as FASTQ bytes in
pinned host memory (what a caller hands the FASTQ entry points after reading
its files)."""
import numpy as np

from . import config_scale, synth_config, synth_fastq, synth_fastq_record_len, synth_reads


def config_repeat_pe(seed=3003, genome_len=500_000_000, coverage=20):
    """Repeat-heavy paired-end configuration: a 500 Mbp genome, 45% of it
    copies of 300 bp - 6 kbp repeat elements at 5-15% divergence, 2 x 100 bp
    pairs with fragment length ~N(400, 50), 20x coverage (100M reads), 0.5%
    substitutions."""
    n = int(genome_len * coverage // 100) // 2 * 2
    return synth_config(seed=seed, genome_len=genome_len, gc=0.45, repeat_frac=0.45, rep_min=300, rep_max=6000,
                        div_min=0.05, div_max=0.15, rep_families=64, n_reads=n, L=100, paired=1, frag_mean=400,
                        frag_sd=50, sub_first=0.005, sub_last=0.005, n_rate=0.0002)


def pinned_fastq(ctx, reads, paired):
    """FASTQ records of reads ((n, L) uint8; pairs interleaved when paired)
    in pinned host buffers: one HostBuffer, or (file 1, file 2)."""
    L = reads.shape[1]
    if not paired:
        hb = ctx.host_buffer(len(reads) * synth_fastq_record_len(L))
        synth_fastq(reads, out=hb.array)
        return hb
    bufs = []
    for mate in (1, 2):
        hb = ctx.host_buffer(len(reads) // 2 * synth_fastq_record_len(L, mate))
        synth_fastq(reads[mate - 1:], out=hb.array, mate=mate, stride_reads=2)
        bufs.append(hb)
    return tuple(bufs)


def scale_fastq(ctx, **kw):
    """The 100 Mbp / 20M x 150 bp workload as one pinned FASTQ buffer."""
    _, reads = synth_reads(config_scale(**kw))
    return pinned_fastq(ctx, reads, False)


def block_reads(reads, ranges, rec_len, paired):
    """Rows of reads belonging to each block's byte range (records have a
    fixed length rec_len per file)."""
    out = []
    for s, e in ranges:
        a, b = s // rec_len, e // rec_len
        out.append(reads[2 * a:2 * b] if paired else reads[a:b])
    return out
