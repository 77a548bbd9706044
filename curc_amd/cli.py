"""curc -c / -d command line (main.cpp:53-161).

    python -m curc_amd.cli -c -i reads.fastq [-i reads_2.fastq] -o name
                           [--block_ratio R] [--preserve_order] [--working_dir DIR]
    python -m curc_amd.cli -d -i name.curc [-o prefix] [--working_dir DIR]

Compression runs every block's pseudogenome path on the GPU (libcurc_gpu)
and writes name.curc; decompression writes name.seq, or name_1.seq and
name_2.seq for paired input.
"""
import argparse
import os
import sys
import time

from . import Context, archive, fastq


def compress(inputs, output_name, block_ratio=1.0, preserve_order=False, working_dir=".", device=0, log=print,
             io_chunk=fastq.IO_CHUNK):
    """All blocks of the input go to the GPU together (curc_batch_run: one
    stream per block, up to four at a time), where the reference runs one
    block task per block (preprocess.cpp:205-213)."""
    if not 1 <= len(inputs) <= 2:
        raise ValueError("one or two FASTQ inputs")
    paired = len(inputs) == 2
    out_path = os.path.join(working_dir, output_name + ".curc")
    blocks = []
    t0 = time.perf_counter()
    L = fastq.read_length(inputs[0])
    ranges = fastq.block_ranges(inputs[0], inputs[1] if paired else None, block_ratio, io_chunk)
    total = 0
    with Context(device) as ctx:
        # each block's FASTQ bytes in pinned host memory, parsed on the GPU
        jobs = []
        for s, e in ranges:
            bufs = []
            for path in inputs:
                hb = ctx.host_buffer(e - s)
                with open(path, "rb") as f:
                    f.seek(s)
                    if f.readinto(memoryview(hb.array)) != e - s:
                        raise fastq.FastqError("short read of " + path)
                bufs.append(hb)
            jobs.append(tuple(bufs) if paired else bufs[0])
        done = ctx.fastq_batch(jobs, L, streams=True, preserve_order=preserve_order, paired=paired, refmatch=True)
        for i, blk in enumerate(done):
            n = blk.n_reads
            total += n
            blocks.append(archive.BlockStreams.from_block(blk, n, L, preserve_order, paired))
            blk.free()
            log(f"block {i}: {n} reads")
        for j in jobs:
            for hb in (j if isinstance(j, tuple) else (j,)):
                hb.free()
    archive.write_archive(out_path, blocks, preserve_order, paired, L or 0)
    size = os.path.getsize(out_path)
    log(f"reads count : {total}; blocks : {len(blocks)}; archive size : {size} bytes; "
        f"{8.0 * size / max(1, total * (L or 1)):.4f} bits/base; {time.perf_counter() - t0:.1f} s")
    return out_path


def decompress(archive_path, output_prefix=None, working_dir=".", device=0):
    """Entropy decoding and the '%' expansion run on the host; each block's
    reads are rebuilt on the GPU (curc_decode_block)."""
    preserve, paired, L, blocks = archive.read_archive(archive_path)
    if output_prefix is None:
        output_prefix = os.path.splitext(os.path.basename(archive_path))[0]
    prefix = os.path.join(working_dir, output_prefix)
    import numpy as np
    with Context(device) as ctx:
        parts = [ctx.decode(b, L, preserve, paired) for b in blocks]
    reads = np.concatenate(parts) if parts else np.zeros((0, L), np.uint8)
    if paired:
        fastq.write_seq(prefix + "_1.seq", reads[0::2])
        fastq.write_seq(prefix + "_2.seq", reads[1::2])
        return [prefix + "_1.seq", prefix + "_2.seq"]
    fastq.write_seq(prefix + ".seq", reads)
    return [prefix + ".seq"]


def main(argv=None):
    p = argparse.ArgumentParser(prog="curc", description="CURC read compressor (MI355X pseudogenome path)")
    g = p.add_mutually_exclusive_group(required=True)
    g.add_argument("-c", "--compress", action="store_true")
    g.add_argument("-d", "--decompress", action="store_true")
    p.add_argument("-i", "--input", action="append", required=True)
    p.add_argument("-o", "--output")
    p.add_argument("--block_ratio", type=float, default=1.0)
    p.add_argument("--preserve_order", action="store_true")
    p.add_argument("--working_dir", default=".")
    a = p.parse_args(argv)
    if a.compress:
        for f in a.input:
            if not (f.endswith(".fastq") or f.endswith(".fq")):
                p.error("inputs must be .fastq or .fq")  # main.cpp:119-140
        if not 0 < a.block_ratio <= 1:
            p.error("block_ratio must be in (0, 1]")
        compress(a.input, a.output or os.path.splitext(os.path.basename(a.input[0]))[0], a.block_ratio,
                 a.preserve_order, a.working_dir)
    else:
        if len(a.input) != 1 or not a.input[0].endswith(".curc"):
            p.error("decompression takes one .curc archive")  # main.cpp:141-155
        decompress(a.input[0], a.output, a.working_dir)
    return 0


if __name__ == "__main__":
    sys.exit(main())
