"""FASTQ reading and block cutting (preprocess.cpp:66-319).

Reads are taken in I/O chunks of 50 MB (25 MB per file when paired); a block
closes after the chunk that takes the block past block_ratio of the input
bytes (twice the file size when paired), so blocks hold whole chunks.  Block-
global read ids follow record order; paired reads interleave as 2k (file 1)
and 2k+1 (file 2).  Only the sequence line of each record is kept, as in
CURC, which compresses the DNA stream alone.
"""
import os

import numpy as np

IO_CHUNK = 50 * 1000 * 1000


class FastqError(ValueError):
    pass


def _sequences(buf: bytes, L=None):
    """Sequence lines of the FASTQ records in buf -> (n, L) uint8."""
    a = np.frombuffer(buf, dtype=np.uint8)
    nl = np.flatnonzero(a == 10)
    if len(nl) % 4:
        raise FastqError("truncated FASTQ record")
    starts = np.concatenate([[0], nl[:-1] + 1])
    if len(starts) and buf[0:1] != b"@":
        raise FastqError("FASTQ record does not start with '@'")
    s = starts[1::4]
    e = nl[1::4]
    lens = e - s
    if len(lens) == 0:
        return np.zeros((0, L or 0), dtype=np.uint8)
    if L is None:
        L = int(lens[0])
    if np.any(lens != L):
        raise FastqError("read length isn't fix")  # preprocess.cpp:154
    if L > 512:
        raise FastqError("read length must be at most 512")  # preprocess.cpp:315-318
    idx = s[:, None] + np.arange(L)[None, :]
    return a[idx]


def _chunks(path, chunk):
    """Byte chunks of a FASTQ file, each ending at a record boundary."""
    with open(path, "rb") as f:
        data = f.read()
    n = len(data)
    pos = 0
    while pos < n:
        end = min(n, pos + chunk)
        if end < n:
            # back up to the last complete record (four newlines)
            cut = end
            k = data.rfind(b"\n@", pos, cut)
            while k != -1:
                # a record start is a line starting with '@' whose line index is a multiple of 4
                if data.count(b"\n", pos, k + 1) % 4 == 0:
                    end = k + 1
                    break
                k = data.rfind(b"\n@", pos, k)
            if k == -1:
                end = n
        yield data[pos:end]
        pos = end


def read_blocks(path1, path2=None, block_ratio=1.0, io_chunk=IO_CHUNK):
    """Yields the blocks of a FASTQ input as (n, L) uint8 arrays.  io_chunk
    is the reference's 50 MB read granularity (smaller values cut small
    inputs into several blocks)."""
    if not 0 < block_ratio <= 1:
        raise ValueError("block_ratio must be in (0, 1]")  # main.cpp:111-115
    paired = path2 is not None
    size1 = os.path.getsize(path1)
    if paired and os.path.getsize(path2) != size1:
        raise FastqError("paired-end file size isn't equal")  # preprocess.cpp:83
    block_bytes = (2 if paired else 1) * size1 * block_ratio
    chunk = io_chunk // 2 if paired else io_chunk
    it1 = _chunks(path1, chunk)
    it2 = _chunks(path2, chunk) if paired else None
    L = None
    cur, cur_bytes = [], 0
    for c1 in it1:
        r1 = _sequences(c1, L)
        L = r1.shape[1] if L is None else L
        if paired:
            c2 = next(it2)
            r2 = _sequences(c2, L)
            if len(r2) != len(r1):
                raise FastqError("paired files have different record counts in a chunk")
            r = np.empty((2 * len(r1), L), dtype=np.uint8)
            r[0::2], r[1::2] = r1, r2
            cur_bytes += len(c1) + len(c2)
        else:
            r = r1
            cur_bytes += len(c1)
        cur.append(r)
        if cur_bytes > block_bytes:
            yield np.concatenate(cur)
            cur, cur_bytes = [], 0
    if cur:
        yield np.concatenate(cur)


def write_fastq(path, reads, name="r"):
    """Writes reads ((n, L) uint8) as FASTQ with constant qualities."""
    n, L = reads.shape
    qual = b"I" * L
    with open(path, "wb") as f:
        for i in range(n):
            f.write(b"@%s%d\n" % (name.encode(), i))
            f.write(reads[i].tobytes())
            f.write(b"\n+\n")
            f.write(qual)
            f.write(b"\n")


def write_seq(path, reads):
    """CURC's decompressed form: one sequence per line (decompress.cu:1038-1043)."""
    n, L = reads.shape
    out = np.full((n, L + 1), 10, dtype=np.uint8)
    out[:, :L] = reads
    with open(path, "wb") as f:
        f.write(out.tobytes())
