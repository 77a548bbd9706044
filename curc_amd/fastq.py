"""FASTQ input: the reference's chunking and block cut (preprocess.cpp:66-243).

The reference reads an input in I/O chunks of 50 MB (25 MB per file when
paired), memory-mapped one at a time.  Every chunk but the last is cut back
to the start of a record by a backward scan (preprocess.cpp:114-138), and a
block closes after the chunk that takes it past block_ratio of the input
bytes (twice the file size when paired, :205-207).  So a block is a run of
whole chunks, which here is one byte range [start, end) of the file (the
same range of both files when paired: file 2 is mapped at file 1's offsets,
:98-101).  block_ranges() computes those ranges touching only the first
line and the tail of each chunk; the records themselves are parsed on the
GPU (curc_fastq_* in libcurc_gpu) or, for the ASCII interface, by
sequences() below.

Block-global read ids follow record order; paired reads interleave as 2k
(file 1) and 2k+1 (file 2).  Only the sequence line of each record is kept,
as in CURC, which compresses the DNA stream alone.
"""
import mmap
import os

import numpy as np

IO_CHUNK = 50 * 1000 * 1000
HEADER_PREFIX = b"synth."


class FastqError(ValueError):
    pass


_WINDOW = 1 << 16


def _cut(get, start, size):
    """Backward scan of preprocess.cpp:114-138 over the chunk [start,
    start+size) of an input (not its last chunk); get(a, b) returns the
    input's bytes [a, b).  Returns the chunk's new size.  Candidates are
    positions e (scanning down from size minus the first line's length) with
    a newline before and '@' at e; each candidate's bytes are compared with
    the chunk's first line, and the candidate is taken as soon as more than
    five comparisons in a row were equal.  The run counter is reset only by
    an unequal byte, not between candidates, exactly as in the reference.
    Positions without a candidate leave the counter alone, so only candidates
    are visited (found with rfind in windows read backwards)."""
    head, first_len, n = b"", -1, 256
    while first_len < 0 and len(head) < size:
        head = get(start, start + min(size, n))
        first_len = head.find(b"\n")
        n *= 4
    if head[:1] != b"@":
        raise FastqError("io_buffer doesn't start with @")  # preprocess.cpp:116
    if first_len < 0:
        raise FastqError("no line end in an I/O chunk")
    first = head[:first_len]
    e = size - first_len
    equal = 0
    win_lo, win = None, b""
    while e > 0:
        # the largest candidate at or below e: '\n' at e-1, '@' at e
        if win_lo is None or e - 1 < win_lo:
            win_lo = max(0, e + 1 - _WINDOW)
            win = get(start + win_lo, start + e + 1)
        k = win.rfind(b"\n@", 0, e + 1 - win_lo)
        if k < 0:
            if win_lo == 0:
                break
            e = win_lo  # a pair straddling the window's start is candidate e = win_lo
            win_lo = None
            continue
        e = win_lo + k + 1
        cand = get(start + e, start + e + first_len)
        for j in range(first_len):
            if first[j] == cand[j]:
                equal += 1
                if equal > 5:
                    return e
            else:
                equal = 0
        e -= 1
    # the reference keeps scanning with an empty chunk and never advances
    raise FastqError("no record start found in an I/O chunk (records longer than the chunk, "
                     "or headers without a common prefix of six bytes)")


class _Input:
    """A FASTQ input: a file path (memory-mapped) or a buffer in memory
    (bytes, bytearray or uint8 array)."""

    def __init__(self, src):
        self._f = self._mm = None
        if isinstance(src, (str, os.PathLike)):
            self.size = os.path.getsize(src)
            if self.size:
                self._f = open(src, "rb")
                self._mm = mmap.mmap(self._f.fileno(), 0, access=mmap.ACCESS_READ)
            self._buf = self._mm
        else:
            self._buf = memoryview(src).cast("B") if not isinstance(src, np.ndarray) else src
            self.size = len(self._buf)

    def get(self, a, b):
        return bytes(self._buf[a:b])

    def close(self):
        if self._mm is not None:
            self._mm.close()
        if self._f is not None:
            self._f.close()


def block_ranges(src1, src2=None, block_ratio=1.0, io_chunk=IO_CHUNK):
    """Byte ranges [(start, end), ...] of the blocks of an input (file paths
    or in-memory buffers), following process() (preprocess.cpp:72-243):
    chunks of io_chunk bytes (half of it per file when paired), each but the
    last cut back by _cut, blocks closed after the chunk that takes the
    block's bytes past block_bytes = fastq_bytes * block_ratio (both doubled
    when paired).  File 2 of a paired input is cut at file 1's offsets."""
    if not 0 < block_ratio <= 1:
        raise ValueError("block_ratio must be in (0, 1]")  # main.cpp:111-115
    paired = src2 is not None
    a = _Input(src1)
    try:
        size = a.size
        if paired:
            b = _Input(src2)
            size2 = b.size
            b.close()
            if size2 != size:
                raise FastqError("paired-end file size isn't equal")  # preprocess.cpp:83
        if size == 0:
            return []
        block_bytes = int((2.0 if paired else 1.0) * float(size) * block_ratio)
        chunk = io_chunk // 2 if paired else io_chunk
        out = []
        total = last = 0
        while total < size:
            start = total
            n = min(chunk, size - total)
            total += n
            if total < size:
                cut = _cut(a.get, start, n)
                total -= n - cut
            blk = total - last
            if paired:
                blk *= 2
            if blk > block_bytes:
                out.append((last, total))
                last = total
        if total - last > 0:
            out.append((last, total))
        return out
    finally:
        a.close()


def read_length(path):
    """Sequence length of the first record (compress(), preprocess.cpp:281-292)."""
    with open(path, "rb") as f:
        f.readline()
        return len(f.readline().rstrip(b"\n"))


def sequences(buf, L=None):
    """Sequence lines of the FASTQ records in buf (bytes-like or uint8
    array) -> (n, L) uint8, following the record loop of preprocess.cpp:147-
    193 (four lines per record, the second one kept)."""
    a = np.frombuffer(buf, dtype=np.uint8) if not isinstance(buf, np.ndarray) else buf
    if len(a) == 0:
        return np.zeros((0, L or 0), dtype=np.uint8)
    if a[0] != 64:
        raise FastqError("FASTQ block does not start with '@'")
    if a[-1] != 10:
        raise FastqError("truncated FASTQ record")
    nl = np.flatnonzero(a == 10)
    if len(nl) % 4:
        raise FastqError("truncated FASTQ record")
    s = nl[0::4] + 1
    e = nl[1::4]
    lens = e - s
    if L is None:
        L = int(lens[0])
    if np.any(lens != L):
        raise FastqError("read length isn't fix")  # preprocess.cpp:154
    if L == 0 or L > 512:
        raise FastqError("read length must be 1..512")  # preprocess.cpp:315-318
    return a[s[:, None] + np.arange(L)[None, :]]


def read_range(path, start, end):
    """Bytes [start, end) of a file."""
    with open(path, "rb") as f:
        f.seek(start)
        return f.read(end - start)


def read_blocks(path1, path2=None, block_ratio=1.0, io_chunk=IO_CHUNK):
    """Yields the blocks of a FASTQ input as (n, L) uint8 arrays (paired
    reads interleaved), cut as block_ranges() cuts them."""
    L = read_length(path1)
    for s, e in block_ranges(path1, path2, block_ratio, io_chunk):
        r1 = sequences(read_range(path1, s, e), L)
        if path2 is None:
            yield r1
            continue
        r2 = sequences(read_range(path2, s, e), L)
        if len(r2) != len(r1):
            raise FastqError("paired-end files hold different record counts in a block")
        r = np.empty((2 * len(r1), L), dtype=np.uint8)
        r[0::2], r[1::2] = r1, r2
        yield r


def fastq_bytes(reads, first_id=0, mate=None, prefix=HEADER_PREFIX):
    """FASTQ records of reads ((n, L) uint8) as one uint8 array: header
    '@<prefix><id, 10 digits>' (with '/1' or '/2' for a mate), the sequence,
    '+', and constant qualities 'I'.  Every record has the same length, and
    headers share their first bytes, which is what the reference's chunk cut
    looks for (preprocess.cpp:119-130)."""
    reads = np.ascontiguousarray(reads, dtype=np.uint8)
    n, L = reads.shape
    head = b"@" + prefix
    tail = b"" if mate is None else b"/%d" % mate
    hl = len(head) + 10 + len(tail) + 1
    rec = hl + L + 3 + L + 1
    out = np.empty((n, rec), dtype=np.uint8)
    out[:, :len(head)] = np.frombuffer(head, dtype=np.uint8)
    ids = np.arange(first_id, first_id + n, dtype=np.int64)
    for d in range(10):
        out[:, len(head) + 9 - d] = 48 + (ids // 10 ** d) % 10
    if tail:
        out[:, len(head) + 10:hl - 1] = np.frombuffer(tail, dtype=np.uint8)
    out[:, hl - 1] = 10
    out[:, hl:hl + L] = reads
    out[:, hl + L:hl + L + 3] = np.frombuffer(b"\n+\n", dtype=np.uint8)
    out[:, hl + L + 3:rec - 1] = 73
    out[:, rec - 1] = 10
    return out.reshape(-1)


def write_fastq(path, reads, mate=None):
    """Writes reads ((n, L) uint8) as FASTQ (fastq_bytes records)."""
    with open(path, "wb") as f:
        step = 1 << 20
        for i in range(0, len(reads), step):
            f.write(fastq_bytes(reads[i:i + step], first_id=i, mate=mate).tobytes())


def write_paired_fastq(path1, path2, reads):
    """Interleaved pairs (reads 2k, 2k+1) as two FASTQ files of equal size."""
    write_fastq(path1, reads[0::2], mate=1)
    write_fastq(path2, reads[1::2], mate=2)


def write_seq(path, reads):
    """CURC's decompressed form: one sequence per line (decompress.cu:1038-1043)."""
    n, L = reads.shape
    out = np.full((n, L + 1), 10, dtype=np.uint8)
    out[:, :L] = reads
    with open(path, "wb") as f:
        f.write(out.tobytes())
