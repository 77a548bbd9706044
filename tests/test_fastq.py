"""FASTQ chunking and block cut (preprocess.cpp:72-243) against a literal
restatement of the reference loop (tests/fastq_ref.py), on inputs whose
records straddle chunk ends, single-end and paired, block_ratio 0.25 / 0.5 /
1.  Also the record parse (fastq.sequences) and the oracle's parser against
the same restatement.  CPU only."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from curc_amd import fastq  # noqa: E402
from fastq_ref import process_blocks  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def records(reads, header, qual=None):
    """FASTQ text of reads ((n, L) uint8) with header(k) -> bytes (without
    '@') and qual(k) -> bytes (default constant 'I')."""
    out = bytearray()
    L = reads.shape[1]
    for k in range(len(reads)):
        out += b"@" + header(k) + b"\n" + reads[k].tobytes() + b"\n+\n"
        out += (qual(k) if qual else b"I" * L) + b"\n"
    return bytes(out)


def random_reads(n, L, seed, n_rate=0.01):
    rng = np.random.default_rng(seed)
    r = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, size=(n, L))]
    r[rng.random((n, L)) < n_rate] = ord("N")
    return r


def check_against_reference(d1, d2, ratio, chunk, tmp_path):
    p1 = tmp_path / "a_1.fastq"
    p1.write_bytes(d1)
    p2 = None
    if d2 is not None:
        p2 = tmp_path / "a_2.fastq"
        p2.write_bytes(d2)
    ref_blocks, ref_ranges = process_blocks(d1, d2, ratio, chunk)
    ranges = fastq.block_ranges(str(p1), None if p2 is None else str(p2), ratio, chunk)
    assert ranges == ref_ranges
    got = list(fastq.read_blocks(str(p1), None if p2 is None else str(p2), ratio, chunk))
    assert [len(b) for b in got] == [len(b) for b in ref_blocks]
    for g, r in zip(got, ref_blocks):
        assert [bytes(x) for x in g] == r
    return ref_ranges


@pytest.mark.parametrize("paired", [False, True])
@pytest.mark.parametrize("ratio", [0.25, 0.5, 1.0])
@pytest.mark.parametrize("chunk", [1500, 2333, 4000])
def test_block_cut_matches_reference_loop(tmp_path, paired, ratio, chunk):
    L = 60
    n = 90
    r1 = random_reads(n, L, seed=chunk + int(ratio * 10))
    # header lengths vary (ids without padding), so chunk ends fall inside records at varying offsets
    d1 = records(r1, lambda k: b"synth.%d extra=%d/1" % (k, 7 * k))
    d2 = None
    if paired:
        r2 = random_reads(n, L, seed=chunk + 99)
        d2 = records(r2, lambda k: b"synth.%d extra=%d/2" % (k, 7 * k))
    ranges = check_against_reference(d1, d2, ratio, chunk, tmp_path)
    # the chunk size makes several chunks; a ratio below one makes several blocks
    if ratio < 1:
        assert len(ranges) > 1


def test_run_counter_carries_across_candidates(tmp_path):
    """Headers share only '@ab.' with the chunk's first line, except every
    ninth; each header ends with '/1' like the first line, so its comparison
    ends with a run of 2 that the next earlier candidate continues: the
    reference takes that candidate after '@ab.' (2 + 4 > 5).  A loop that
    resets the counter per candidate cuts at the nearest header with the same
    first digit instead, and gives other blocks."""
    L = 40
    n = 120
    r = random_reads(n, L, seed=5, n_rate=0.0)
    d = records(r, lambda k: b"ab.%d%06d/1" % (1 + k % 9, k))
    ranges = check_against_reference(d, None, 0.3, 1200, tmp_path)
    assert len(ranges) > 1
    mutant_blocks, mutant_ranges = process_blocks(d, None, 0.3, 1200, reset_between_candidates=True)
    assert mutant_ranges != ranges


def test_quality_lines_starting_with_at(tmp_path):
    """Quality lines that begin with '@' are candidates too (the reference
    notes 'id or qual'); with constant '@@@...' qualities the cut rule still
    follows the reference loop."""
    L = 50
    r = random_reads(60, L, seed=9)
    d = records(r, lambda k: b"synth.%d/1" % k, qual=lambda k: b"@" * L)
    check_against_reference(d, None, 0.5, 1700, tmp_path)


def test_sequences_and_oracle_parser_match_reference():
    L = 75
    r = random_reads(40, L, seed=3)
    d = records(r, lambda k: b"synth.%d" % k)
    blocks, _ = process_blocks(d, None, 1.0, 1 << 20)
    want = np.frombuffer(b"".join(blocks[0]), dtype=np.uint8).reshape(-1, L)
    assert np.array_equal(fastq.sequences(d), want)
    assert np.array_equal(orc.parse_fastq(d, L), want)
    # paired: interleaved 2k / 2k+1
    r2 = random_reads(40, L, seed=4)
    d2 = records(r2, lambda k: b"synth.%d" % k)
    blocks, _ = process_blocks(d, d2, 1.0, 1 << 20)
    want = np.frombuffer(b"".join(blocks[0]), dtype=np.uint8).reshape(-1, L)
    assert np.array_equal(orc.parse_fastq(d, L, d2), want)


def test_parse_errors():
    L = 30
    r = random_reads(5, L, seed=1)
    d = records(r, lambda k: b"x%d" % k)
    bad_len = d.replace(r[2].tobytes(), r[2].tobytes()[:-1], 1)
    with pytest.raises(fastq.FastqError, match="read length"):
        fastq.sequences(bad_len)
    with pytest.raises(ValueError, match="read length"):
        orc.parse_fastq(bad_len, L)
    with pytest.raises(fastq.FastqError, match="truncated"):
        fastq.sequences(d[:-5])
    with pytest.raises(ValueError, match="truncated"):
        orc.parse_fastq(d[:-5], L)
    short2 = records(r[:4], lambda k: b"x%d" % k)
    with pytest.raises(ValueError, match="record counts"):
        orc.parse_fastq(d, L, short2)


def test_fastq_bytes_layout():
    r = random_reads(3, 20, seed=2)
    b = fastq.fastq_bytes(r, first_id=7, mate=2).tobytes()
    lines = b.split(b"\n")
    assert lines[0] == b"@synth.0000000007/2"
    assert lines[1] == r[0].tobytes()
    assert lines[2] == b"+" and lines[3] == b"I" * 20
    assert lines[4] == b"@synth.0000000008/2"
    assert np.array_equal(fastq.sequences(b), r)


def test_no_cut_is_an_error(tmp_path):
    """Headers with no common six-byte run never give a cut; the reference
    loops without advancing, here it is an error."""
    L = 30
    r = random_reads(50, L, seed=8)
    d = records(r, lambda k: b"%d" % (10 ** 6 + 104729 * k))
    p = tmp_path / "x.fastq"
    p.write_bytes(d)
    with pytest.raises(fastq.FastqError, match="no record start"):
        fastq.block_ranges(str(p), None, 1.0, 500)


@pytest.mark.parametrize("window", [7, 23, 64, 300])
def test_block_cut_with_small_search_windows(tmp_path, monkeypatch, window):
    """The backward scan reads the chunk in windows; with windows shorter than
    a record the search crosses many window starts, and the result must still
    be the reference loop's."""
    monkeypatch.setattr(fastq, "_WINDOW", window)
    L = 60
    r = random_reads(80, L, seed=window)
    d = records(r, lambda k: b"synth.%d extra=%d/1" % (k, 3 * k))
    check_against_reference(d, None, 0.3, 1900, tmp_path)
    d = records(r, lambda k: b"ab.%d%06d/1" % (1 + k % 9, k))
    check_against_reference(d, None, 0.3, 1200, tmp_path)


def test_block_ranges_on_buffers_equal_files(tmp_path):
    L = 50
    r1 = random_reads(70, L, seed=12)
    r2 = random_reads(70, L, seed=13)
    d1 = records(r1, lambda k: b"synth.%d/1" % k)
    d2 = records(r2, lambda k: b"synth.%d/2" % k)
    p1, p2 = tmp_path / "b_1.fastq", tmp_path / "b_2.fastq"
    p1.write_bytes(d1)
    p2.write_bytes(d2)
    want = fastq.block_ranges(str(p1), str(p2), 0.25, 1700)
    assert fastq.block_ranges(d1, d2, 0.25, 1700) == want
    assert fastq.block_ranges(np.frombuffer(d1, dtype=np.uint8), np.frombuffer(d2, dtype=np.uint8), 0.25,
                              1700) == want
    assert fastq.block_ranges(b"", None) == []
