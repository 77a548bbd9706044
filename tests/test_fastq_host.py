"""The library's host-side FASTQ record parse (curc_fastq_host_parse), which
splits one file's bytes over threads by guessing record starts and keeps a
guess only when the record chain of the thread before lands on it.  Checked
against the oracle's parser (itself pinned to a literal restatement of
preprocess.cpp:147-199 in test_fastq.py) and fastq.sequences, on inputs
made to mislead the guesses, for 1..16 threads.  CPU only."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import curc_amd as ca  # noqa: E402
from curc_amd import fastq  # noqa: E402
from oracle import oracle as orc  # noqa: E402

THREADS = (1, 2, 3, 7, 16)


def fq(seqs, header, plus, qual):
    out = bytearray()
    for k, s in enumerate(seqs):
        out += header(k) + b"\n" + s + b"\n" + plus(k) + b"\n" + qual(k) + b"\n"
    return bytes(out)


def rand_seqs(rng, n, L, alphabet=b"ACGTN"):
    a = np.frombuffer(alphabet, dtype=np.uint8)
    return [a[rng.integers(0, len(a), L)].tobytes() for _ in range(n)]


def check(data, L):
    want = orc.parse_fastq(data, L)
    assert np.array_equal(fastq.sequences(data, L), want)
    for t in THREADS:
        got = ca.fastq_host_parse(data, L, t)
        assert got.shape == want.shape and np.array_equal(got, want), f"threads {t}"


@pytest.mark.parametrize("L", [1, 31, 100, 150])
def test_plain_records(L):
    rng = np.random.default_rng(L)
    s = rand_seqs(rng, 300, L)
    check(fq(s, lambda k: b"@read.%d" % k, lambda k: b"+", lambda k: b"I" * L), L)


def test_misleading_lines():
    """Quality lines that start with '@' and '+' lines with names, headers as
    long as a sequence line, sequences of '+' characters, headers without
    '@' after the first record: lines that pass for record starts in the
    middle of records, so that guesses fail and the file is parsed again."""
    L = 40
    rng = np.random.default_rng(5)
    s = rand_seqs(rng, 400, L, alphabet=b"+ACGT@")
    data = fq(s, lambda k: (b"@" + b"h" * (L - 1)) if k % 3 else b"x%d" % k if k else b"@first",
              lambda k: b"+" + s[k][:k % 7], lambda k: b"@" + b"+" * (L - 1) if k % 2 else b"+" * L)
    check(data, L)
    # every qual line is '@' + a line of L '+', so a guess at a quality line passes the guess rule
    data = fq([b"+" * L] * 200, lambda k: b"@" + b"+" * (L - 1), lambda k: b"+", lambda k: b"@" * L)
    check(data, L)


def test_variable_quality_lengths():
    """The quality line's length is not checked (the device parser counts
    newlines only), so records of other quality lengths parse as usual."""
    L = 25
    rng = np.random.default_rng(8)
    s = rand_seqs(rng, 500, L)
    check(fq(s, lambda k: b"@r%d" % k, lambda k: b"+", lambda k: b"I" * (k % 60)), L)


def test_empty_and_single():
    assert ca.fastq_host_parse(b"", 10, 4).shape == (0, 10)
    check(b"@a\nACGTACGTAC\n+\nIIIIIIIIII\n", 10)


@pytest.mark.parametrize("t", THREADS)
def test_format_errors_in_device_order(t):
    L = 30
    rng = np.random.default_rng(2)
    s = rand_seqs(rng, 200, L)
    d = fq(s, lambda k: b"@x%d" % k, lambda k: b"+", lambda k: b"I" * L)
    short = fq(s[:120] + [s[120][:-1]] + s[121:], lambda k: b"@x%d" % k, lambda k: b"+", lambda k: b"I" * L)
    crlf = d.replace(b"\n", b"\r\n")
    cases = [
        (b"x" + d[1:], "start with '@'"),
        (d[:-1], "truncated"),
        (d[:-40], "truncated"),
        (d + b"@y\n", "truncated"),
        (short, "read length"),
        (short[:-1], "truncated"),  # a truncation is reported before a length error
        (crlf, "read length"),
    ]
    for data, msg in cases:
        with pytest.raises(ca.CurcError, match=msg):
            ca.fastq_host_parse(data, L, t)
        with pytest.raises(Exception, match=msg):
            fastq.sequences(data, L)


def test_synth_fastq_layout():
    """The FASTQ layout the bench feeds (curc_synth_fastq) at a size where
    every thread has a range of its own."""
    L = 150
    cfg = ca.synth_config(seed=4, genome_len=50_000, n_reads=20_000, L=L, n_rate=0.001)
    reads = ca.synth_reads(cfg)[1]
    data = ca.synth_fastq(reads)
    for t in THREADS:
        assert np.array_equal(ca.fastq_host_parse(data, L, t), reads), f"threads {t}"
