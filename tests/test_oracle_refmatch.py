"""Pins the oracle's RefMatcher (compress.cu:597-1001) on the CPU: closed
forms for the seed/extension rules, an independent brute-force statement,
planted mistakes that the checks must catch, and lossless expansion by the
archive decoder (decompress.cpp:339-366)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bruteforce as bf  # noqa: E402
from curc_amd import archive  # noqa: E402
from oracle import oracle as orc  # noqa: E402

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)
RC = bytes.maketrans(b"ACGTN", b"TGCAN")


def rnd(rng, n):
    return ACGT[rng.integers(0, 4, n)].tobytes()


def rc(s):
    return s.translate(RC)[::-1]


def frugal(v):
    o = bytearray()
    while v >= 128:
        o.append(128 + v % 128)
        v //= 128
    o.append(v)
    return bytes(o)


def mutate(rng, s, rate):
    a = bytearray(s)
    for i in np.flatnonzero(rng.random(len(a)) < rate):
        a[i] = b"ACGT"[(b"ACGT".index(a[i]) + 1 + rng.integers(0, 3)) % 4]
    return bytes(a)


def structured(rng, n_pieces=14, lo=40, hi=420, err=0.004):
    """Concatenated pieces: fresh random sequence, reverse complements and
    forward repeats of earlier pieces, each with a few substitutions."""
    pieces = []
    for _ in range(n_pieces):
        r = rng.random()
        if pieces and r < 0.35:
            p = rc(pieces[rng.integers(0, len(pieces))])
        elif pieces and r < 0.5:
            p = pieces[rng.integers(0, len(pieces))]
        elif r < 0.6:
            unit = rnd(rng, int(rng.integers(5, 12)))
            p = (unit * 60)[: int(rng.integers(lo, hi))]
        else:
            p = rnd(rng, int(rng.integers(lo, hi)))
        pieces.append(mutate(rng, p, err))
    return b"".join(pieces)


def with_n(rng, s, rate=0.01):
    a = bytearray(s)
    a_idx = np.flatnonzero(rng.random(len(a)) < rate)
    for i in a_idx:
        a[i] = ord("N")
    return bytes(a)


def test_marush_hash_matches_bruteforce():
    rng = np.random.default_rng(3)
    L = orc.lib()
    for _ in range(200):
        s = bytes(rng.integers(0, 256, 36, dtype=np.uint8))
        assert L.orc_marush36(s) == bf.rm_hash(s)
    # hash = 36; per word k += j; hash ^= k; hash *= 171717 (mod 2^64),
    # truncated to 32 bits: for 36 NUL bytes the words are 0..8
    h = 36
    for j in range(9):
        h = ((h ^ j) * 171717) % (1 << 64)
    assert L.orc_marush36(b"\0" * 36) == h % (1 << 32)


@pytest.mark.parametrize("n", [51, 300, 3000])
def test_reverse_complement_copy_closed_form(n):
    """X + spacer + rc(X): one match whose left end gives up the first base
    (the left loop stops at a text start without counting it)."""
    rng = np.random.default_rng(n)
    X = rnd(rng, n)
    while True:
        sp = rnd(rng, 20)
        # the extension must stop at X's end on both copies
        if sp[0:1] != rc(sp)[0:1] and X[0:1] != rc(X)[-1:]:
            break
    pg = X + sp + rc(X)
    r = orc.ref_match(pg, pg, False, True)
    assert r["matches"].tolist() == [[1, n - 1, n + 20]]
    assert r["text"] == pg[: n + 20] + b"%" + pg[-1:]
    assert r["map_off"] == (1).to_bytes(4, "little")
    assert r["map_len"] == frugal(50) + frugal(n - 1 - 50)
    assert archive.recover_ref(r["text"], None, r["map_off"], r["map_len"], True, True) == pg


def test_match_length_threshold_closed_form():
    """A 50-base copy (49 bases after the boundary rule) is not matched; a
    51-base copy is, with length exactly 50."""
    rng = np.random.default_rng(7)
    for n, expect in ((50, []), (51, [[1, 50, 71]])):
        for _ in range(20):
            X = rnd(rng, n)
            sp = rnd(rng, 20)
            if sp[0:1] != rc(sp)[0:1] and X[0:1] != rc(X)[-1:]:
                break
        pg = X + sp + rc(X)
        assert orc.ref_match(pg, pg, False, True)["matches"].tolist() == expect


def test_unmapped_text_against_reference_closed_form():
    """dest = rc(src[a:b]) plus flanks: the match points back at src[a:b]."""
    rng = np.random.default_rng(11)
    src = rnd(rng, 2000)
    a, b = 500, 900
    while True:
        fl, fr = rnd(rng, 30), rnd(rng, 30)
        piece = rc(src[a:b])
        # flanks must break the extension on both sides
        if fl[-1:] != rc(src[b:b + 1]) and fr[:1] != rc(src[a - 1:a]):
            break
    dest = fl + piece + fr
    r = orc.ref_match(src, dest, False, False)
    assert r["matches"].tolist() == [[a, b - a, 30]]
    assert r["text"] == fl + b"%" + fr
    assert archive.recover_ref(r["text"], src, r["map_off"], r["map_len"], True, False) == dest


def test_short_texts_yield_no_matches():
    rng = np.random.default_rng(5)
    s = rnd(rng, 35)
    r = orc.ref_match(s, s, False, True)
    assert r["text"] == s and r["map_off"] == b"" and r["map_len"] == b""
    r = orc.ref_match(rnd(rng, 500), s, True, False)
    assert r["text"] == s and r["map_len"] == b""


CASES = [(seed, nlq) for seed in range(8) for nlq in (False, True)]


@pytest.mark.parametrize("seed,nlq", CASES)
def test_oracle_equals_bruteforce(seed, nlq):
    rng = np.random.default_rng(100 + seed)
    hq = structured(rng)
    if nlq:
        dest = with_n(rng, mutate(rng, rc(hq[len(hq) // 3: 2 * len(hq) // 3]), 0.003) + structured(rng, 4))
        runs = [(hq, dest, True, False)]
    else:
        lq = mutate(rng, rc(hq[: len(hq) // 2]), 0.004) + structured(rng, 4)
        runs = [(hq, hq, False, True), (hq, lq, False, False)]
    for src, dest, cn, isref in runs:
        r = orc.ref_match(src, dest, cn, isref)
        t, mo, ml, found, matches = bf.ref_match(src, dest, cn, isref)
        assert r["found"].tolist() == [list(f) for f in found]
        assert r["matches"].tolist() == [list(m) for m in matches]
        assert (r["text"], r["map_off"], r["map_len"]) == (t, mo, ml)
        back = archive.recover_ref(r["text"], None if isref else src, r["map_off"], r["map_len"], True, isref)
        assert back == dest


def test_matches_are_reverse_complement_copies():
    rng = np.random.default_rng(21)
    hq = structured(rng, 24)
    r = orc.ref_match(hq, hq, False, True)
    assert len(r["matches"]) > 3
    for s, n, d in r["matches"].tolist():
        assert s + n <= d  # the source precedes the destination after the margin split
        assert hq[d:d + n] == rc(hq[s:s + n])


def test_capacity_truncation_keeps_first_results():
    """A highly repetitive destination finds more results than the capacity
    (20% of the probe positions); the first ones in segment order are kept."""
    rng = np.random.default_rng(9)
    unit = rnd(rng, 7)
    src = (unit * 400) + rnd(rng, 50)
    dest = rc(src)
    r = orc.ref_match(src, dest, False, False)
    t, mo, ml, found, matches = bf.ref_match(src, dest, False, False)
    assert len(found) > r["capacity"] > 0
    assert r["matches"].tolist() == [list(m) for m in matches]
    assert archive.recover_ref(r["text"], src, r["map_off"], r["map_len"], True, False) == dest


def test_plausible_mistakes_are_detected():
    """Each planted mistake in the brute-force statement disagrees with the
    oracle on one of the cases above."""
    inputs = []
    for seed, nlq in CASES:
        rng = np.random.default_rng(100 + seed)
        hq = structured(rng)
        if nlq:
            dest = with_n(rng, mutate(rng, rc(hq[len(hq) // 3: 2 * len(hq) // 3]), 0.003) + structured(rng, 4))
            inputs.append((hq, dest, True, False))
        else:
            lq = mutate(rng, rc(hq[: len(hq) // 2]), 0.004) + structured(rng, 4)
            inputs += [(hq, hq, False, True), (hq, lq, False, False)]
    rng = np.random.default_rng(7)
    X = rnd(rng, 51)
    inputs.append((X + b"A" * 20 + rc(X), X + b"A" * 20 + rc(X), False, True))
    for mutant in ("left", "bucket", "segment", "threshold"):
        caught = False
        for src, dest, cn, isref in inputs:
            r = orc.ref_match(src, dest, cn, isref)
            got = bf.ref_match(src, dest, cn, isref, mutant=mutant)
            if (r["found"].tolist() != [list(f) for f in got[3]] or r["text"] != got[0]):
                caught = True
                break
        assert caught, mutant


def test_block_archive_round_trip_with_self_match():
    """Oracle block -> streams + self-match -> archive -> decode: every read
    comes back, and the archive shrinks (the HQ pseudogenome holds both
    strands of the genome)."""
    import tempfile

    import curc_amd as ca

    cfg = ca.synth_config(seed=5, genome_len=60_000, gc=0.5, n_reads=12_000, L=150, rc_prob=0.5, sub_first=0.005,
                          sub_last=0.005, n_rate=0.0001)
    _, reads = ca.synth_reads(cfg)
    ob = orc.OracleBlock(reads)
    ob.streams()
    plain = archive.BlockStreams.from_block(ob, len(reads), 150, False, False)
    ob.refmatch()
    bs = archive.BlockStreams.from_block(ob, len(reads), 150, False, False)
    assert bs.rm_texts is not None and len(bs.rm_texts[0]) < 0.8 * len(bs.hq_text)
    sizes = {}
    with tempfile.TemporaryDirectory() as d:
        for name, b in (("plain", plain), ("rm", bs)):
            p = os.path.join(d, name + ".curc")
            archive.write_archive(p, [b], False, False, 150)
            sizes[name] = os.path.getsize(p)
            _, _, L, blocks = archive.read_archive(p)
            assert blocks[0].hq_text == bs.hq_text and blocks[0].lq_text == bs.lq_text
            dec = archive.decode_block(blocks[0], L, False, False)
            assert sorted(map(bytes, dec)) == sorted(map(bytes, reads))
    assert sizes["rm"] < sizes["plain"]
