"""Pins the CPU oracle: closed forms for the bit helpers, and brute-force
restatements (tests/bruteforce.py) for every block stage on tiny inputs.
The oracle is the checker the GPU path is held to, so these tests are what
make its answers trustworthy."""
import random

import numpy as np
import pytest

import bruteforce as bf
from oracle import oracle as orc

INV = bf.INV


def rand_reads(n, L, genome_len, seed, err=0.01, nrate=0.0, alphabet="ACGT"):
    rnd = random.Random(seed)
    g = "".join(rnd.choice(alphabet) for _ in range(genome_len))
    out = []
    for _ in range(n):
        s = rnd.randrange(0, genome_len - L + 1)
        r = g[s:s + L]
        if rnd.random() < 0.5:
            r = bf.revcomp_str(r)
        r = list(r)
        for i in range(L):
            if rnd.random() < err:
                r[i] = rnd.choice([c for c in "ACGT" if c != r[i]])
            if rnd.random() < nrate:
                r[i] = "N"
        out.append("".join(r))
    return out


def to_array(reads):
    return np.frombuffer("".join(reads).encode(), dtype=np.uint8).reshape(len(reads), -1).copy()


# ----------------------------------------------------------------- helpers ---
def test_params_table():
    # preprocess.cpp:307-313 for the read lengths of the workloads
    expect = {75: (3, 48, 27, 0, 15), 100: (4, 65, 32, 3, 20), 150: (5, 97, 32, 21, 30), 250: (8, 162, 32, 56, 50)}
    for L, (W, mo, k, kip, mm) in expect.items():
        p = orc.params(L)
        assert (p.W, p.max_off, p.k, p.kip, p.max_mm) == (W, mo, k, kip, mm)
    for L in range(10, 513):
        p = orc.params(L)
        assert p.kip + p.k + p.max_off == L and p.kip >= 0


def test_pack_known_words():
    reads = to_array(["ACGT" * 8])
    assert orc.pack(reads)[0, 0] == 0x1B1B1B1B1B1B1B1B
    # N and lowercase: N packs as A, lowercase as its base (compress.cu:2165)
    reads = to_array(["N" + "c" + "T" * 30])
    w = int(orc.pack(reads)[0, 0])
    assert w >> 62 == 0 and (w >> 60) & 3 == 1 and w & 3 == 3


@pytest.mark.parametrize("L", [1, 31, 32, 33, 75, 150, 250, 512])
def test_pack_layout_bruteforce(L):
    reads = rand_reads(20, L, max(L + 5, 600), seed=L, err=0.0, nrate=0.05)
    P = orc.pack(to_array(reads))
    W = (L + 31) // 32
    for r, row in zip(reads, P):
        s = bf.packed_string(r)
        for w in range(W):
            chunk = s[32 * w: 32 * w + 32]
            assert int(row[w]) == bf.kmer_value(chunk) << (2 * (32 - len(chunk)))


def test_murmur_matches_published_fmix64_and_is_invertible():
    L = orc.lib()
    rnd = random.Random(1)
    for x in [0, 1, 2, 0xFFFFFFFFFFFFFFFF] + [rnd.getrandbits(64) for _ in range(2000)]:
        y = L.orc_murmur64(x)
        assert y == bf.fmix64(x)
        assert bf.fmix64_inverse(y) == x
    assert L.orc_murmur64(0) == 0


def test_hash_size_closed_form():
    L = orc.lib()
    for n in [0, 1, 1 << 20, (1 << 24) - 1, 1 << 24, (1 << 24) + 1, 3 << 25, 1 << 31, (1 << 31) + 5, 1 << 40]:
        assert L.orc_hash_size(n) == bf.hash_size(n)


def test_extract_revcomp_mismatch_bruteforce():
    import ctypes
    L = orc.lib()
    rnd = random.Random(2)
    for _ in range(300):
        n_bases = rnd.randrange(40, 200)
        s = "".join(rnd.choice("ACGT") for _ in range(n_bases))
        words = np.array([bf.kmer_value(s[i:i + 32].ljust(32, "A")) for i in range(0, n_bases + 32, 32)],
                         dtype=np.uint64)
        ptr = words.ctypes.data
        off = rnd.randrange(0, n_bases - 1)
        n = rnd.randrange(1, min(32, n_bases - off) + 1)
        assert L.orc_extract_word(ptr, 0, off, n) == bf.kmer_value(s[off:off + n])
        assert L.orc_extract_word_unaligned(ptr, off, n) == bf.kmer_value(s[off:off + n])
        km = s[off:off + n]
        assert L.orc_revcomp_kmer(bf.kmer_value(km), n) == bf.kmer_value(bf.revcomp_str(km))
        t = "".join(c if rnd.random() > 0.2 else rnd.choice("ACGT") for c in km)
        assert L.orc_mismatch_words(bf.kmer_value(km), bf.kmer_value(t)) == sum(a != b for a, b in zip(km, t))
    del ctypes


def test_forward_and_rc_mismatch_bruteforce():
    L = orc.lib()
    rnd = random.Random(3)
    for _ in range(200):
        RL = rnd.choice([30, 64, 100, 150, 151])
        ref = "".join(rnd.choice("ACGT") for _ in range(RL + 80))
        rw = np.array([bf.kmer_value(ref[i:i + 32].ljust(32, "A")) for i in range(0, len(ref) + 32, 32)],
                      dtype=np.uint64)
        mp = rnd.randrange(0, 80)
        window = ref[mp:mp + RL]
        read = "".join(c if rnd.random() > 0.05 else rnd.choice("ACGT") for c in window)
        for strand in (0, 1):
            q = read if strand == 0 else bf.revcomp_str(read)
            packed = orc.pack(to_array([q]))[0]
            truth = sum(a != b for a, b in zip(window, read))
            bound = rnd.randrange(0, 12)
            f = L.orc_forward_mismatch if strand == 0 else L.orc_rc_mismatch
            got = f(rw.ctypes.data, mp, packed.ctypes.data, RL, bound)
            if truth <= bound:
                assert got == truth
            else:
                assert got > bound


# ------------------------------------------------------------ block stages ---
def block_case(n, L, glen, seed, err=0.01, nrate=0.0):
    reads = rand_reads(n, L, glen, seed, err=err, nrate=nrate)
    return reads, orc.OracleBlock(to_array(reads), stop="all")


# the last three take the k = 32 branch (L >= 89) that every benchmark read length uses:
# k-mer index position 3 (L = 100), 21 (L = 150), and 2 at L = 96
CASES = [(120, 40, 300, 1, 0.01, 0.0), (200, 36, 150, 2, 0.0, 0.0), (150, 50, 400, 3, 0.02, 0.01),
         (90, 45, 90, 4, 0.0, 0.0), (160, 64, 700, 5, 0.005, 0.002), (150, 96, 400, 6, 0.005, 0.002),
         (150, 100, 500, 7, 0.01, 0.0), (150, 150, 600, 8, 0.005, 0.002)]


@pytest.mark.parametrize("case", CASES)
def test_block_stages_bruteforce(case):
    n, L, glen, seed, err, nrate = case
    reads, ob = block_case(n, L, glen, seed, err, nrate)
    P = bf.params(L)
    nonN = [i for i, r in enumerate(reads) if "N" not in r]
    Nids = [i for i, r in enumerate(reads) if "N" in r]
    assert list(ob.get("nonN_ids")) == nonN and list(ob.get("N_ids")) == Nids

    order, nxt, prv, off = bf.dedup(reads, nonN)
    assert list(ob.get("sorted_ids")) == order
    assert list(ob.get("next_dedup")) == nxt and list(ob.get("prev_dedup")) == prv

    bf.greedy_overlap(reads, order, nxt, prv, off, P)
    assert list(ob.get("next_overlap")) == nxt
    assert list(ob.get("prev_overlap")) == prv
    assert list(ob.get("offset_overlap")) == off

    hq, unm = bf.prune_partition(order, nxt, prv, off)
    assert list(ob.get("hq_ids")) == hq and list(ob.get("unmatched_ids")) == unm
    assert list(ob.get("next_prune")) == nxt and list(ob.get("prev_prune")) == prv

    seqs = {i: bf.packed_string(reads[i]) for i in range(len(reads))}
    bf.full_match(seqs, hq, nxt, prv, off, L)
    assert list(ob.get("next")) == nxt and list(ob.get("prev")) == prv
    assert list(ob.get("offset")) == off

    text, recs = bf.assemble(seqs, hq, nxt, prv, off, L)
    assert ob.get("hq_pg").tobytes().decode() == text
    rec = ob.get("hq_rec")
    assert [(int(r["read_id"]), int(r["word"]) >> 2) for r in rec] == recs
    assert all(int(r["word"]) & 3 == 0 for r in rec)

    mapped = bf.map_reads(text, [reads[i] for i in unm] + [reads[i] for i in Nids], P)
    mr, mm = ob.get("map_rec"), ob.get("map_mis")
    exp = [(i, m) for i, m in zip(unm + list(range(len(Nids))), mapped) if m is not None]
    isN = [False] * len(unm) + [True] * len(Nids)
    exp_isN = [f for f, m in zip(isN, mapped) if m is not None]
    assert len(mr) == len(exp)
    for r, c, (rid, (strand, pos, cnt)), flagN in zip(mr, mm, exp, exp_isN):
        w = int(r["word"])
        assert (int(r["read_id"]), w & 1, (w >> 1) & 1, w >> 2, int(c)) == (rid, int(flagN), strand, pos, cnt)
    unmapped = [i for i, m in zip(unm, mapped[:len(unm)]) if m is None]
    unmappedN = [k for k, m in enumerate(mapped[len(unm):]) if m is None]
    assert list(ob.get("unmapped")) == unmapped and list(ob.get("unmapped_N")) == unmappedN

    # LQ: exact all-offset matching over the unmapped reads (local ids)
    lseqs = {k: seqs[g] for k, g in enumerate(unmapped)}
    ln, lp, lo = [INV] * len(unmapped), [INV] * len(unmapped), [0xFFFF] * len(unmapped)
    bf.full_match(lseqs, list(range(len(unmapped))), ln, lp, lo, L)
    assert list(ob.get("lq_next")) == ln and list(ob.get("lq_prev")) == lp and list(ob.get("lq_offset")) == lo
    ltext, lrecs = bf.assemble(lseqs, list(range(len(unmapped))), ln, lp, lo, L)
    assert ob.get("lq_pg").tobytes().decode() == ltext
    assert [(int(r["read_id"]), int(r["word"]) >> 2) for r in ob.get("lq_rec")] == \
        [(unmapped[i], p) for i, p in lrecs]

    # N-LQ (compress.cu:1176-1347): ASCII, alphabet A<C<G<N<T, ties by N index
    if unmappedN:
        nseqs = {k: reads[Nids[k]] for k in range(len(Nids))}
        ids = sorted(unmappedN, key=lambda k: (nseqs[k], k))
        nn, npv, no = [INV] * len(Nids), [INV] * len(Nids), [0xFFFF] * len(Nids)
        for a, b in zip(ids, ids[1:]):
            if nseqs[a] == nseqs[b]:
                nn[a], npv[b], no[a] = b, a, 0
        bf.full_match(nseqs, ids, nn, npv, no, L, alphabet="ACGNT")
        ntext, nrecs = bf.assemble(nseqs, ids, nn, npv, no, L)
        assert ob.get("nlq_pg").tobytes().decode() == ntext
        assert [(int(r["read_id"]), int(r["word"]) >> 2) for r in ob.get("nlq_rec")] == nrecs
        assert all(int(r["word"]) & 1 == 1 for r in ob.get("nlq_rec"))


@pytest.mark.parametrize("seed", [11, 12])
def test_block_invariants_c1_like(seed):
    """Invariants at a few thousand reads: every read is placed exactly once,
    every placement reproduces the read, mismatch counts recount exactly."""
    L = 100
    reads = rand_reads(3000, L, 20000, seed, err=0.006, nrate=0.0005)
    ob = orc.OracleBlock(to_array(reads), stop="all")
    Nids = list(ob.get("N_ids"))
    seen = {}
    hq = ob.get("hq_pg").tobytes().decode()
    for r in ob.get("hq_rec"):
        rid, pos = int(r["read_id"]), int(r["word"]) >> 2
        assert hq[pos:pos + L] == bf.packed_string(reads[rid])
        seen[rid] = seen.get(rid, 0) + 1
    for r, mm in zip(ob.get("map_rec"), ob.get("map_mis")):
        w = int(r["word"])
        rid = Nids[int(r["read_id"])] if w & 1 else int(r["read_id"])
        strand, pos = (w >> 1) & 1, w >> 2
        q = reads[rid] if strand == 0 else bf.revcomp_str(reads[rid])
        assert sum(a != b for a, b in zip(hq[pos:pos + L], q)) == int(mm) <= L // 5
        seen[rid] = seen.get(rid, 0) + 1
    lq = ob.get("lq_pg").tobytes().decode()
    for r in ob.get("lq_rec"):
        rid, pos = int(r["read_id"]), int(r["word"]) >> 2
        assert lq[pos:pos + L] == reads[rid]
        seen[rid] = seen.get(rid, 0) + 1
    nlq = ob.get("nlq_pg").tobytes().decode()
    for r in ob.get("nlq_rec"):
        rid, pos = Nids[int(r["read_id"])], int(r["word"]) >> 2
        assert nlq[pos:pos + L] == reads[rid]
        seen[rid] = seen.get(rid, 0) + 1
    assert sorted(seen) == list(range(len(reads))) and set(seen.values()) == {1}


def test_openmp_build_is_byte_identical():
    import curc_amd as ca
    _, reads = ca.synth_reads(ca.synth_config(seed=77, genome_len=40_000, n_reads=6_000, L=150))
    a = orc.OracleBlock(reads, kind="st")
    b = orc.OracleBlock(reads, kind="omp")
    for f in orc.FIELDS:
        assert np.array_equal(a.get(f), b.get(f)), f


def _periodic_reads(seed=0, unit_len=40, n=250, L=40, err=0.03):
    rnd = random.Random(seed)
    unit = "".join(rnd.choice("ACGT") for _ in range(unit_len))
    flank = lambda: "".join(rnd.choice("ACGT") for _ in range(100))  # noqa: E731
    g = flank() + unit * 12 + flank()
    reads = []
    for _ in range(n):
        s = rnd.randrange(0, len(g) - L + 1)
        r = list(g[s:s + L])
        if rnd.random() < 0.5:
            r = list(bf.revcomp_str("".join(r)))
        for i in range(L):
            if rnd.random() < err:
                r[i] = rnd.choice([c for c in "ACGT" if c != r[i]])
        reads.append("".join(r))
    return reads


def test_plausible_mistakes_are_detected():
    """Each brute-force check has discriminating power: a plausible mistake in
    the rule it states gives a different answer from the oracle.  Mutants:
    suffixes processed in reverse list order in a greedy-overlap round; exact
    matching ties ordered by the forward instead of the reversed prefix; the
    mapping index buckets in ascending instead of descending position."""
    reads, ob = block_case(300, 40, 200, 22, 0.005, 0.0)
    P = bf.params(40)
    order, nxt, prv, off = bf.dedup(reads, list(range(len(reads))))
    mut = (list(nxt), list(prv), list(off))
    bf.greedy_overlap(reads, order, *mut, P, mutant=True)
    assert mut[0] != list(ob.get("next_overlap"))
    bf.greedy_overlap(reads, order, nxt, prv, off, P)
    assert nxt == list(ob.get("next_overlap"))

    reads = _periodic_reads()
    ob = orc.OracleBlock(to_array(reads), stop="all")
    order, nxt, prv, off = bf.dedup(reads, list(range(len(reads))))
    bf.greedy_overlap(reads, order, nxt, prv, off, P)
    hq, _ = bf.prune_partition(order, nxt, prv, off)
    seqs = {i: bf.packed_string(r) for i, r in enumerate(reads)}
    mut = (list(nxt), list(prv), list(off))
    bf.full_match(seqs, hq, *mut, 40, mutant=True)
    assert mut[0] != list(ob.get("next"))
    bf.full_match(seqs, hq, nxt, prv, off, 40)
    assert nxt == list(ob.get("next"))
    text = ob.get("hq_pg").tobytes().decode()
    unm = list(ob.get("unmatched_ids"))
    good = bf.map_reads(text, [reads[i] for i in unm], P)
    bad = bf.map_reads(text, [reads[i] for i in unm], P, mutant=True)
    got = [(int(r["word"]) >> 1 & 1, int(r["word"]) >> 2, int(m)) for r, m in zip(ob.get("map_rec"), ob.get("map_mis"))]
    assert got == [m for m in good if m is not None]
    assert got != [m for m in bad if m is not None]
