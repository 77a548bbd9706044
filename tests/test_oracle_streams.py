"""Pins the oracle's stream builders (compress.cu:1730-2044) by decoding:
the streams plus the pseudogenomes must give back every read, the way
CURC's decompressor rebuilds them (decompress.cu:909-994 for the
non-preserving layout, 472-577 for the order-preserving one)."""
import numpy as np
import pytest

import bruteforce as bf
import curc_amd as ca
from oracle import oracle as orc


def noisy_mix(seed=63):
    clean = ca.synth_config(seed=seed, genome_len=30_000, n_reads=5_000, L=120, sub_first=0.001,
                            sub_last=0.002, n_rate=0.0005)
    genome, a = ca.synth_reads(clean)
    noisy = ca.synth_config(seed=seed + 1, genome_len=30_000, n_reads=2_000, L=120, sub_first=0.03,
                            sub_last=0.1, n_rate=0.003)
    _, b = ca.synth_reads(noisy, genome)
    return np.concatenate([a, b])


def decode_offsets(stream, m, n_rec, L):
    """Mismatch positions of the n_rec records with m mismatches, in order."""
    w = 2 if m == 1 else 3 if m == 2 else m + 2
    out = []
    for r in range(n_rec):
        v = [int(x) for x in stream[r * w:(r + 1) * w]]
        flag, first = v[0], v[1]
        if m == 1:
            offs = [first] if flag == 0 else [L - 1 - first]
        elif m == 2:
            if flag == 0:
                offs = [first, first + v[2] + 1]
            else:
                last = L - 1 - first
                offs = [last - v[2] - 1, last]
        else:
            mn = v[2] + 1
            gaps = [g + mn for g in v[3:]]
            if flag == 0:
                offs = [first]
                for g in gaps:
                    offs.append(offs[-1] + g)
            else:
                offs = [L - 1 - first]
                for g in gaps:
                    offs.append(offs[-1] - g)
                offs = offs[::-1]
        out.append(offs)
    return out


def rebuild(ob, reads, preserve):
    L = reads.shape[1]
    hq = ob.get("hq_pg").tobytes().decode()
    lq = ob.get("lq_pg").tobytes().decode()
    nlq = ob.get("nlq_pg").tobytes().decode()
    table = ob.get("s_mis_table").reshape(4, 4)
    ctx_row = {"A": 0, "T": 1, "C": 2, "G": 3}
    counts = ob.get("s_mis_count")
    strand = ob.get("s_strand").tobytes().decode()
    syms = ob.get("s_mis_base")
    per_m = {m: iter(decode_offsets(ob.mis_off(m), m, int((counts == m).sum()), L))
             for m in range(1, L // 5 + 1) if (counts == m).any()}
    rec = ob.get("s_sorted_rec")
    Nids = ob.get("N_ids")
    got = []
    k = 0
    pos = 0
    for i, r in enumerate(rec):
        w = int(r["word"])
        p = w >> 2
        if not preserve:
            pos = 0 if i == 0 else pos + int(ob.get("s_read_off")[i])
            assert pos == p
        s = list(hq[p:p + L])
        m = int(counts[i])
        if m:
            for off in next(per_m[m]):
                s[off] = chr(table[ctx_row[s[off]]][syms[k]])
                k += 1
        s = "".join(s)
        if strand[i] == "1":
            s = bf.revcomp_str(s)
        gid = int(Nids[r["read_id"]]) if w & 1 else int(r["read_id"])
        got.append((gid, s))
    assert k == len(syms)
    for name, text, recs, isN in (("lq", lq, ob.get("lq_rec"), False), ("nlq", nlq, ob.get("nlq_rec"), True)):
        if not preserve:
            offs = ob.get("s_" + name + "_off")
            q = 0
            for j, r in enumerate(recs):
                q = q + (L - int(offs[j])) if j else L - int(offs[j])
                assert q == int(r["word"]) >> 2
        for r in recs:
            p = int(r["word"]) >> 2
            gid = int(Nids[r["read_id"]]) if isN else int(r["read_id"])
            got.append((gid, text[p:p + L]))
    return got


@pytest.mark.parametrize("preserve", [False, True])
def test_streams_rebuild_every_read(preserve):
    reads = noisy_mix()
    ob = orc.OracleBlock(reads, stop="all")
    ob.streams(preserve, False)
    got = rebuild(ob, reads, preserve)
    assert sorted(g for g, _ in got) == list(range(len(reads)))
    for gid, s in got:
        assert s == reads[gid].tobytes().decode()
    if preserve:
        idpos = ob.get("s_id_pos")
        hq_len, lq_len = len(ob.get("hq_pg")), len(ob.get("lq_pg"))
        assert len(idpos) == len(reads) and int(idpos.max()) < hq_len + lq_len + len(ob.get("nlq_pg"))


def test_streams_reach_wide_offset_widths():
    o = orc.OracleBlock(noisy_mix(), stop="all")
    o.streams(False, False)
    assert max(m for m in range(1, 25) if len(o.mis_off(m))) >= 8


def test_paired_order_covers_all_reads():
    cfg = ca.synth_config(seed=62, genome_len=60_000, gc=0.5, n_reads=10_000, L=100, paired=1,
                          sub_first=0.003, sub_last=0.01, n_rate=0.001)
    reads = ca.synth_reads(cfg)[1]
    o = orc.OracleBlock(reads, stop="all")
    o.streams(False, True)
    assert sorted(o.get("s_pe_order").tolist()) == list(range(len(reads)))


def test_pe_slots_match_the_pair_stream():
    """The vectorised output slots of a paired block equal the sequential
    decode of the archive's pair stream."""
    from curc_amd import archive

    rng = np.random.default_rng(3)
    order = rng.permutation(2000)
    d, s = archive._pe_order_encode(order)
    assert np.array_equal(archive.pe_slots_from_order(order), archive._pe_order_decode(d, s, 2000))
