"""The .curc archive: container, entropy coding and decoding.

Layout (preprocess.cpp:261-276, compress.cu:1364-2147; read back by
decompress.cu:760-1060):

    archive := u8 preserve_order, u8 paired_end, u16 read_len, u8 block_count, block*
    block   := u32 reads_count, 16 B mismatch decoder rows (contexts A, T, C, G),
               [u8 is_join_ref_len_std if preserve_order], u8 is_ref_len_std,
               u64 hq_len, u64 lq_len, u64 n_len,
               sub-files, each u64 size + payload, in this order:
                 [id_pos if preserve_order], strand_id, ref,
                 [read_off, unmapping_N_read_off, unmapping_read_off, [pe_order if paired]
                  if not preserve_order],
                 ref_off.map, ref_len.map, unmap_ref_off.map, unmap_ref_len.map,
                 unmap_n_ref_off.map, unmap_n_ref_len.map, mismatch_count, mismatch_base,
               u8 max_mismatch, mismatch_off_1 .. mismatch_off_max (each u64 size + payload)

The pre-entropy streams are CURC's.  Each sub-file payload here is a run of
xz (LZMA2) streams over 8 MiB chunks (u32 count, then u64 size + stream per
chunk), coded in parallel; the pseudogenome text is first packed to 2 bits per base
with an exception list for other symbols.  The reference's coders (PgRC's
PPMd/LZMA framing, VarLenDNA + fast-lzma2, bsc, a range coder) are
third-party code this package does not carry, so archives are not
byte-compatible with CURC's; the streams they hold are.
"""
import io
import lzma
import os
import struct

import numpy as np

XZ_PRESET = 6
COMP = np.zeros(256, dtype=np.uint8)
for _a, _b in zip(b"ACGTNacgtn", b"TGCANtgcan"):
    COMP[_a] = _b
# complement table of the pseudogenome self-match (helper.cpp:253-268):
# IUPAC symbols map to their complements (lower case to upper), the rest to 0
_RM_COMP = bytearray(256)
for _a, _b in zip(b"AaCcGgTtNnUuYyRrKkMmBbDdHhVv", b"TTGGCCAANNAARRYYMMKKVVHHDDBB"):
    _RM_COMP[_a] = _b
RM_COMP = bytes(_RM_COMP)
MAP_NAMES = ("ref_off", "ref_len", "unmap_ref_off", "unmap_ref_len", "unmap_n_ref_off", "unmap_n_ref_len")


# ------------------------------------------------------------------ coders ---
# A sub-file payload is split into chunks of XZ_CHUNK bytes, each its own xz
# stream: u32 chunk count, then u64 size + xz bytes per chunk.  The chunks of
# a block are (de)compressed on a thread pool (the lzma module releases the
# GIL), so archive time scales with cores; XZ_CHUNK matches the dictionary of
# preset 6, so chunking costs little ratio.
XZ_CHUNK = 8 << 20
_POOL = None


def _pool():
    global _POOL
    if _POOL is None:
        from concurrent.futures import ThreadPoolExecutor

        _POOL = ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)))
    return _POOL


def _xz1(data: bytes) -> bytes:
    return lzma.compress(data, format=lzma.FORMAT_XZ, preset=XZ_PRESET)


def _unxz1(data: bytes) -> bytes:
    return lzma.decompress(data, format=lzma.FORMAT_XZ)


def xz_many(payloads):
    """Chunked xz containers of several payloads, compressed in parallel."""
    pieces = [[p[i:i + XZ_CHUNK] for i in range(0, len(p), XZ_CHUNK)] for p in payloads]
    flat = [c for ps in pieces for c in ps]
    comp = list(_pool().map(_xz1, flat))
    out, k = [], 0
    for ps in pieces:
        parts = [struct.pack("<I", len(ps))]
        for c in comp[k:k + len(ps)]:
            parts += [struct.pack("<Q", len(c)), c]
        k += len(ps)
        out.append(b"".join(parts))
    return out


def unxz_many(containers):
    """Inverse of xz_many."""
    spans = []
    for data in containers:
        (n,) = struct.unpack_from("<I", data, 0)
        off, cs = 4, []
        for _ in range(n):
            (ln,) = struct.unpack_from("<Q", data, off)
            cs.append(data[off + 8:off + 8 + ln])
            off += 8 + ln
        spans.append(cs)
    flat = list(_pool().map(_unxz1, [c for cs in spans for c in cs]))
    out, k = [], 0
    for cs in spans:
        out.append(b"".join(flat[k:k + len(cs)]))
        k += len(cs)
    return out


def xz(data: bytes) -> bytes:
    return xz_many([data])[0]


def unxz(data: bytes) -> bytes:
    return unxz_many([data])[0]


def pack_dna(text: bytes) -> bytes:
    """2-bit ACGT packing with (position, byte) exceptions, then xz."""
    return xz(pack_dna_raw(text))


def pack_dna_raw(text: bytes) -> bytes:
    """2-bit ACGT packing with (position, byte) exceptions."""
    a = np.frombuffer(text, dtype=np.uint8)
    code = np.full(256, 255, dtype=np.uint8)
    code[[65, 67, 71, 84]] = [0, 1, 2, 3]
    c = code[a]
    exc = np.flatnonzero(c == 255)
    c = np.where(c == 255, 0, c).astype(np.uint8)
    pad = (-len(c)) % 4
    c4 = np.concatenate([c, np.zeros(pad, dtype=np.uint8)]).reshape(-1, 4)
    packed = (c4[:, 0] << 6) | (c4[:, 1] << 4) | (c4[:, 2] << 2) | c4[:, 3]
    head = struct.pack("<QQ", len(a), len(exc))
    return head + exc.astype("<u8").tobytes() + a[exc].tobytes() + packed.astype(np.uint8).tobytes()


def unpack_dna(blob: bytes) -> bytes:
    return unpack_dna_raw(unxz(blob))


def unpack_dna_raw(body: bytes) -> bytes:
    n, ne = struct.unpack_from("<QQ", body, 0)
    off = 16
    exc = np.frombuffer(body, dtype="<u8", count=ne, offset=off)
    off += 8 * ne
    excv = np.frombuffer(body, dtype=np.uint8, count=ne, offset=off)
    off += ne
    packed = np.frombuffer(body, dtype=np.uint8, offset=off)
    c = np.stack([(packed >> 6) & 3, (packed >> 4) & 3, (packed >> 2) & 3, packed & 3], axis=1).reshape(-1)[:n]
    out = np.frombuffer(b"ACGT", dtype=np.uint8)[c].copy()
    out[exc.astype(np.int64)] = excv
    return out.tobytes()


# ------------------------------------------------------------ block streams ---
class BlockStreams:
    """Pre-entropy streams of one block (the fields CURC writes).

    hq_text / lq_text / nlq_text are the full pseudogenomes.  When the
    self-match ran, rm_texts holds the three rewritten ('%'-marked) texts
    that the archive stores and maps the six MapOff / MapLen streams."""

    FIELDS = ("reads_count", "mis_table", "hq_text", "lq_text", "nlq_text", "strand", "read_off", "lq_off",
              "nlq_off", "pe_order", "id_pos", "mis_count", "mis_base", "mis_off", "max_mm", "maps", "rm_texts")

    def __init__(self, **kw):
        for f in self.FIELDS:
            setattr(self, f, kw.get(f))
        if self.maps is None:
            self.maps = {k: b"" for k in MAP_NAMES}

    @classmethod
    def from_block(cls, blk, n_reads, L, preserve, paired):
        """From a curc_amd.Block run with streams=True (or an oracle block
        after .streams(); both expose the same field names), with the
        self-match outputs when the block carries them."""
        max_mm = L // 5
        hq = blk.get("hq_pg").tobytes()
        bs = cls(
            reads_count=n_reads, mis_table=blk.get("s_mis_table").tobytes(), hq_text=hq,
            lq_text=blk.get("lq_pg").tobytes(), nlq_text=blk.get("nlq_pg").tobytes(),
            strand=blk.get("s_strand").tobytes(), read_off=blk.get("s_read_off"), lq_off=blk.get("s_lq_off"),
            nlq_off=blk.get("s_nlq_off"), pe_order=blk.get("s_pe_order"), id_pos=blk.get("s_id_pos"),
            mis_count=blk.get("s_mis_count"), mis_base=blk.get("s_mis_base"),
            mis_off={m: blk.mis_off(m) for m in range(1, max_mm + 1)}, max_mm=max_mm)
        if getattr(blk, "has_refmatch", False):
            bs.rm_texts = tuple(blk.get(f"rm_{p}_text").tobytes() for p in ("hq", "lq", "nlq"))
            bs.maps = {}
            for p, name in (("hq", "ref"), ("lq", "unmap_ref"), ("nlq", "unmap_n_ref")):
                bs.maps[name + "_off"] = blk.get(f"rm_{p}_off").tobytes()
                bs.maps[name + "_len"] = blk.get(f"rm_{p}_len").tobytes()
        return bs


def _read_frugal(buf, i):
    """helper.h:191-200; returns (value, next index)."""
    v = 0
    base = 1
    while True:
        b = buf[i] if i < len(buf) else 0
        i += 1
        v += base * (b % 128)
        base *= 128
        if b < 128:
            return v, i


def recover_ref(dest: bytes, src, map_off: bytes, map_len: bytes, ref_len_std: bool, src_is_dest: bool) -> bytes:
    """recoverRef (decompress.cpp:339-366): every '%' of dest expands to the
    reverse complement of src[off : off + len].  With src_is_dest the source
    is the text being rebuilt, whose earlier part each reference points to."""
    out = bytearray()
    srcb = out if src_is_dest else src
    if not map_len:
        return bytes(dest)
    target, li = _read_frugal(map_len, 0)
    osz = 4 if ref_len_std else 8
    ofmt = "<I" if ref_len_std else "<Q"
    oi = 0
    pos = 0
    while True:
        mark = dest.find(b"%", pos)
        if mark < 0:
            break
        out += dest[pos:mark]
        pos = mark + 1
        (s,) = struct.unpack_from(ofmt, map_off, oi)
        oi += osz
        n, li = _read_frugal(map_len, li)
        n += target
        out += bytes(srcb[s:s + n]).translate(RM_COMP)[::-1]
    out += dest[pos:]
    return bytes(out)


def _pe_order_encode(pe_order):
    """Pairs in emission order (paired_end_reads_order_compress,
    util.cuh:444-525, carries the same information): for each pair, first
    seen at emission index e, the distance to its mate's index and whether
    the first-seen read is the second mate."""
    pe_order = np.asarray(pe_order, dtype=np.int64)
    n = len(pe_order)
    idx = np.empty(n, dtype=np.int64)
    idx[pe_order] = np.arange(n)
    mate_idx = idx[pe_order ^ 1]
    first = mate_idx > np.arange(n)
    e = np.flatnonzero(first)
    delta = (mate_idx[e] - e).astype("<u4")
    second = (pe_order[e] & 1).astype(np.uint8)
    return delta.tobytes(), second.tobytes()


def _pe_order_decode(delta, second, n):
    """Emission indices in output order: even positions are first mates."""
    delta = np.frombuffer(delta, dtype="<u4").astype(np.int64)
    second = np.frombuffer(second, dtype=np.uint8)
    taken = np.zeros(n, dtype=bool)
    out = np.empty(n, dtype=np.int64)
    k = 0
    j = 0
    for i in range(n):
        if taken[i]:
            continue
        m = i + int(delta[k])
        taken[m] = True
        a, b = (m, i) if second[k] else (i, m)
        out[j], out[j + 1] = a, b
        j += 2
        k += 1
    return out


def pe_slots_from_order(pe_order):
    """Output slots of a paired block whose order is not preserved, from the
    global ids in emission order: pairs in the order their first read is
    emitted, each as (file 1 mate, file 2 mate); the value is the emission
    index (the same order as _pe_order_decode, computed without a loop)."""
    pe_order = np.asarray(pe_order, dtype=np.int64)
    n = len(pe_order)
    idx = np.empty(n, dtype=np.int64)
    idx[pe_order] = np.arange(n)
    e1, e2 = idx[0::2], idx[1::2]
    order = np.argsort(np.minimum(e1, e2), kind="stable")
    out = np.empty(n, dtype=np.int64)
    out[0::2], out[1::2] = e1[order], e2[order]
    return out


def pe_slots(bs, n):
    """Output slots of a paired block (see pe_slots_from_order): from the
    emission order when the streams carry it, else decoded from the
    archive's pair stream."""
    if getattr(bs, "pe_order", None) is not None and len(bs.pe_order):
        return pe_slots_from_order(bs.pe_order)
    return _pe_order_decode(bs.pe_pairs[0], bs.pe_pairs[1], n)


def block_parts(bs: BlockStreams, preserve: bool, paired: bool):
    """A block's fixed header bytes and its raw sub-file payloads, in file
    order (the u8 max_mm goes after the first len(head) of them)."""
    hdr = [struct.pack("<I", bs.reads_count), bytes(bs.mis_table)]
    joined = len(bs.hq_text) + len(bs.lq_text) + len(bs.nlq_text)
    if preserve:
        hdr.append(struct.pack("<B", 1 if joined <= 0xFFFFFFFF else 0))
    hdr.append(struct.pack("<B", 1 if len(bs.hq_text) <= 0xFFFFFFFF else 0))
    texts = bs.rm_texts if bs.rm_texts is not None else (bs.hq_text, bs.lq_text, bs.nlq_text)
    hdr.append(struct.pack("<QQQ", *(len(t) for t in texts)))
    head = []
    if preserve:
        dt = "<u4" if joined <= 0xFFFFFFFF else "<u8"
        head.append(np.asarray(bs.id_pos).astype(dt).tobytes())
    head.append(bytes(bs.strand))
    head.append(pack_dna_raw(b"".join(texts)))
    if not preserve:
        head.append(np.asarray(bs.read_off, dtype="<u2").tobytes())
        head.append(np.asarray(bs.nlq_off, dtype="<u2").tobytes())
        head.append(np.asarray(bs.lq_off, dtype="<u2").tobytes())
        if paired:
            d, s = _pe_order_encode(bs.pe_order)
            head.append(struct.pack("<Q", len(d)) + d + s)
    for k in MAP_NAMES:
        head.append(bytes(bs.maps[k]))
    head.append(np.asarray(bs.mis_count, dtype=np.uint8).tobytes())
    head.append(np.asarray(bs.mis_base, dtype=np.uint8).tobytes())
    tail = [np.asarray(bs.mis_off.get(m, np.zeros(0)), dtype="<u2").tobytes() for m in range(1, bs.max_mm + 1)]
    return b"".join(hdr), head, tail, bs.max_mm


def _write_parts(out, hdr, comp_head, comp_tail, max_mm):
    out.write(hdr)
    for c in comp_head:
        out.write(struct.pack("<Q", len(c)))
        out.write(c)
    out.write(struct.pack("<B", max_mm))
    for c in comp_tail:
        out.write(struct.pack("<Q", len(c)))
        out.write(c)


def write_block(out: io.BufferedIOBase, bs: BlockStreams, preserve: bool, paired: bool):
    hdr, head, tail, max_mm = block_parts(bs, preserve, paired)
    comp = xz_many(head + tail)  # all sub-files of the block compressed together
    _write_parts(out, hdr, comp[:len(head)], comp[len(head):], max_mm)


def read_block(inp: io.BufferedIOBase, preserve: bool, paired: bool, L: int) -> BlockStreams:
    def rd(fmt):
        return struct.unpack(fmt, inp.read(struct.calcsize(fmt)))

    (reads_count,) = rd("<I")
    mis_table = inp.read(16)
    if preserve:
        (join_std,) = rd("<B")
    (ref_len_std,) = rd("<B")
    hq_len, lq_len, n_len = rd("<QQQ")

    def sub():
        (n,) = rd("<Q")
        return inp.read(n)

    n_head = (1 if preserve else 0) + 2 + ((3 + (1 if paired else 0)) if not preserve else 0) + len(MAP_NAMES) + 2
    raw = [sub() for _ in range(n_head)]
    (max_mm,) = rd("<B")
    raw += [sub() for _ in range(max_mm)]
    dec = iter(unxz_many(raw))
    id_pos = None
    if preserve:
        dt = "<u4" if join_std else "<u8"
        id_pos = np.frombuffer(next(dec), dtype=dt).astype(np.uint64)
    strand = next(dec)
    text = unpack_dna_raw(next(dec))
    read_off = lq_off = nlq_off = pe_order = None
    pe_pairs = None
    if not preserve:
        read_off = np.frombuffer(next(dec), dtype="<u2")
        nlq_off = np.frombuffer(next(dec), dtype="<u2")
        lq_off = np.frombuffer(next(dec), dtype="<u2")
        if paired:
            pr = next(dec)
            (nd,) = struct.unpack_from("<Q", pr, 0)
            pe_pairs = (pr[8:8 + nd], pr[8 + nd:])
    maps = {}
    for k in MAP_NAMES:
        maps[k] = next(dec)
    mis_count = np.frombuffer(next(dec), dtype=np.uint8)
    mis_base = np.frombuffer(next(dec), dtype=np.uint8)
    mis_off = {m: np.frombuffer(next(dec), dtype="<u2") for m in range(1, max_mm + 1)}
    # decompress.cpp:602-608: the HQ text first (its references point into
    # itself), then the LQ and N-read LQ texts against it
    rm = (text[:hq_len], text[hq_len:hq_len + lq_len], text[hq_len + lq_len:])
    std = bool(ref_len_std)
    hq = recover_ref(rm[0], None, maps["ref_off"], maps["ref_len"], std, True)
    lq = recover_ref(rm[1], hq, maps["unmap_ref_off"], maps["unmap_ref_len"], std, False)
    nlq = recover_ref(rm[2], hq, maps["unmap_n_ref_off"], maps["unmap_n_ref_len"], std, False)
    bs = BlockStreams(reads_count=reads_count, mis_table=mis_table, hq_text=hq, lq_text=lq, nlq_text=nlq,
                      rm_texts=rm,
                      strand=strand, read_off=read_off, lq_off=lq_off, nlq_off=nlq_off, pe_order=pe_order,
                      id_pos=id_pos, mis_count=mis_count, mis_base=mis_base, mis_off=mis_off, max_mm=max_mm,
                      maps=maps)
    bs.pe_pairs = pe_pairs
    return bs


def write_archive(path, blocks, preserve, paired, L):
    if len(blocks) >= 255:  # preprocess.cpp:215-216
        raise ValueError("too many blocks")
    with open(path, "wb") as f:
        f.write(struct.pack("<BBHB", int(preserve), int(paired), L, len(blocks)))
        for bs in blocks:
            write_block(f, bs, preserve, paired)


def read_archive(path):
    with open(path, "rb") as f:
        preserve, paired, L, nb = struct.unpack("<BBHB", f.read(5))
        blocks = [read_block(f, bool(preserve), bool(paired), L) for _ in range(nb)]
    return bool(preserve), bool(paired), L, blocks


# ------------------------------------------------------------------ decode ---
def _offsets(stream, m, n_rec, L):
    """Mismatch positions (n_rec x m) from the offset stream for m mismatches
    (inverse of compress.cu:1848-1886)."""
    w = 2 if m == 1 else 3 if m == 2 else m + 2
    v = np.asarray(stream, dtype=np.int64).reshape(n_rec, w)
    flag, first = v[:, 0], v[:, 1]
    l2r = flag == 0
    if m == 1:
        return np.where(l2r, first, L - 1 - first)[:, None]
    if m == 2:
        gap = v[:, 2] + 1
        a = np.where(l2r, first, L - 1 - first - gap)
        return np.stack([a, a + gap], axis=1)
    mn = v[:, 2:3] + 1
    gaps = v[:, 3:] + mn
    cum = np.cumsum(gaps, axis=1)
    fwd = np.concatenate([first[:, None], first[:, None] + cum], axis=1)
    last = L - 1 - first
    bwd = np.concatenate([last[:, None], last[:, None] - cum], axis=1)[:, ::-1]
    return np.where(l2r[:, None], fwd, bwd)


def _apply_records(pg, pos, strand, mis_count, mis_base, mis_off, table, L):
    """Reads of the HQ/mapped records (compress.cu:1815-1846 inverted)."""
    n = len(pos)
    pga = np.frombuffer(pg, dtype=np.uint8)
    reads = pga[pos[:, None].astype(np.int64) + np.arange(L)[None, :]] if n else np.zeros((0, L), np.uint8)
    reads = reads.copy()
    if n and len(mis_base):
        row = np.zeros(256, dtype=np.int64)
        row[[ord("A"), ord("T"), ord("C"), ord("G")]] = [0, 1, 2, 3]
        tab = np.frombuffer(table, dtype=np.uint8).reshape(4, 4)
        base_idx = np.concatenate([[0], np.cumsum(mis_count.astype(np.int64))])
        for m in np.unique(mis_count):
            if m == 0:
                continue
            sel = np.flatnonzero(mis_count == m)
            offs = _offsets(mis_off[int(m)], int(m), len(sel), L)
            syms = mis_base[(base_idx[sel][:, None] + np.arange(m)[None, :]).reshape(-1)].reshape(len(sel), m)
            r = np.repeat(sel, m)
            o = offs.reshape(-1)
            ctx = reads[r, o]
            reads[r, o] = tab[row[ctx], syms.reshape(-1)]
    rc = np.frombuffer(strand, dtype=np.uint8) == ord("1")
    if rc.any():
        reads[rc] = COMP[reads[rc][:, ::-1]]
    return reads


def decode_block(bs: BlockStreams, L: int, preserve: bool, paired: bool):
    """Reads of one block in CURC's output order (decompress.cu:909-1005):
    without preserve_order HQ/mapped records in position order, then LQ,
    then N-LQ (pairs regrouped when paired); with it, by read id."""
    n_rec = len(bs.mis_count)
    if not preserve:
        pos = np.cumsum(bs.read_off.astype(np.int64))
        hq = _apply_records(bs.hq_text, pos, bs.strand, bs.mis_count, bs.mis_base, bs.mis_off, bs.mis_table, L)

        def chain(text, off):
            p = np.cumsum(L - off.astype(np.int64))
            a = np.frombuffer(text, dtype=np.uint8)
            return a[p[:, None] + np.arange(L)[None, :]] if len(p) else np.zeros((0, L), np.uint8)

        reads = np.concatenate([hq, chain(bs.lq_text, bs.lq_off), chain(bs.nlq_text, bs.nlq_off)])
        if paired:
            order = _pe_order_decode(bs.pe_pairs[0], bs.pe_pairs[1], len(reads))
            reads = reads[order]
        return reads
    # order preserved: id_pos gives each read's joined-pseudogenome position
    id_pos = bs.id_pos.astype(np.int64)
    hq_len, lq_len = len(bs.hq_text), len(bs.lq_text)
    joined = np.frombuffer(bs.hq_text + bs.lq_text + bs.nlq_text, dtype=np.uint8)
    out = joined[id_pos[:, None] + np.arange(L)[None, :]].copy()
    in_hq = id_pos < hq_len
    del lq_len
    hq_ids = np.flatnonzero(in_hq)
    assert len(hq_ids) == n_rec
    out[hq_ids] = _apply_records(bs.hq_text, id_pos[hq_ids], bs.strand, bs.mis_count, bs.mis_base, bs.mis_off,
                                 bs.mis_table, L)
    return out


def archive_bytes(blocks, preserve, paired, L):
    """Size of the .curc archive of the given BlockStreams."""
    return archive_sizes([(blocks, preserve, paired, L)])[0]


def archive_sizes(archives):
    """Sizes of several .curc archives, each given as (blocks, preserve,
    paired, L), with all their sub-files compressed in one parallel pass."""
    parts = [[block_parts(bs, pr, pa) for bs in blocks] for blocks, pr, pa, _ in archives]
    payloads = [p for arch in parts for (_, head, tail, _) in arch for p in head + tail]
    comp = iter(xz_many(payloads))
    sizes = []
    for arch in parts:
        size = 5  # u8 preserve, u8 paired, u16 L, u8 block count
        for hdr, head, tail, _ in arch:
            size += len(hdr) + 1
            for _ in range(len(head) + len(tail)):
                size += 8 + len(next(comp))
        sizes.append(size)
    return sizes
