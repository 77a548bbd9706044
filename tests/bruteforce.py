"""Independent pure-Python statements of the CURC block rules, for tiny
inputs.  They avoid the oracle's machinery on purpose (no hash index, no
binary search, no suffix re-sorting, no merge join) so that a mistake in that
machinery shows up as a disagreement.

References: compress.cu:1396-1420 (dedup), 1473-1532 (greedy overlap),
1539-1585 (prune/partition), 1003-1174 (exact all-offset matching),
107-165 (assembly), 268-595 (mapping), util.cuh:82-89 (fmix64).
"""
INV = 0xFFFFFFFF
M64 = (1 << 64) - 1
CODE = {"A": 0, "C": 1, "G": 2, "T": 3}
COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def fmix64(x):
    x ^= x >> 33
    x = (x * 0xFF51AFD7ED558CCD) & M64
    x ^= x >> 33
    x = (x * 0xC4CEB9FE1A85EC53) & M64
    x ^= x >> 33
    return x


def unxorshift33(y):
    # x ^= x >> 33 is its own inverse for shifts >= 32
    return y ^ (y >> 33)


def fmix64_inverse(y):
    y = unxorshift33(y)
    y = (y * pow(0xC4CEB9FE1A85EC53, -1, 1 << 64)) & M64
    y = unxorshift33(y)
    y = (y * pow(0xFF51AFD7ED558CCD, -1, 1 << 64)) & M64
    return unxorshift33(y)


def kmer_value(s):
    """2-bit value of a base string, first base most significant; non-ACGT -> 0."""
    v = 0
    for ch in s:
        v = (v << 2) | CODE.get(ch.upper(), 0)
    return v


def revcomp_str(s):
    return "".join(COMP.get(c, "N") for c in reversed(s))


def hash_size(n):
    for i in range(24, 32):
        if (1 << i) >= n:
            return 1 << i
    return 1 << 31


def params(L):
    max_off = L * 65 // 100
    k = 32 if L >= 89 else 36 * L // 100
    return dict(L=L, max_off=max_off, k=k, kip=L - k - max_off, max_mm=L // 5)


def packed_string(r):
    """What a read looks like after 2-bit packing (non-ACGT becomes A)."""
    return "".join(c if c in CODE else ("A" if c.upper() not in CODE else c.upper()) for c in r)


def dedup(reads, nonN):
    order = sorted(nonN, key=lambda i: (packed_string(reads[i]), i))
    n = len(reads)
    nxt, prv, off = [INV] * n, [INV] * n, [0xFFFF] * n
    for a, b in zip(order, order[1:]):
        if packed_string(reads[a]) == packed_string(reads[b]):
            nxt[a], prv[b], off[a] = b, a, 0
    return order, nxt, prv, off


def serial_dictatorship(groups, nxt, prv, off, o):
    """groups: list of (suffixes in processing order, prefixes in preference
    order).  Each suffix takes the first free non-self prefix."""
    for sufs, pres in groups:
        for s in sufs:
            for p in pres:
                if p != s and prv[p] == INV:
                    nxt[s], prv[p], off[s] = p, s, o
                    break


def greedy_overlap(reads, order, nxt, prv, off, P, mutant=False):
    """Greedy overlap (compress.cu:1473-1532) as serial dictatorship over all
    pairs: suffixes in list order, candidates in lexicographic order."""
    L = P["L"]
    R = {i: packed_string(reads[i]) for i in order}
    prefixes = [i for i in order if prv[i] == INV]
    suffixes = [i for i in order if nxt[i] == INV]
    for o in range(1, P["max_off"]):
        for s in (reversed(suffixes) if mutant else suffixes):
            x = R[s][o:]
            cands = [p for p in prefixes if R[p][: L - o] == x]
            serial_dictatorship([([s], cands)], nxt, prv, off, o)
        suffixes = [s for s in suffixes if nxt[s] == INV]


def prune_partition(order, nxt, prv, off):
    n = len(nxt)
    res = [False] * n
    for r in range(n):
        res[r] = ((nxt[r] != INV and prv[r] != INV) or (nxt[r] != INV and off[r] == 0)
                  or (prv[r] != INV and off[prv[r]] == 0))
    for r in range(n):
        if not res[r]:
            if nxt[r] != INV:
                prv[nxt[r]] = INV
                nxt[r] = INV
            if prv[r] != INV:
                nxt[prv[r]] = INV
                prv[r] = INV
    return [i for i in order if res[i]], [i for i in order if not res[i]]


def full_match(seqs, ids, nxt, prv, off, L, alphabet="ACGT", mutant=False):
    """Exact all-offset matching (compress.cu:1003-1174) stated per group of
    equal overlap strings: prefixes in lexicographic order, suffixes ordered
    by s[off:] then by s[off-1], s[off-2], ..., s[0]; each suffix takes the
    first free non-self prefix."""
    rank = {c: i for i, c in enumerate(alphabet)}

    def key(s):
        return [rank[c] for c in s]

    lex = sorted(ids, key=lambda i: key(seqs[i]))
    for o in range(1, L):
        pres = [p for p in lex if prv[p] == INV]
        sufs = [s for s in lex if nxt[s] == INV]
        by_x = {}
        for s in sufs:
            by_x.setdefault(seqs[s][o:], []).append(s)
        groups = []
        for x, ss in by_x.items():
            ss = sorted(ss, key=lambda s: key(seqs[s][:o] if mutant else seqs[s][:o][::-1]))
            ps = [p for p in pres if seqs[p][: L - o] == x]
            groups.append((ss, ps))
        serial_dictatorship(groups, nxt, prv, off, o)


def assemble(seqs, visit, nxt, prv, off, L):
    """compress.cu:107-165: returns (text, [(read, pos)])."""
    seen = set()
    text = []
    recs = []
    pos = 0

    def emit(i, nb):
        nonlocal pos
        recs.append((i, pos))
        text.append(seqs[i][:nb])
        pos += nb

    for rid in visit:
        if rid in seen:
            continue
        i = rid
        while prv[i] != INV and prv[i] != rid:
            i = prv[i]
        if prv[i] == rid:
            cyc = []
            j = nxt[rid]
            while True:
                cyc.append(j)
                if j == rid:
                    break
                j = nxt[j]
            best = cyc[0]
            for j in cyc:
                if off[j] > off[best]:
                    best = j
            seen.update(cyc)
            s = nxt[best]
            while s != best:
                emit(s, off[s])
                s = nxt[s]
            emit(best, L)
        else:
            while nxt[i] != INV:
                seen.add(i)
                emit(i, off[i])
                i = nxt[i]
            seen.add(i)
            emit(i, L)
    return "".join(text), recs


def map_reads(pg, reads_fwd, P, k=28, step=5, rstep=2, limit=12, mutant=False):
    """compress.cu:362-521 restated as: candidates in (strand, read position,
    bucket slot) order; the answer is the first candidate minimising
    max(mismatches, 1), accepted if mismatches <= L/5.  reads_fwd: list of
    read strings (N kept).  Returns list of (strand, pos, mm) or None."""
    L = P["L"]
    n_idx = (len(pg) - k + 1 + step - 1) // step if len(pg) >= k - 1 else 0
    H = hash_size(n_idx)
    out = []
    buckets = []
    for strand in (0, 1):
        b = {}
        for i in range(n_idx):
            km = pg[i * step: i * step + k]
            if strand:
                km = revcomp_str(km)
            b.setdefault(fmix64(kmer_value(km)) & (H - 1), []).append(i)
        buckets.append({h: (v[:limit] if mutant else list(reversed(v[:limit]))) for h, v in b.items()})
    for r in reads_fwd:
        best = None
        for strand in (0, 1):
            q = r if strand == 0 else revcomp_str(r)
            for j in range(0, L - k + 1, rstep):
                h = fmix64(kmer_value(r[j:j + k])) & (H - 1)
                for i in buckets[strand].get(h, []):
                    rp = i * step
                    ls = (L - k - j) if strand else j
                    if rp < ls or rp - ls + L > len(pg):
                        continue
                    mp = rp - ls
                    mm = sum(1 for a, c in zip(pg[mp:mp + L], q) if a != c or c == "N")
                    cand = (max(mm, 1), strand, mp, mm)
                    if best is None or cand[0] < best[0]:
                        best = cand
        if best is not None and best[3] <= P["max_mm"]:
            out.append((best[1], best[2], best[3]))
        else:
            out.append(None)
    return out


# ---------------------------------------------------------------- RefMatcher ---
# compress.cu:597-1001.  Closed forms instead of the reference's loops: the
# seed's extensions are a longest common prefix / suffix, and the boundary
# rule of the left loop (it gives up the base at a text start) is written as
# an explicit case.
RM_COMP = {ord(a): ord(b) for a, b in zip("AaCcGgTtNnUuYyRrKkMmBbDdHhVv", "TTGGCCAANNAARRYYMMKKVVHHDDBB")}


def rm_comp(c):
    return RM_COMP.get(c, 0)


def rm_hash(s):
    """maRushPrime1 over 36 bytes (compress.cu:617-629)."""
    import struct
    h = 36
    for j, k in enumerate(struct.unpack("<9I", bytes(s[:36]))):
        h ^= (k + j) & 0xFFFFFFFF
        h = (h * 171717) & M64
    return h & 0xFFFFFFFF


def _lcp(a, i, b, j, limit):
    n = 0
    while n < limit and a[i + n] == b[j + n]:
        n += 1
    return n


def ref_match(src, dest, contain_n, is_ref, mutant=None):
    """Returns (text, map_off, map_len, found, matches).  mutant plants one
    plausible mistake: 'left' (no boundary rule on the left extension),
    'bucket' (ascending bucket order), 'segment' (one result per probe
    position on a binary destination), 'threshold' (49-base matches)."""
    src, dest = bytes(src), bytes(dest)
    n, m = len(src), len(dest)
    rc = bytes(rm_comp(c) for c in reversed(dest))
    text = bytes(rm_comp(c) for c in reversed(rc))
    if n < 36 or m < 36:
        return text, b"", b"", [], []
    cnt = (n - 35 + 4) // 5
    H = hash_size(cnt)
    buckets = {}
    for i in range(cnt):
        b = buckets.setdefault(rm_hash(src[5 * i:]) & (H - 1), [])
        if len(b) < 12:
            b.append(i)
    minlen = 49 if mutant == "threshold" else 50

    def probe(dp):
        cand = buckets.get(rm_hash(rc[dp:]) & (H - 1), [])
        for i in (cand if mutant == "bucket" else cand[::-1]):
            sp = 5 * i
            if not contain_n and is_ref and m - sp < dp:
                continue
            if rc[dp:dp + 36] != src[sp:sp + 36]:
                continue
            right = dp + 36 + _lcp(rc, dp + 36, src, sp + 36, min(m - dp - 36, n - sp - 36))
            b = min(dp, sp)
            left = 0
            while left < b and rc[dp - 1 - left] == src[sp - 1 - left]:
                left += 1
            start = dp - left if (left < b or mutant == "left") else dp - b + 1
            if right - start >= minlen:
                return (sp - (dp - start), right - start, start)
        return None

    dic = (m - 35 + 2) // 3
    found = []
    if not contain_n and mutant != "segment":
        for g in range((dic + 3) // 4):
            for s in range(4):
                dp = (4 * g + s) * 3
                if dp + 36 <= m:
                    r = probe(dp)
                    if r:
                        found.append(r)
                        break
    else:
        for i in range(dic):
            r = probe(3 * i)
            if r:
                found.append(r)
    if not found:
        return text, b"", b"", [], []
    cap = int(dic * (0.2 if is_ref else 0.21))
    res = []
    for s, ln, d in found[:cap]:
        d = m - (d + ln)
        if is_ref:
            s, d = min(s, d), max(s, d)
            if s + ln > d:
                mg = (s + ln - d + 1) // 2
                ln -= mg
                d += mg
        res.append((d, s, ln))
    res = sorted(set(res))
    matches = [(s, ln, d) for d, s, ln in res]

    def frugal(v):
        o = bytearray()
        while v >= 128:
            o.append(128 + v % 128)
            v //= 128
        o.append(v)
        return bytes(o)

    out, moff, mlen = bytearray(), bytearray(), bytearray(frugal(50))
    pos = 0
    for s, ln, d in matches:
        if d < pos:
            if pos - d >= ln:
                continue
            ln -= pos - d
            d = pos
        if ln < 50:
            continue
        out += text[pos:d] + b"%"
        moff += s.to_bytes(4 if n <= 0xFFFFFFFF else 8, "little")
        mlen += frugal(ln - 50)
        pos = d + ln
    out += text[pos:]
    return bytes(out), bytes(moff), bytes(mlen), found, matches
