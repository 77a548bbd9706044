"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {"st": os.path.join(HERE, "liboracle_st.so"), "omp": os.path.join(HERE, "liboracle_omp.so")}

FIELDS = {
    "packed": (0, np.uint64), "nonN_ids": (1, np.uint32), "N_ids": (2, np.uint32),
    "sorted_ids": (3, np.uint32), "next_dedup": (4, np.uint32), "next_overlap": (5, np.uint32),
    "prev_overlap": (6, np.uint32), "offset_overlap": (7, np.uint16), "next": (8, np.uint32),
    "prev": (9, np.uint32), "offset": (10, np.uint16), "hq_ids": (11, np.uint32),
    "unmatched_ids": (12, np.uint32), "hq_pg": (13, np.uint8), "hq_rec": (14, "rec"),
    "map_rec": (15, "rec"), "map_mis": (16, np.uint8), "unmapped": (17, np.uint32),
    "unmapped_N": (18, np.uint32), "lq_pg": (19, np.uint8), "lq_rec": (20, "rec"),
    "nlq_pg": (21, np.uint8), "nlq_rec": (22, "rec"), "prev_dedup": (23, np.uint32),
    "next_prune": (24, np.uint32), "prev_prune": (25, np.uint32), "hq_pg2": (26, np.uint64),
    "lq_next": (27, np.uint32), "lq_prev": (28, np.uint32), "lq_offset": (29, np.uint16),
    "s_strand": (40, np.uint8), "s_read_off": (41, np.uint16), "s_mis_count": (42, np.uint8),
    "s_mis_base": (43, np.uint8), "s_mis_table": (44, np.uint8), "s_lq_off": (45, np.uint16),
    "s_nlq_off": (46, np.uint16), "s_pe_order": (47, np.uint32), "s_id_pos": (48, np.uint64),
    "s_sorted_rec": (49, "rec"),
}
MIS_OFF_BASE = 64
REC_DTYPE = np.dtype([("read_id", "<u4"), ("pad", "<u4"), ("word", "<u8")])
STAGE = dict(dedup=1, overlap=2, prune=3, fullmatch=4, assembly=5, map=6, lq=7, nlq=8, all=8)


class OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("L", "W", "max_off", "k", "kip", "max_mm")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}


def lib(kind="st"):
    if kind in _libs:
        return _libs[kind]
    path = LIBS[kind]
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    u32, u64, vp, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int
    sig = {
        "orc_make_params": (i32, [u32, ctypes.POINTER(OrcParams)]),
        "orc_murmur64": (u64, [u64]),
        "orc_hash_size": (u32, [u64]),
        "orc_extract_word": (u64, [vp, u64, u64, u32]),
        "orc_extract_word_unaligned": (u64, [vp, u64, u32]),
        "orc_revcomp_kmer": (u64, [u64, u32]),
        "orc_mismatch_words": (i32, [u64, u64]),
        "orc_forward_mismatch": (i32, [vp, u64, vp, u32, u32]),
        "orc_rc_mismatch": (i32, [vp, u64, vp, u32, u32]),
        "orc_pack": (None, [vp, u64, u32, vp]),
        "orc_parse_fastq": (i32, [vp, u64, vp, u64, u32, vp, u64, ctypes.POINTER(u64)]),
        "orc_block_run": (vp, [vp, u64, u32, i32]),
        "orc_block_get": (u64, [vp, i32, ctypes.POINTER(vp)]),
        "orc_block_free": (None, [vp]),
        "orc_set_map_settle": (None, [i32]),
        "orc_map_settled_count": (u64, []),
        "orc_set_map_both": (None, [i32, i32]),
        "orc_block_streams": (i32, [vp, i32, i32]),
        "orc_marush36": (u32, [vp]),
        "orc_ref_match": (vp, [vp, u64, vp, u64, i32, i32]),
        "orc_refmatch_get": (u64, [vp, i32, ctypes.POINTER(vp)]),
        "orc_refmatch_free": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _libs[kind] = L
    return L


def params(L):
    p = OrcParams()
    if lib().orc_make_params(L, ctypes.byref(p)) != 0:
        raise ValueError(L)
    return p


class OracleBlock:
    def __init__(self, reads, stop="all", kind="st"):
        reads = np.ascontiguousarray(reads, dtype=np.uint8)
        n, L = reads.shape
        self.lib = lib(kind)
        self._keep = reads
        h = self.lib.orc_block_run(reads.ctypes.data, n, L, STAGE[stop])
        if not h:
            raise ValueError("oracle rejected the input")
        self._h = ctypes.c_void_p(h)
        self.kind = kind
        self.has_refmatch = False

    def refmatch(self):
        """The block's three self-match runs (compress.cu:2053-2069); results
        via get("rm_<hq|lq|nlq>_<text|off|len|matches|found>")."""
        self._rm = ref_match_block(self.get("hq_pg").tobytes(), self.get("lq_pg").tobytes(),
                                   self.get("nlq_pg").tobytes(), self.kind)
        self.has_refmatch = True

    def streams(self, preserve_order=False, paired_end=False):
        """Builds the block's pre-entropy streams (compress.cu:1730-2044)."""
        if self.lib.orc_block_streams(self._h, int(preserve_order), int(paired_end)) != 0:
            raise ValueError("mismatch recount disagrees with the mapping counts")

    def mis_off(self, m):
        return self._field(MIS_OFF_BASE + m, np.uint16)

    def _field(self, fid, dt):
        ptr = ctypes.c_void_p()
        n = self.lib.orc_block_get(self._h, fid, ctypes.byref(ptr))
        dt = np.dtype(dt)
        if n == 0 or not ptr.value:
            return np.zeros(0, dtype=dt)
        buf = (ctypes.c_char * (n * dt.itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=dt, count=n).copy()

    def get(self, name):
        if name.startswith("rm_"):
            _, part, what = name.split("_")
            r = self._rm[part]
            if what in ("matches", "found"):
                return r.get(what, np.zeros((0, 3), dtype=np.uint64))
            v = r[{"text": "text", "off": "map_off", "len": "map_len"}[what]]
            return np.frombuffer(v, dtype=np.uint8).copy()
        fid, dt = FIELDS[name]
        ptr = ctypes.c_void_p()
        n = self.lib.orc_block_get(self._h, fid, ctypes.byref(ptr))
        dt = np.dtype(REC_DTYPE if dt == "rec" else dt)
        if n == 0 or not ptr.value:
            return np.zeros(0, dtype=dt)
        buf = (ctypes.c_char * (n * dt.itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=dt, count=n).copy()

    def free(self):
        if getattr(self, "_h", None) is not None:
            self.lib.orc_block_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def ref_match(src, dest, dest_contain_n=False, dest_is_ref=False, kind="st"):
    """RefMatcher::match (compress.cu:855-994) of dest against src; returns
    dict(text, map_off, map_len, found, matches, capacity)."""
    src = bytes(src)
    dest = bytes(dest)
    L = lib(kind)
    h = L.orc_ref_match(src, len(src), dest, len(dest), int(dest_contain_n), int(dest_is_ref))
    try:
        def get(fid, dt):
            ptr = ctypes.c_void_p()
            n = L.orc_refmatch_get(h, fid, ctypes.byref(ptr))
            if dt is bytes:
                return ctypes.string_at(ptr.value, n) if n else b""
            if n == 0:
                return np.zeros((0, 3), dtype=np.uint64)
            buf = (ctypes.c_char * (n * 24)).from_address(ptr.value)
            return np.frombuffer(buf, dtype=np.uint64, count=3 * n).reshape(n, 3).copy()

        out = dict(text=get(0, bytes), map_off=get(1, bytes), map_len=get(2, bytes), found=get(3, None),
                   matches=get(4, None))
        ptr = ctypes.c_void_p()
        L.orc_refmatch_get(h, 5, ctypes.byref(ptr))
        out["capacity"] = ctypes.c_uint64.from_address(ptr.value).value
        return out
    finally:
        L.orc_refmatch_free(h)


def ref_match_block(hq, lq, nlq, kind="st"):
    """The three matches of block_compress (compress.cu:2053-2069), in its
    order: the LQ and N-read LQ pseudogenomes against the HQ one (skipped
    when empty), then the HQ pseudogenome against itself."""
    empty = dict(text=b"", map_off=b"", map_len=b"")
    out = {"lq": ref_match(hq, lq, False, False, kind) if len(lq) else dict(empty),
           "nlq": ref_match(hq, nlq, True, False, kind) if len(nlq) else dict(empty),
           "hq": ref_match(hq, hq, False, True, kind)}
    return out


def pack(reads):
    reads = np.ascontiguousarray(reads, dtype=np.uint8)
    n, L = reads.shape
    W = (L + 31) // 32
    out = np.zeros(n * W + 1, dtype=np.uint64)
    lib().orc_pack(reads.ctypes.data, n, L, out.ctypes.data)
    return out[: n * W].reshape(n, W)


PARSE_ERRORS = {-1: "read length isn't fix", -2: "truncated FASTQ record",
                -3: "paired-end files hold different record counts", -4: "too many records"}


def parse_fastq(fq1, L, fq2=None, kind="st"):
    """Reads of one block's FASTQ bytes (preprocess.cpp:147-193) as an
    (n, L) uint8 array in id order (paired: 2k from fq1, 2k+1 from fq2)."""
    a1 = np.frombuffer(fq1, dtype=np.uint8) if not isinstance(fq1, np.ndarray) else fq1
    a2 = None if fq2 is None else (np.frombuffer(fq2, dtype=np.uint8) if not isinstance(fq2, np.ndarray) else fq2)
    # a record takes at least L + 4 bytes (empty header and quality lines)
    cap = (len(a1) + (0 if a2 is None else len(a2))) // (L + 4) + 1
    out = np.empty((cap, L), dtype=np.uint8)
    n = ctypes.c_uint64()
    rc = lib(kind).orc_parse_fastq(a1.ctypes.data if len(a1) else None, len(a1),
                                   None if a2 is None else (a2.ctypes.data if len(a2) else None),
                                   0 if a2 is None else len(a2), L, out.ctypes.data, cap, ctypes.byref(n))
    if rc != 0:
        raise ValueError(PARSE_ERRORS.get(rc, f"parse error {rc}"))
    return out[: n.value]
