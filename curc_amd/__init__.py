"""curc_amd: MI355X-native pseudogenome construction and read mapping for CURC.

Python mirror of the reference's block interface (src/compress.cuh) over the
C ABI in include/curc_gpu.h.  The compute runs in libcurc_gpu.so (HIP kernels
for gfx950); this module only moves arguments and results.  There is no CPU
fallback: if the library is missing or no GPU is visible, calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcurc_gpu.so")

CURC_OK = 0
STAGE_DEDUP, STAGE_OVERLAP, STAGE_PRUNE, STAGE_FULLMATCH, STAGE_ASSEMBLY, STAGE_MAP, STAGE_LQ, STAGE_NLQ = range(1, 9)
STAGE_ALL = STAGE_NLQ
RUN_INPUT_ON_DEVICE = 1
RUN_KEEP_STAGES = 2
RUN_STREAMS = 4
RUN_PRESERVE_ORDER = 8
RUN_PAIRED = 16
RUN_REFMATCH = 32
RUN_TRACE = 64
RUN_LEARNED = 128
MIS_OFF_BASE = 64
RM_BASE = 1000
RM_PARTS = {"hq": 0, "lq": 1, "nlq": 2}
RM_KINDS = {"text": 0, "off": 1, "len": 2, "matches": 3, "found": 4}

# field id -> numpy dtype (ids shared with the oracle)
FIELDS = {
    "packed": (0, np.uint64), "nonN_ids": (1, np.uint32), "N_ids": (2, np.uint32),
    "sorted_ids": (3, np.uint32), "next_dedup": (4, np.uint32), "next_overlap": (5, np.uint32),
    "prev_overlap": (6, np.uint32), "offset_overlap": (7, np.uint16), "next": (8, np.uint32),
    "prev": (9, np.uint32), "offset": (10, np.uint16), "hq_ids": (11, np.uint32),
    "unmatched_ids": (12, np.uint32), "hq_pg": (13, np.uint8), "hq_rec": (14, "rec"),
    "map_rec": (15, "rec"), "map_mis": (16, np.uint8), "unmapped": (17, np.uint32),
    "unmapped_N": (18, np.uint32), "lq_pg": (19, np.uint8), "lq_rec": (20, "rec"),
    "nlq_pg": (21, np.uint8), "nlq_rec": (22, "rec"), "prev_dedup": (23, np.uint32), "next_prune": (24, np.uint32), "prev_prune": (25, np.uint32),
    "hq_pg2": (26, np.uint64), "lq_next": (27, np.uint32), "lq_prev": (28, np.uint32),
    "lq_offset": (29, np.uint16), "N_ascii": (30, np.uint8),
    "s_strand": (40, np.uint8), "s_read_off": (41, np.uint16), "s_mis_count": (42, np.uint8),
    "s_mis_base": (43, np.uint8), "s_mis_table": (44, np.uint8), "s_lq_off": (45, np.uint16),
    "s_nlq_off": (46, np.uint16), "s_pe_order": (47, np.uint32), "s_id_pos": (48, np.uint64),
    "s_sorted_rec": (49, "rec"),
    "trace_x": (60, np.dtype((np.float32, 32))), "trace_y": (61, np.float32),
}

# curc_alignment: u32 read_id, u32 pad, u64 word (isN | strand << 1 | pos << 2)
REC_DTYPE = np.dtype([("read_id", "<u4"), ("pad", "<u4"), ("word", "<u8")])


class CurcError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("L", "W", "max_off", "k", "kip", "max_mm")]


class Job(ctypes.Structure):
    _fields_ = [("ascii", ctypes.c_void_p), ("n_reads", ctypes.c_uint64), ("L", ctypes.c_uint32),
                ("flags", ctypes.c_uint32)]


class FastqJob(ctypes.Structure):
    _fields_ = [("fq1", ctypes.c_void_p), ("bytes1", ctypes.c_uint64), ("fq2", ctypes.c_void_p),
                ("bytes2", ctypes.c_uint64), ("L", ctypes.c_uint32), ("flags", ctypes.c_uint32)]


class DecodeJob(ctypes.Structure):
    _fields_ = [("L", ctypes.c_uint32), ("flags", ctypes.c_uint32), ("n_reads", ctypes.c_uint64),
                ("hq", ctypes.c_void_p), ("hq_len", ctypes.c_uint64), ("lq", ctypes.c_void_p),
                ("lq_len", ctypes.c_uint64), ("nlq", ctypes.c_void_p), ("nlq_len", ctypes.c_uint64),
                ("n_rec", ctypes.c_uint64), ("strand", ctypes.c_void_p), ("mis_count", ctypes.c_void_p),
                ("mis_base", ctypes.c_void_p), ("n_mis", ctypes.c_uint64), ("mis_table", ctypes.c_void_p),
                ("max_mm", ctypes.c_uint32), ("mis_off", ctypes.c_void_p), ("mis_off_len", ctypes.c_void_p),
                ("read_off", ctypes.c_void_p), ("lq_off", ctypes.c_void_p), ("n_lq", ctypes.c_uint64),
                ("nlq_off", ctypes.c_void_p), ("n_nlq", ctypes.c_uint64), ("pe_slot", ctypes.c_void_p),
                ("id_pos", ctypes.c_void_p)]


BLOCK_CB = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int)


class Tuning(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in ("ov_grid", "fm_grid", "map_grid", "radix_items", "workers",
                                               "filter_bits", "fm_fused", "ov_queue", "map_both",
                                               "fm_static", "fq_host")]


class StageStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("ms", ctypes.c_double), ("launches", ctypes.c_uint64),
                ("alg_bytes", ctypes.c_double), ("items", ctypes.c_double)]


class SynthCfg(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64), ("genome_len", ctypes.c_uint64), ("gc", ctypes.c_double),
        ("repeat_frac", ctypes.c_double), ("rep_min", ctypes.c_uint32), ("rep_max", ctypes.c_uint32),
        ("div_min", ctypes.c_double), ("div_max", ctypes.c_double), ("rep_families", ctypes.c_uint32),
        ("n_reads", ctypes.c_uint64), ("L", ctypes.c_uint32), ("paired", ctypes.c_int32),
        ("frag_mean", ctypes.c_double), ("frag_sd", ctypes.c_double), ("rc_prob", ctypes.c_double),
        ("sub_first", ctypes.c_double), ("sub_last", ctypes.c_double), ("n_rate", ctypes.c_double),
    ]


_lib = None


def lib():
    """Loads libcurc_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CurcError(f"{LIB_PATH} is missing; run curc_amd/build.py (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, u64, i32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    sig = {
        "curc_make_params": (i32, [u32, ctypes.POINTER(Params)]),
        "curc_device_count": (i32, [ctypes.POINTER(i32)]),
        "curc_ctx_create": (vp, [i32]),
        "curc_ctx_destroy": (None, [vp]),
        "curc_ctx_last_error": (ctypes.c_char_p, [vp]),
        "curc_ctx_trim": (i32, [vp]),
        "curc_host_alloc": (i32, [vp, ctypes.POINTER(vp), u64]),
        "curc_host_free": (i32, [vp, vp]),
        "curc_dev_alloc": (i32, [vp, ctypes.POINTER(vp), u64]),
        "curc_dev_free": (i32, [vp, vp]),
        "curc_memcpy_h2d": (i32, [vp, vp, vp, u64]),
        "curc_memcpy_d2h": (i32, [vp, vp, vp, u64]),
        "curc_sync": (i32, [vp]),
        "curc_stream_copy": (i32, [vp, vp, vp, u64, i32, ctypes.POINTER(ctypes.c_float)]),
        "curc_device_sync": (i32, [vp]),
        "curc_pack": (i32, [vp, vp, u64, u32, vp, vp]),
        "curc_block_run": (i32, [vp, vp, u64, u32, i32, u32, ctypes.POINTER(vp)]),
        "curc_batch_run": (i32, [vp, vp, u32, u32, ctypes.POINTER(vp), ctypes.POINTER(i32)]),
        "curc_ref_match": (i32, [vp, vp, u64, vp, u64, i32, i32, u32, ctypes.POINTER(vp)]),
        "curc_ctx_set_scorer": (i32, [vp, vp, ctypes.c_float, ctypes.c_float]),
        "curc_ctx_set_trace": (i32, [vp, u32, u64]),
        "curc_ctx_get_tuning": (i32, [vp, ctypes.POINTER(Tuning)]),
        "curc_ctx_set_tuning": (i32, [vp, ctypes.POINTER(Tuning)]),
        "curc_mlp_create": (vp, [vp]),
        "curc_mlp_destroy": (None, [vp]),
        "curc_mlp_set_params": (i32, [vp, vp, u32]),
        "curc_mlp_get_params": (i32, [vp, vp, u32]),
        "curc_mlp_forward": (i32, [vp, vp, u64, vp]),
        "curc_mlp_gradients": (i32, [vp, vp, vp, u64, vp]),
        "curc_mlp_train_step": (i32, [vp, vp, vp, u64, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
        "curc_fastq_block_run": (i32, [vp, ctypes.POINTER(FastqJob), i32, ctypes.POINTER(vp)]),
        "curc_fastq_batch_run": (i32, [vp, vp, u32, u32, ctypes.POINTER(vp), ctypes.POINTER(i32)]),
        "curc_fastq_stream_run": (i32, [vp, vp, u32, BLOCK_CB, vp]),
        "curc_fastq_pack": (i32, [vp, vp, u64, u32, u64, vp, vp, vp, ctypes.POINTER(u64)]),
        "curc_fastq_pack_static": (i32, [vp, vp, u64, u32, u64, vp, vp, vp, ctypes.POINTER(u64)]),
        "curc_block_reads": (u64, [vp]),
        "curc_decode_block": (i32, [vp, ctypes.POINTER(DecodeJob), vp, vp]),
        "curc_block_get": (u64, [vp, i32, ctypes.POINTER(vp)]),
        "curc_block_free": (None, [vp]),
        "curc_ctx_set_timing": (i32, [vp, i32]),
        "curc_ctx_stats": (i32, [vp, ctypes.POINTER(StageStat), i32]),
        "curc_ctx_reset_stats": (None, [vp]),
        "curc_synth_default": (None, [ctypes.POINTER(SynthCfg)]),
        "curc_synth_genome": (i32, [ctypes.POINTER(SynthCfg), vp]),
        "curc_synth_reads": (i32, [ctypes.POINTER(SynthCfg), vp, vp]),
        "curc_synth_fastq_record_len": (u64, [u32, ctypes.c_char_p, i32]),
        "curc_synth_fastq": (i32, [vp, u64, u32, u64, u64, ctypes.c_char_p, i32, vp]),
        "curc_fastq_host_cap": (u64, [u64, u32, u32]),
        "curc_fastq_host_parse": (i32, [vp, u64, u32, u32, vp, u64, ctypes.POINTER(u64), ctypes.POINTER(i32)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def make_params(L):
    p = Params()
    if lib().curc_make_params(L, ctypes.byref(p)) != CURC_OK:
        raise CurcError(f"unsupported read length {L}")
    return p


def device_count():
    n = ctypes.c_int(0)
    lib().curc_device_count(ctypes.byref(n))
    return n.value


# ------------------------------------------------------------ host parse ---
FASTQ_FORMAT_ERRORS = {1: "FASTQ block does not start with '@'", 2: "truncated FASTQ record",
                       3: "read length isn't fix"}


def fastq_host_parse(buf, L, threads=8):
    """The library's host-side FASTQ record parse (curc_fastq_host_parse)
    of one file's bytes -> (n, L) uint8 sequences; a format error raises
    CurcError with the device parser's message."""
    a = np.frombuffer(buf, dtype=np.uint8) if not isinstance(buf, np.ndarray) else buf
    cap = int(lib().curc_fastq_host_cap(len(a), L, threads))
    out = np.empty(max(cap, 1), dtype=np.uint8)
    n = ctypes.c_uint64(0)
    kind = ctypes.c_int(0)
    rc = lib().curc_fastq_host_parse(a.ctypes.data if len(a) else None, len(a), L, threads, out.ctypes.data, cap,
                                     ctypes.byref(n), ctypes.byref(kind))
    if rc != CURC_OK:
        raise CurcError(FASTQ_FORMAT_ERRORS.get(kind.value, f"curc_fastq_host_parse failed ({rc})"))
    return out[:n.value * L].reshape(n.value, L).copy()


# ----------------------------------------------------------------- synth ---
def synth_config(**kw):
    c = SynthCfg()
    lib().curc_synth_default(ctypes.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def synth_genome(cfg):
    g = np.empty(cfg.genome_len, dtype=np.uint8)
    rc = lib().curc_synth_genome(ctypes.byref(cfg), g.ctypes.data)
    if rc != CURC_OK:
        raise CurcError(f"curc_synth_genome failed ({rc})")
    return g


def synth_reads(cfg, genome=None):
    """Returns (genome, reads) with reads an (n_reads, L) uint8 array."""
    if genome is None:
        genome = synth_genome(cfg)
    out = np.empty((cfg.n_reads, cfg.L), dtype=np.uint8)
    rc = lib().curc_synth_reads(ctypes.byref(cfg), genome.ctypes.data, out.ctypes.data)
    if rc != CURC_OK:
        raise CurcError(f"curc_synth_reads failed ({rc})")
    return genome, out


def synth_fastq(reads, out=None, first_id=0, mate=0, stride_reads=1, prefix=b"synth."):
    """FASTQ records of reads ((n, L) uint8; every stride_reads-th row, so
    stride_reads=2 writes one mate of interleaved pairs) in the layout of
    fastq.fastq_bytes, written in parallel into out (a uint8 array, e.g. a
    pinned HostBuffer's) or a new array; returns the filled array."""
    reads = np.ascontiguousarray(reads, dtype=np.uint8)
    total, L = reads.shape
    n = (total + stride_reads - 1) // stride_reads if total else 0
    rec = lib().curc_synth_fastq_record_len(L, prefix, mate)
    if out is None:
        out = np.empty(n * rec, dtype=np.uint8)
    if len(out) < n * rec:
        raise ValueError("output buffer too small")
    rc = lib().curc_synth_fastq(reads.ctypes.data if n else None, n, L, L * stride_reads, first_id, prefix, mate,
                                out.ctypes.data if n else None)
    if rc != CURC_OK:
        raise CurcError(f"curc_synth_fastq failed ({rc})")
    return out[: n * rec]


def synth_fastq_record_len(L, mate=0, prefix=b"synth."):
    return int(lib().curc_synth_fastq_record_len(L, prefix, mate))


# workload presets (sizes as the reference's headline configurations)
def config_small(seed=11):
    """1 Mbp uniform genome, 133,333 x 150 bp SE, 50% RC, 0.5% subst, 0.01% N."""
    return synth_config(seed=seed, genome_len=1_000_000, gc=0.5, n_reads=133_333, L=150, rc_prob=0.5,
                        sub_first=0.005, sub_last=0.005, n_rate=0.0001)


def config_scale(seed=22, n_reads=20_000_000, genome_len=100_000_000):
    """100 Mbp genome (GC 41%), 20M x 150 bp SE, subst 0.1%->1.5% along the read, 0.02% N."""
    return synth_config(seed=seed, genome_len=genome_len, gc=0.41, n_reads=n_reads, L=150, rc_prob=0.5,
                        sub_first=0.001, sub_last=0.015, n_rate=0.0002)


# --------------------------------------------------------------- context ---
class Block:
    """Results of one block (host memory owned by the library)."""

    def __init__(self, handle, has_refmatch=False):
        self._h = ctypes.c_void_p(handle)
        self.has_refmatch = has_refmatch

    @property
    def n_reads(self):
        """Reads in the block (the parsed record count for FASTQ input)."""
        return int(lib().curc_block_reads(self._h))

    def mis_off(self, m):
        """Mismatch-offset stream for reads with m mismatches."""
        return self._field(MIS_OFF_BASE + m, np.uint16)

    def get(self, name):
        if name.startswith("rm_"):
            # rm_<hq|lq|nlq>_<text|off|len|matches|found>: self-match outputs
            _, part, kind = name.split("_")
            k = RM_KINDS[kind]
            fid = RM_BASE + 5 * RM_PARTS[part] + k
            if k >= 3:
                return self._field(fid, np.dtype((np.uint64, 3)))
            return self._field(fid, np.uint8)
        fid, dt = FIELDS[name]
        return self._field(fid, dt)

    def _field(self, fid, dt):
        ptr = ctypes.c_void_p()
        n = lib().curc_block_get(self._h, fid, ctypes.byref(ptr))
        if dt == "rec":
            dt = REC_DTYPE
        dt = np.dtype(dt)
        if n == 0 or not ptr.value:
            return np.zeros((0,) + dt.shape, dtype=dt.base)
        buf = (ctypes.c_char * (n * dt.itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=dt, count=n).copy()

    def free(self):
        if self._h is not None and self._h.value:
            lib().curc_block_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostBuffer:
    """Pinned host memory (curc_host_alloc) viewed as a uint8 array."""

    def __init__(self, ctx, nbytes):
        self._ctx = ctx
        p = ctypes.c_void_p()
        ctx._check(lib().curc_host_alloc(ctx._h, ctypes.byref(p), max(1, nbytes)), "curc_host_alloc")
        self.ptr = p.value
        self.nbytes = nbytes
        self.array = np.frombuffer((ctypes.c_char * max(1, nbytes)).from_address(self.ptr), dtype=np.uint8,
                                   count=nbytes)

    def free(self):
        if self.ptr:
            self.array = None
            lib().curc_host_free(self._ctx._h, ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _host_ptr(buf):
    """(address, length, keep-alive) of FASTQ bytes given as bytes, a uint8
    array or a HostBuffer."""
    if buf is None:
        return None, 0, None
    if isinstance(buf, HostBuffer):
        return buf.ptr, buf.nbytes, buf
    a = np.frombuffer(buf, dtype=np.uint8) if isinstance(buf, (bytes, bytearray, memoryview)) else \
        np.ascontiguousarray(buf, dtype=np.uint8)
    return (a.ctypes.data if a.size else None), a.size, a


def _run_flags(keep_stages=False, streams=False, preserve_order=False, paired=False, refmatch=False, trace=False,
               learned=False):
    flags = RUN_KEEP_STAGES if keep_stages else 0
    flags |= (RUN_TRACE if trace else 0) | (RUN_LEARNED if learned else 0)
    if streams:
        flags |= RUN_STREAMS | (RUN_PRESERVE_ORDER if preserve_order else 0) | (RUN_PAIRED if paired else 0)
    if refmatch:
        flags |= RUN_REFMATCH
    return flags


class Context:
    """One device context (stream + cached device scratch)."""

    def __init__(self, device=0):
        h = lib().curc_ctx_create(device)
        if not h:
            raise CurcError(f"curc_ctx_create({device}) failed: no usable HIP device")
        self._h = ctypes.c_void_p(h)
        self._dependents = []  # objects holding device memory of this context (scorer networks)

    def _check(self, rc, what):
        if rc != CURC_OK:
            msg = lib().curc_ctx_last_error(self._h)
            raise CurcError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def run_block(self, reads, L=None, stop_after=STAGE_ALL, keep_stages=False, device_ptr=None, n_reads=None,
                  streams=False, preserve_order=False, paired=False, refmatch=False, trace=False, learned=False):
        """Runs the GPU part of block_compress on reads ((n, L) uint8 host
        array) or on a device buffer (device_ptr, n_reads, L); with streams,
        also the record sort and stream builders; with refmatch, the
        pseudogenome self-match; with trace, scorer training records; with
        learned, mapping under the context's scorer (set_scorer)."""
        flags = RUN_KEEP_STAGES if keep_stages else 0
        flags |= (RUN_TRACE if trace else 0) | (RUN_LEARNED if learned else 0)
        if streams:
            flags |= RUN_STREAMS | (RUN_PRESERVE_ORDER if preserve_order else 0) | (RUN_PAIRED if paired else 0)
        if refmatch:
            flags |= RUN_REFMATCH
        out = ctypes.c_void_p()
        if device_ptr is not None:
            flags |= RUN_INPUT_ON_DEVICE
            rc = lib().curc_block_run(self._h, ctypes.c_void_p(device_ptr), n_reads, L, stop_after, flags,
                                      ctypes.byref(out))
        else:
            reads = np.ascontiguousarray(reads, dtype=np.uint8)
            n, L = reads.shape
            rc = lib().curc_block_run(self._h, reads.ctypes.data, n, L, stop_after, flags, ctypes.byref(out))
        self._check(rc, "curc_block_run")
        return Block(out.value, has_refmatch=refmatch)

    def batch_run(self, inputs, workers=0, keep_stages=False, streams=False, preserve_order=False, paired=False,
                  refmatch=False, learned=False):
        """Runs independent blocks together (curc_batch_run): inputs is a
        list of (n, L) uint8 host arrays (lengths may differ); returns one
        Block per input, in input order.  With learned, every input maps
        under the context's scorer."""
        arrs = [np.ascontiguousarray(r, dtype=np.uint8) for r in inputs]
        flags = RUN_KEEP_STAGES if keep_stages else 0
        flags |= RUN_LEARNED if learned else 0
        if streams:
            flags |= RUN_STREAMS | (RUN_PRESERVE_ORDER if preserve_order else 0) | (RUN_PAIRED if paired else 0)
        if refmatch:
            flags |= RUN_REFMATCH
        n = len(arrs)
        jobs = (Job * max(1, n))()
        for i, a in enumerate(arrs):
            jobs[i] = Job(a.ctypes.data, a.shape[0], a.shape[1], flags)
        out = (ctypes.c_void_p * max(1, n))()
        status = (ctypes.c_int32 * max(1, n))()
        rc = lib().curc_batch_run(self._h, jobs, n, workers, out, status)
        blocks = [Block(out[i], has_refmatch=refmatch) if out[i] else None for i in range(n)]
        if rc != CURC_OK:
            for b in blocks:
                if b is not None:
                    b.free()
            self._check(rc, "curc_batch_run")
        return blocks

    def _fastq_job(self, fq1, fq2, L, flags, on_device=False, nbytes=None):
        keep = []
        if on_device:
            a1, n1 = fq1, nbytes[0]
            a2, n2 = (fq2, nbytes[1]) if fq2 is not None else (None, 0)
            flags |= RUN_INPUT_ON_DEVICE
        else:
            a1, n1, k1 = _host_ptr(fq1)
            a2, n2, k2 = _host_ptr(fq2)
            keep += [k1, k2]
        if fq2 is not None:
            flags |= RUN_PAIRED
        return FastqJob(a1, n1, a2, n2, L, flags), keep

    def fastq_block(self, fq1, fq2=None, L=None, stop_after=STAGE_ALL, on_device=False, nbytes=None, **kw):
        """One block from FASTQ bytes (curc_fastq_block_run): fq1 / fq2 are
        the block's bytes of file 1 / file 2 (bytes, uint8 arrays or
        HostBuffers), or device pointers with on_device=True and
        nbytes=(bytes1, bytes2).  L is the input's read length.  Keywords as
        run_block (keep_stages, streams, preserve_order, refmatch, ...)."""
        if L is None:
            raise ValueError("L (the input's read length) is required")
        job, keep = self._fastq_job(fq1, fq2, L, _run_flags(**kw), on_device, nbytes)
        out = ctypes.c_void_p()
        rc = lib().curc_fastq_block_run(self._h, ctypes.byref(job), stop_after, ctypes.byref(out))
        del keep
        self._check(rc, "curc_fastq_block_run")
        return Block(out.value, has_refmatch=bool(kw.get("refmatch")))

    def fastq_batch(self, inputs, L, workers=0, **kw):
        """Blocks from FASTQ bytes together (curc_fastq_batch_run); inputs is
        a list of fq1 or (fq1, fq2) host buffers.  Returns one Block each."""
        flags = _run_flags(**kw)
        n = len(inputs)
        jobs = (FastqJob * max(1, n))()
        keep = []
        for i, x in enumerate(inputs):
            f1, f2 = (x if isinstance(x, tuple) else (x, None))
            jobs[i], k = self._fastq_job(f1, f2, L, flags)
            keep += k
        out = (ctypes.c_void_p * max(1, n))()
        status = (ctypes.c_int32 * max(1, n))()
        rc = lib().curc_fastq_batch_run(self._h, jobs, n, workers, out, status)
        blocks = [Block(out[i], has_refmatch=bool(kw.get("refmatch"))) if out[i] else None for i in range(n)]
        if rc != CURC_OK:
            for b in blocks:
                if b is not None:
                    b.free()
            self._check(rc, "curc_fastq_batch_run")
        return blocks

    def fastq_stream(self, inputs, L, on_block, **kw):
        """Blocks from FASTQ host buffers in order, each next block's bytes
        copied to the device while the current one runs
        (curc_fastq_stream_run).  on_block(i, block) is called for each
        block in order and owns it (block.free())."""
        flags = _run_flags(**kw)
        n = len(inputs)
        jobs = (FastqJob * max(1, n))()
        keep = []
        for i, x in enumerate(inputs):
            f1, f2 = (x if isinstance(x, tuple) else (x, None))
            jobs[i], k = self._fastq_job(f1, f2, L, flags)
            keep += k
        errors = []

        def cb(_user, i, handle, status):
            if status != CURC_OK or not handle:
                return
            try:
                on_block(int(i), Block(handle, has_refmatch=bool(kw.get("refmatch"))))
            except Exception as e:  # raised again after the run
                errors.append(e)

        fn = BLOCK_CB(cb)
        rc = lib().curc_fastq_stream_run(self._h, jobs, n, fn, None)
        del keep
        self._check(rc, "curc_fastq_stream_run")
        if errors:
            raise errors[0]

    def fastq_pack(self, d_fq, nbytes, L, max_reads, d_packed, d_isN, d_seq_off):
        """Record parse + N flag + 2-bit pack of device FASTQ bytes
        (curc_fastq_pack); returns the record count."""
        n = ctypes.c_uint64()
        self._check(lib().curc_fastq_pack(self._h, ctypes.c_void_p(d_fq), nbytes, L, max_reads,
                                          ctypes.c_void_p(d_packed), ctypes.c_void_p(d_isN) if d_isN else None,
                                          ctypes.c_void_p(d_seq_off), ctypes.byref(n)), "curc_fastq_pack")
        return int(n.value)

    def decode(self, bs, L, preserve, paired, pe_slot=None):
        """Read reconstruction of one block on the GPU (curc_decode_block)
        from its streams (archive.BlockStreams, texts after the '%'
        expansion).  Returns the reads as an (n, L) uint8 array in output
        order (paired: file 1's line k is row 2k, file 2's row 2k+1)."""
        from . import archive

        keep = []

        def arr(x, dt):
            a = np.ascontiguousarray(np.frombuffer(x, dtype=dt) if isinstance(x, (bytes, bytearray)) else x,
                                     dtype=dt)
            keep.append(a)
            return a

        def ptr(a):
            return a.ctypes.data if a.size else None

        n = int(bs.reads_count)
        hq, lq, nlq = arr(bs.hq_text, np.uint8), arr(bs.lq_text, np.uint8), arr(bs.nlq_text, np.uint8)
        strand, cnt = arr(bs.strand, np.uint8), arr(bs.mis_count, np.uint8)
        base, tab = arr(bs.mis_base, np.uint8), arr(bs.mis_table, np.uint8)
        max_mm = int(bs.max_mm)
        offs = [arr(bs.mis_off.get(m, np.zeros(0, np.uint16)), np.uint16) for m in range(1, max_mm + 1)]
        off_ptrs = (ctypes.c_void_p * max(1, max_mm))(*[ptr(o) for o in offs])
        off_lens = np.array([len(o) for o in offs] or [0], dtype=np.uint64)
        j = DecodeJob(L=L, flags=(RUN_PRESERVE_ORDER if preserve else 0) | (RUN_PAIRED if paired else 0),
                      n_reads=n, hq=ptr(hq), hq_len=hq.size, lq=ptr(lq), lq_len=lq.size, nlq=ptr(nlq),
                      nlq_len=nlq.size, n_rec=cnt.size, strand=ptr(strand), mis_count=ptr(cnt), mis_base=ptr(base),
                      n_mis=base.size, mis_table=ptr(tab), max_mm=max_mm,
                      mis_off=ctypes.cast(off_ptrs, ctypes.c_void_p), mis_off_len=off_lens.ctypes.data)
        if preserve:
            ip = arr(bs.id_pos, np.uint64)
            j.id_pos = ptr(ip)
        else:
            ro, lo, no = arr(bs.read_off, np.uint16), arr(bs.lq_off, np.uint16), arr(bs.nlq_off, np.uint16)
            j.read_off, j.lq_off, j.n_lq, j.nlq_off, j.n_nlq = ptr(ro), ptr(lo), lo.size, ptr(no), no.size
            if paired:
                ps = arr(pe_slot if pe_slot is not None else archive.pe_slots(bs, n), np.uint32)
                j.pe_slot = ptr(ps)
        lines = n // 2 if paired else n
        o1 = np.empty((lines, L + 1), dtype=np.uint8)
        o2 = np.empty((lines, L + 1), dtype=np.uint8) if paired else None
        self._check(lib().curc_decode_block(self._h, ctypes.byref(j), o1.ctypes.data,
                                            o2.ctypes.data if paired else None), "curc_decode_block")
        if not paired:
            return o1[:, :L]
        out = np.empty((n, L), dtype=np.uint8)
        out[0::2], out[1::2] = o1[:, :L], o2[:, :L]
        return out

    def fastq_pack_static(self, d_fq, nbytes, L, max_reads, d_packed, d_isN, d_seq_off):
        """As fastq_pack, through the entry that allocates nothing
        (curc_fastq_pack_static)."""
        n = ctypes.c_uint64()
        self._check(lib().curc_fastq_pack_static(self._h, ctypes.c_void_p(d_fq), nbytes, L, max_reads,
                                                 ctypes.c_void_p(d_packed),
                                                 ctypes.c_void_p(d_isN) if d_isN else None,
                                                 ctypes.c_void_p(d_seq_off), ctypes.byref(n)),
                    "curc_fastq_pack_static")
        return int(n.value)

    def host_buffer(self, nbytes):
        """Pinned host memory of nbytes (HostBuffer)."""
        return HostBuffer(self, nbytes)

    def ref_match(self, src, dest, dest_contain_n=False, dest_is_ref=False, keep_stages=False):
        """RefMatcher of dest against src (bytes); returns dict(text,
        map_off, map_len[, found, matches])."""
        src, dest = bytes(src), bytes(dest)
        out = ctypes.c_void_p()
        rc = lib().curc_ref_match(self._h, src, len(src), dest, len(dest), int(dest_contain_n), int(dest_is_ref),
                                  RUN_KEEP_STAGES if keep_stages else 0, ctypes.byref(out))
        self._check(rc, "curc_ref_match")
        b = Block(out.value, has_refmatch=True)
        res = dict(text=b.get("rm_hq_text").tobytes(), map_off=b.get("rm_hq_off").tobytes(),
                   map_len=b.get("rm_hq_len").tobytes())
        if keep_stages:
            res["found"] = b.get("rm_hq_found")
            res["matches"] = b.get("rm_hq_matches")
        b.free()
        return res

    def set_scorer(self, mlp, threshold=0.0):
        """Scorer network (curc_amd.scorer.Mlp, kept alive by the caller) for
        runs with learned=True; None disables.  threshold is one logit
        threshold, or (for reads with a mapping, for reads without one)."""
        self._scorer = mlp
        t_map, t_un = (threshold, threshold) if np.isscalar(threshold) else threshold
        self._check(lib().curc_ctx_set_scorer(self._h, mlp._h if mlp is not None else None,
                                              ctypes.c_float(t_map), ctypes.c_float(t_un)), "curc_ctx_set_scorer")

    def get_tuning(self):
        """Launch shapes in use (dict of curc_tuning fields)."""
        t = Tuning()
        self._check(lib().curc_ctx_get_tuning(self._h, ctypes.byref(t)), "curc_ctx_get_tuning")
        return {n: getattr(t, n) for n, _ in Tuning._fields_}

    def set_tuning(self, **kw):
        """Sets launch shapes (curc_tuning fields; the others keep their
        values).  They change speed only, never results."""
        cur = self.get_tuning()
        cur.update(kw)
        t = Tuning(**cur)
        self._check(lib().curc_ctx_set_tuning(self._h, ctypes.byref(t)), "curc_ctx_set_tuning")

    def stream_copy_gbs(self, nbytes=1 << 31, reps=5):
        """Measured device copy bandwidth (read + write bytes per second, GB/s)
        of a plain 16-byte-word copy kernel over nbytes."""
        src = self.dev_alloc(nbytes)
        dst = self.dev_alloc(nbytes)
        try:
            ms = ctypes.c_float()
            self._check(lib().curc_stream_copy(self._h, dst, src, nbytes, reps, ctypes.byref(ms)), "curc_stream_copy")
            return 2.0 * nbytes / (ms.value * 1e-3) / 1e9
        finally:
            self.dev_free(src)
            self.dev_free(dst)

    def set_trace(self, every=16, max_records=1 << 22):
        self._check(lib().curc_ctx_set_trace(self._h, every, max_records), "curc_ctx_set_trace")

    def dev_alloc(self, nbytes):
        p = ctypes.c_void_p()
        self._check(lib().curc_dev_alloc(self._h, ctypes.byref(p), nbytes), "curc_dev_alloc")
        return p.value

    def dev_free(self, p):
        self._check(lib().curc_dev_free(self._h, ctypes.c_void_p(p)), "curc_dev_free")

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        self._check(lib().curc_memcpy_h2d(self._h, ctypes.c_void_p(dst), arr.ctypes.data, arr.nbytes), "h2d")

    def d2h(self, arr, src):
        self._check(lib().curc_memcpy_d2h(self._h, arr.ctypes.data, ctypes.c_void_p(src), arr.nbytes), "d2h")

    def pack(self, d_ascii, n, L, d_packed, d_isN=None):
        self._check(lib().curc_pack(self._h, ctypes.c_void_p(d_ascii), n, L, ctypes.c_void_p(d_packed),
                                    ctypes.c_void_p(d_isN) if d_isN else None), "curc_pack")

    def sync(self):
        self._check(lib().curc_sync(self._h), "curc_sync")

    def device_sync(self):
        self._check(lib().curc_device_sync(self._h), "curc_device_sync")

    def set_timing(self, on=True):
        lib().curc_ctx_set_timing(self._h, 1 if on else 0)

    def reset_stats(self):
        lib().curc_ctx_reset_stats(self._h)

    def stats(self):
        arr = (StageStat * 64)()
        n = lib().curc_ctx_stats(self._h, arr, 64)
        return [dict(name=arr[i].name.decode(), ms=arr[i].ms, launches=arr[i].launches,
                     alg_bytes=arr[i].alg_bytes, items=arr[i].items) for i in range(min(n, 64))]

    def trim(self):
        lib().curc_ctx_trim(self._h)

    def close(self):
        for d in list(getattr(self, "_dependents", [])):
            d.close()
        if self._h is not None and self._h.value:
            lib().curc_ctx_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rec_fields(rec):
    """Splits curc_alignment records into (read_id, isN, strand, pos) arrays."""
    w = rec["word"]
    return rec["read_id"], (w & 1).astype(np.uint8), ((w >> 1) & 1).astype(np.uint8), w >> np.uint64(2)
