"""FASTQ input (curc_fastq_*): the record parse of preprocess.cpp:147-199 on
the GPU (tuning fq_host 0: the bytes copied as they are) or on host threads
before the copy (fq_host 8: only the sequences copied), then the N split and
pack on the GPU.  Every test runs in both modes.  Blocks from FASTQ bytes
must equal the ASCII-read runs of the same records and the oracle (whose
parser restates the reference loop); malformed input fails with
CURC_EINVAL."""
import ctypes

import numpy as np
import pytest

import curc_amd as ca
from curc_amd import fastq
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("nonN_ids", "N_ids", "hq_pg", "hq_rec", "map_rec", "map_mis", "unmapped", "unmapped_N", "lq_pg", "lq_rec",
          "nlq_pg", "nlq_rec")


@pytest.fixture(params=[0, 8], ids=["device_parse", "host_parse"], autouse=True)
def parse_mode(request, gpu_ctx):
    before = gpu_ctx.get_tuning()
    gpu_ctx.set_tuning(fq_host=request.param)
    print(f"fq_host {request.param}", flush=True)
    yield request.param
    gpu_ctx.set_tuning(**before)


def records(reads, mate=None, vary=True):
    """FASTQ bytes with header lengths that vary by record (sequence lines at
    every byte alignment)."""
    out = bytearray()
    L = reads.shape[1]
    tail = b"" if mate is None else b"/%d" % mate
    for k in range(len(reads)):
        pad = b"x" * (k % 7) if vary else b""
        out += b"@synth.%d%s%s\n" % (k, pad, tail) + reads[k].tobytes() + b"\n+\n" + b"I" * L + b"\n"
    return bytes(out)


def workload(seed, n, L=150, genome=None, n_rate=0.002):
    cfg = ca.synth_config(seed=seed, genome_len=genome or max(20_000, n * L // 25), gc=0.45, n_reads=n, L=L,
                          sub_first=0.002, sub_last=0.01, n_rate=n_rate)
    return ca.synth_reads(cfg)[1]


def same(a, b, fields, what):
    for f in fields:
        x, y = a.get(f), b.get(f)
        assert x.shape == y.shape and np.array_equal(x, y), f"{what}: {f} differs"


@pytest.mark.parametrize("L", [100, 150])
def test_fastq_block_matches_ascii_run_and_oracle(gpu_ctx, L):
    reads = workload(31 + L, 20_000, L)
    fq = records(reads)
    print(f"FASTQ block: {len(reads)} x {L}, {len(fq)} bytes", flush=True)
    b = gpu_ctx.fastq_block(fq, L=L, keep_stages=True)
    assert b.n_reads == len(reads)
    a = gpu_ctx.run_block(reads, keep_stages=True)
    same(b, a, FIELDS + ("packed", "N_ascii", "next", "prev", "offset"), "FASTQ vs ASCII run")
    parsed = orc.parse_fastq(fq, L)
    assert np.array_equal(parsed, reads)
    o = orc.OracleBlock(parsed)
    same(b, o, FIELDS, "FASTQ vs oracle")
    a.free()
    b.free()


def test_fastq_paired_block_interleaves(gpu_ctx):
    L = 100
    reads = workload(41, 16_000, L)
    f1, f2 = records(reads[0::2], 1), records(reads[1::2], 2)
    print(f"paired FASTQ block: {len(reads)} reads", flush=True)
    b = gpu_ctx.fastq_block(f1, f2, L=L, streams=True, preserve_order=True, keep_stages=True)
    a = gpu_ctx.run_block(reads, streams=True, preserve_order=True, paired=True, keep_stages=True)
    same(b, a, FIELDS + ("s_strand", "s_mis_count", "s_mis_base", "s_id_pos"), "paired FASTQ vs ASCII")
    o = orc.OracleBlock(orc.parse_fastq(f1, L, f2))
    same(b, o, FIELDS, "paired FASTQ vs oracle")
    a.free()
    b.free()


def test_fastq_on_device_input(gpu_ctx):
    L = 150
    reads = workload(51, 8_000, L)
    fq = np.frombuffer(records(reads), dtype=np.uint8)
    d = gpu_ctx.dev_alloc(len(fq))
    try:
        gpu_ctx.h2d(d, fq)
        b = gpu_ctx.fastq_block(d, L=L, on_device=True, nbytes=(len(fq), 0))
        a = gpu_ctx.run_block(reads)
        same(b, a, FIELDS, "device FASTQ vs ASCII")
        a.free()
        b.free()
    finally:
        gpu_ctx.dev_free(d)


def test_fastq_errors_then_context_still_works(gpu_ctx):
    L = 100
    reads = workload(61, 3_000, L)
    fq = records(reads, vary=False)
    ragged = fq.replace(reads[10].tobytes(), reads[10].tobytes()[:-1], 1)
    longer = fq.replace(reads[20].tobytes(), reads[20].tobytes() + b"A", 1)
    for bad, what in ((ragged, "read length"), (longer, "read length"), (fq[:-7], "truncated"),
                      (b"#" + fq[1:], "start with")):
        print("malformed case:", what, len(bad), flush=True)
        with pytest.raises(ca.CurcError, match=what) as e:
            gpu_ctx.fastq_block(bad, L=L)
        assert "(-1)" in str(e.value)
    half = records(reads[:1500], 1, vary=False)
    with pytest.raises(ca.CurcError, match="record counts"):
        gpu_ctx.fastq_block(half, records(reads[:1400], 2, vary=False), L=L)
    b = gpu_ctx.fastq_block(fq, L=L)
    a = gpu_ctx.run_block(reads)
    same(b, a, FIELDS, "after errors")
    a.free()
    b.free()
    e = gpu_ctx.fastq_block(b"", L=L)
    assert e.n_reads == 0 and len(e.get("hq_pg")) == 0
    e.free()


def test_fastq_pack_entry(gpu_ctx):
    L = 75
    reads = workload(71, 5_000, L, n_rate=0.01)
    fq = np.frombuffer(records(reads), dtype=np.uint8)
    n = len(reads)
    W = (L + 31) // 32
    d_fq = gpu_ctx.dev_alloc(len(fq))
    d_p = gpu_ctx.dev_alloc(n * W * 8)
    d_n = gpu_ctx.dev_alloc(n + 4)
    d_s = gpu_ctx.dev_alloc(n * 8)
    try:
        gpu_ctx.h2d(d_fq, fq)
        got_n = gpu_ctx.fastq_pack(d_fq, len(fq), L, n, d_p, d_n, d_s)
        assert got_n == n
        packed = np.empty(n * W, dtype=np.uint64)
        isn = np.empty(n, dtype=np.uint8)
        starts = np.empty(n, dtype=np.uint64)
        gpu_ctx.d2h(packed, d_p)
        gpu_ctx.d2h(isn, d_n)
        gpu_ctx.d2h(starts, d_s)
        assert np.array_equal(packed.reshape(n, W), orc.pack(reads))
        assert np.array_equal(isn != 0, (reads == ord("N")).any(axis=1))
        nl = np.flatnonzero(fq == 10)
        assert np.array_equal(starts, (nl[0::4] + 1).astype(np.uint64))
        with pytest.raises(ca.CurcError, match=r"\(-4\)"):
            gpu_ctx.fastq_pack(d_fq, len(fq), L, n - 1, d_p, d_n, d_s)
        with pytest.raises(ca.CurcError, match=r"\(-1\)"):
            gpu_ctx.fastq_pack(d_fq + 1, len(fq) - 1, L, n, d_p, d_n, d_s)
    finally:
        for p in (d_fq, d_p, d_n, d_s):
            gpu_ctx.dev_free(p)


def test_fastq_stream_and_batch_match_single_runs(gpu_ctx):
    L = 150
    sizes = [9_000, 4_000, 12_000, 7_000]
    ins = [workload(81 + i, n, L) for i, n in enumerate(sizes)]
    fqs = [records(r) for r in ins]
    pinned = []
    for f in fqs:
        hb = gpu_ctx.host_buffer(len(f))
        hb.array[:] = np.frombuffer(f, dtype=np.uint8)
        pinned.append(hb)
    singles = [gpu_ctx.run_block(r) for r in ins]
    seen = []

    def on_block(i, blk):
        seen.append(i)
        same(blk, singles[i], FIELDS, f"stream block {i}")
        blk.free()

    print("stream of", sizes, flush=True)
    gpu_ctx.fastq_stream(pinned, L, on_block)
    assert seen == [0, 1, 2, 3]
    # pageable host bytes take the same path
    seen.clear()
    gpu_ctx.fastq_stream(fqs[:2], L, on_block)
    assert seen == [0, 1]
    print("FASTQ batch", flush=True)
    got = gpu_ctx.fastq_batch(pinned, L, workers=3)
    for i, b in enumerate(got):
        same(b, singles[i], FIELDS, f"batch block {i}")
        b.free()
    for s in singles:
        s.free()
    for hb in pinned:
        hb.free()


def test_fastq_stream_stops_at_first_error(gpu_ctx):
    L = 100
    good = records(workload(91, 2_000, L))
    bad = good.replace(b"\n+\n", b"\n+\n", 1)[: len(good) - 3]
    seen = []
    with pytest.raises(ca.CurcError, match="truncated"):
        gpu_ctx.fastq_stream([good, bad, good], L, lambda i, b: (seen.append(i), b.free()))
    assert seen == [0]
