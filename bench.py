"""Benchmark of the CURC pseudogenome path on one GPU.

Workload: 100 Mbp synthetic genome (GC 41%), 20,000,000 single-end 150 bp
reads (3.0 Gbases), 50% reverse complemented, substitutions rising linearly
from 0.1% at base 1 to 1.5% at base 150, N at 0.02% per base, one block,
written as FASTQ records (6.4 GB) into pinned host memory.

A step is one block through the whole path: host threads parse the FASTQ
records (the library's curc_fastq_host_parse, tuning fq_host) and only the
sequence lines are copied to HBM, then N split, 2-bit pack, dedup, prefix
k-mer index, greedy overlap, prune/partition, exact all-offset matching,
pseudogenome assembly, mapping onto it and the LQ pseudogenome, ending with
the pseudogenomes, map records and unmapped lists in host memory.  Steps run
as the reference runs the blocks of an input: the next block is parsed and
copied while the current block computes (curc_fastq_stream_run).  value is
that whole-job throughput; value_device_parse streams the same blocks with
the FASTQ bytes copied as they are and parsed on the device;
value_hbm_resident times the same steps with the FASTQ bytes already in HBM.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import re
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import curc_amd as ca  # noqa: E402
from curc_amd import fastq  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--reads", type=int, default=20_000_000)
    p.add_argument("--genome", type=int, default=100_000_000)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-cpu-full", action="store_true",
                   help="skip the OpenMP oracle on the whole block (keep the single-thread sample)")
    p.add_argument("--cpu-sample-reads", type=int, default=400_000,
                   help="reads of the single-thread oracle sample (30x coverage, same error profile)")
    p.add_argument("--no-full-block", action="store_true")
    p.add_argument("--no-bits-per-base", action="store_true")
    p.add_argument("--no-batch", action="store_true")
    p.add_argument("--no-repeat-pe", action="store_true",
                   help="skip the 100M-read repeat-heavy paired-end configuration")
    p.add_argument("--no-learned", action="store_true")
    p.add_argument("--no-tune", action="store_true",
                   help="keep the default launch shapes instead of autotuning them on the block first")
    p.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                   help="only run three device-to-device copies of BYTES (a kernel of known traffic for "
                        "scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    return p.parse_args()


def copy_probe(nbytes):
    import torch

    x = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"copied {nbytes} bytes three times")


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "mapping_traffic_100Mbp_20M_reads.md")


def measured_traffic():
    """HBM bytes per mapping launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (None if absent)."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            m = re.search(r"^traffic_bytes_per_launch = (\d+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def dominant_roofline(stats, steps, default_workload=True):
    """Roofline entry for the mapping stage, whose event pair brackets its
    launches (reads without N, then the N-reads) and nothing else.  traffic
    (rocprof fetched + written bytes of those launches) is given for the
    default workload only, which the profile covers."""
    single = [s for s in stats if s["name"] in ("mapping",)]
    if not single:
        return None
    s = max(single, key=lambda x: x["ms"])
    per_launch_ms = s["ms"] / max(1, s["launches"])
    per_launch_bytes = s["alg_bytes"] / max(1, s["launches"])
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms > 0 else 0.0
    return dict(kernel=s["name"], bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 5), traffic=measured_traffic() if default_workload else None,
                avg_launch_ms=round(per_launch_ms, 4), alg_bytes_per_launch=per_launch_bytes)


def stage_rooflines(stats):
    """Achieved algorithmic GB/s of every stage that counts its bytes:
    bytes per region / event time per region, against the 8 TB/s peak."""
    out = {}
    for s in stats:
        if s["ms"] <= 0 or s["alg_bytes"] <= 0 or s["launches"] == 0:
            continue
        ms = s["ms"] / s["launches"]
        b = s["alg_bytes"] / s["launches"]
        gbs = b / (ms * 1e-3) / 1e9
        out[s["name"]] = dict(ms=round(ms, 3), alg_bytes=round(b), achieved_GBs=round(gbs, 1),
                              frac=round(gbs / HBM_PEAK_GBS, 5))
    return out


def bits_per_base(learned=False):
    """Archive size per base on the 1 Mbp / 133,333-read workload: GPU
    streams and pseudogenome self-match -> xz-coded .curc archive (written
    to a temporary file); with learned, mapping under the trained scorer."""
    import tempfile

    from curc_amd import archive, scorer, train_scorer

    cfg = ca.config_small()
    _, reads = ca.synth_reads(cfg)
    with ca.Context(0) as ctx:
        net = None
        if learned:
            params, thr = train_scorer.load()
            net = scorer.Mlp(ctx, params)
            ctx.set_scorer(net, thr)
        blk = ctx.run_block(reads, streams=True, refmatch=True, learned=learned)
        bs = archive.BlockStreams.from_block(blk, len(reads), reads.shape[1], False, False)
        blk.free()
        if net is not None:
            ctx.set_scorer(None)
            net.close()
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "c1.curc")
        archive.write_archive(p, [bs], False, False, reads.shape[1])
        size = os.path.getsize(p)
        preserve, paired, L, blocks = archive.read_archive(p)
        dec = archive.decode_block(blocks[0], L, preserve, paired)
    lossless = sorted(map(bytes, dec)) == sorted(map(bytes, reads))
    return dict(value=round(8.0 * size / reads.size, 5), archive_bytes=size, lossless=bool(lossless),
                workload="1 Mbp uniform genome, 133,333 x 150 bp SE, 0.5% subst, 0.01% N, order not preserved")


def learned_run(ctx, a, L, run, plain_ms, plain_stats, cands):
    """The same block (FASTQ bytes resident in HBM) with mapping under the
    trained scorer; run(**flags) runs one block."""
    from curc_amd import scorer, train_scorer

    params, thr = train_scorer.load()
    net = scorer.Mlp(ctx, params)
    ctx.set_scorer(net, thr)

    def outcome(b):
        mis = b.get("map_mis").astype(np.int64)
        return dict(mapped=int(len(mis)), mismatches=int(mis.sum()), lq_reads=int(len(b.get("lq_rec"))),
                    nlq_reads=int(len(b.get("nlq_rec"))), lq_pg_bases=int(len(b.get("lq_pg"))))

    b = run()
    plain_out = outcome(b)
    b.free()
    run(learned=True).free()  # warm-up
    ctx.reset_stats()
    ctx.set_timing(True)
    ctx.device_sync()
    t0 = time.perf_counter()
    for k in range(a.steps):
        b = run(learned=True)
        if k + 1 < a.steps:
            b.free()
    ctx.device_sync()
    ms = (time.perf_counter() - t0) * 1e3 / a.steps
    out = outcome(b)
    b.free()
    ctx.set_timing(False)
    raw = ctx.stats()
    st = {s["name"]: round(s["ms"] / a.steps, 3) for s in raw if s["ms"] > 0}
    work = {s["name"]: int(s["items"] / a.steps) for s in raw if s["name"].startswith("map_") and s["ms"] == 0}
    log("  learned stages: " + ", ".join(f"{k} {v:.1f}" for k, v in st.items()))
    log("  learned mapping work: " + ", ".join(f"{k[4:]} {v}" for k, v in work.items()))
    rate = mlp_rate(ctx, net)
    ctx.set_scorer(None)
    net.close()
    pm = sum(s["ms"] for s in plain_stats if s["name"] == "mapping") / a.steps
    per_cand = dict(mlp_samples_per_s=round(rate, 1), mlp_tflops=round(rate * MLP_FLOP / 1e12, 2),
                    candidates_per_block=int(cands), scoring_all_candidates_ms=round(cands / rate * 1e3, 2),
                    plain_mapping_ms=round(pm, 3))
    log(f"  network inference: {rate / 1e9:.2f} G samples/s ({rate * MLP_FLOP / 1e12:.1f} TFLOP/s); scoring every one "
        f"of the block's {int(cands)} candidates would take {cands / rate * 1e3:.1f} ms against {pm:.1f} ms for the "
        f"whole plain mapping")
    log(f"learned mapping: {ms:.1f} ms per block (plain {plain_ms:.1f}); mapping {st.get('mapping_learned', 0):.1f} ms "
        f"(plain {pm:.1f}); mapped {out['mapped']} (plain {plain_out['mapped']}), mismatches {out['mismatches']} "
        f"(plain {plain_out['mismatches']})")
    return dict(ms_per_step=round(ms, 3), gbases_s=round(a.reads * L / 1e9 / (ms / 1e3), 4),
                speedup_vs_plain=round(plain_ms / ms, 4), mapping_ms=st.get("mapping_learned"),
                mapping_ms_plain=round(pm, 3), threshold=thr, outcome=out, outcome_plain=plain_out,
                stages_ms=st, mapping_work=work, per_candidate_scoring=per_cand,
                input="FASTQ bytes resident in HBM (plain_ms is the same input)")


MLP_FLOP = 2 * (32 * 128 + 128 * 128 + 128)  # multiply-adds of one 32 -> 128 -> 128 -> 1 forward, x2


def mlp_rate(ctx, net, n=1 << 23, reps=3):
    """Forward throughput of the scorer network (samples/s) on n feature rows
    resident in HBM."""
    from curc_amd import scorer

    x = np.random.default_rng(0).random((n, 32), dtype=np.float32) * 4
    dx = scorer.DeviceArray(ctx, x)
    do = scorer.DeviceArray(ctx, np.zeros(n, np.float32))
    try:
        net.forward_dev(dx.ptr, n, do.ptr)
        ctx.device_sync()
        t = time.perf_counter()
        for _ in range(reps):
            net.forward_dev(dx.ptr, n, do.ptr)
        ctx.device_sync()
        return n * reps / (time.perf_counter() - t)
    finally:
        dx.free()
        do.free()


def batched_inputs(n_inputs=64, seed=404):
    """Many-input workload: 64 inputs, each from its own genome of 1-50 Mbp
    (uniform), read lengths from {75, 100, 150, 250}, read counts
    log-uniform in [10k, 5M], 0.3-1% substitutions, 0.02% N."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inputs):
        L = int(rng.choice([75, 100, 150, 250]))
        n = int(np.exp(rng.uniform(np.log(10_000), np.log(5_000_000))))
        sub = float(rng.uniform(0.003, 0.01))
        g = int(rng.uniform(1_000_000, 50_000_000))
        cfg = ca.synth_config(seed=seed * 1000 + i, genome_len=g, gc=0.45, n_reads=n, L=L, sub_first=sub,
                              sub_last=sub, n_rate=0.0002)
        out.append(ca.synth_reads(cfg)[1])
    return out


def batched(ins, reps=2, learned=True):
    """The many-input workload one input after another on one stream (the
    plain path), then as one batch on concurrent streams, then as one batch
    with mapping under the trained scorer; all from host memory."""
    from curc_amd import scorer, train_scorer

    gbases = sum(r.size for r in ins) / 1e9

    def mapped(blocks):
        n = 0
        for b in blocks:
            n += len(b.get("map_rec"))
            b.free()
        return n

    from curc_amd import autotune

    with ca.Context(0) as ctx:
        for b in ctx.batch_run(ins[:4], workers=4):
            b.free()
        workers, wtimes = autotune.tune_workers(ctx, ins[::2], reps=2, log=log)
        t = time.perf_counter()
        for _ in range(reps):
            for r in ins:
                ctx.run_block(r).free()
        ctx.device_sync()
        seq = (time.perf_counter() - t) / reps
        t = time.perf_counter()
        for k in range(reps):
            blocks = ctx.batch_run(ins, workers=0)
            if k + 1 < reps:
                for b in blocks:
                    b.free()
        ctx.device_sync()
        bat = (time.perf_counter() - t) / reps
        plain_mapped = mapped(blocks)
        lb = learned_mapped = None
        if learned:
            params, thr = train_scorer.load()
            net = scorer.Mlp(ctx, params)
            ctx.set_scorer(net, thr)
            for b in ctx.batch_run(ins[:4], workers=4, learned=True):
                b.free()
            t = time.perf_counter()
            for k in range(reps):
                blocks = ctx.batch_run(ins, workers=0, learned=True)
                if k + 1 < reps:
                    for b in blocks:
                        b.free()
            ctx.device_sync()
            lb = (time.perf_counter() - t) / reps
            learned_mapped = mapped(blocks)
            ctx.set_scorer(None)
            net.close()
    out_bits = batch_bits(ins, learned)
    out = dict(inputs=len(ins), gbases=round(gbases, 4), sequential_s=round(seq, 4), batched_s=round(bat, 4),
               sequential_gbases_s=round(gbases / seq, 4), batched_gbases_s=round(gbases / bat, 4),
               speedup=round(seq / bat, 3), mapped_reads=plain_mapped, workers=workers,
               workers_tuning_s=wtimes,
               workload="64 inputs: genomes 1-50 Mbp, L in {75,100,150,250}, 10k-5M reads log-uniform, "
                        "0.3-1% subst, input in host memory")
    if lb is not None:
        out.update(learned_batched_s=round(lb, 4), learned_batched_gbases_s=round(gbases / lb, 4),
                   learned_batched_vs_plain=round(seq / lb, 3), learned_batched_vs_batched=round(bat / lb, 3),
                   scorer_gain_alone=round(bat / lb, 3), learned_mapped_reads=learned_mapped)
    out["bits_per_base"] = out_bits
    return out


def batch_bits(ins, learned=True):
    """bits/base over the many-input workload (one .curc archive per input,
    summed), plain and learned, from batches with streams and self-match."""
    from curc_amd import archive, scorer, train_scorer

    t = time.perf_counter()
    bases = sum(r.size for r in ins)

    def total(ctx, lrn):
        blocks = ctx.batch_run(ins, workers=0, streams=True, refmatch=True, learned=lrn)
        arch = []
        for r, b in zip(ins, blocks):
            arch.append(([archive.BlockStreams.from_block(b, len(r), r.shape[1], False, False)], False, False,
                         r.shape[1]))
            b.free()
        return sum(archive.archive_sizes(arch))

    with ca.Context(0) as ctx:
        plain = total(ctx, False)
        out = dict(plain=round(8.0 * plain / bases, 5), plain_archive_bytes=plain)
        if learned:
            params, thr = train_scorer.load()
            net = scorer.Mlp(ctx, params)
            ctx.set_scorer(net, thr)
            lsize = total(ctx, True)
            ctx.set_scorer(None)
            net.close()
            out.update(learned=round(8.0 * lsize / bases, 5), learned_archive_bytes=lsize,
                       learned_vs_plain=round(lsize / plain, 5))
    out["seconds"] = round(time.perf_counter() - t, 1)
    log(f"  bits/base over the {len(ins)} inputs: plain {out['plain']}" +
        (f", learned {out['learned']} ({out['learned_vs_plain']}x)" if learned else "") + f" ({out['seconds']} s)")
    return out


def host_info():
    """Host cores and CPU model of this machine (nproc, lscpu)."""
    import subprocess

    info = {"os_cpu_count": os.cpu_count()}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        info["nproc"] = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        m = re.search(r"^Model name:\s*(.+)$", out, re.M)
        info["cpu_model"] = m.group(1).strip() if m else None
    except OSError:
        info["cpu_model"] = None
    return info


FIELDS_CHECKED = ("hq_pg", "hq_rec", "map_rec", "map_mis", "unmapped", "unmapped_N", "lq_pg", "lq_rec")


def cpu_baseline(fq, L, gpu_block, sample_reads, full=True):
    """The oracle over the same region as the GPU: from the FASTQ bytes in
    host memory (its own parser, preprocess.cpp:147-193) to the
    pseudogenomes, map records and unmapped lists.  The OpenMP build runs
    once on the whole block (its results are compared field by field with
    the GPU block of the timed run); the single-thread build and the OpenMP
    build are also timed on a smaller sample with the same coverage and
    error profile."""
    from oracle import oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = dict(unit="Gbases/s", cores=threads, kind="port", host=host_info())
    if full:
        t0 = time.perf_counter()
        reads = orc.parse_fastq(fq, L, kind="omp")
        t1 = time.perf_counter()
        ob = orc.OracleBlock(reads, stop="lq", kind="omp")
        dt = time.perf_counter() - t0
        gbases = len(reads) * L / 1e9
        same = {f: bool(np.array_equal(gpu_block.get(f), ob.get(f))) for f in FIELDS_CHECKED}
        ob.free()
        del reads
        out.update(value=round(gbases / dt, 6), seconds=round(dt, 2), parse_seconds=round(t1 - t0, 2),
                   sample=f"oracle_omp ({threads} threads), one run, on the whole 20M-read block from the same "
                          f"FASTQ bytes: parse + block, {dt:.1f} s",
                   gpu_matches_oracle=all(same.values()), fields_compared=same)
        log(f"  oracle_omp whole block: {dt:.1f} s ({gbases / dt:.4f} Gbases/s), GPU == oracle: {all(same.values())}")
    cfg = ca.config_scale(seed=23, n_reads=sample_reads, genome_len=sample_reads * L // 30)
    _, sr = ca.synth_reads(cfg)
    sfq = fastq.fastq_bytes(sr)
    del sr
    times = {}
    for kind in ("st", "omp"):
        t0 = time.perf_counter()
        r = orc.parse_fastq(sfq, L, kind=kind)
        ob = orc.OracleBlock(r, stop="lq", kind=kind)
        times[kind] = time.perf_counter() - t0
        ob.free()
    g = sample_reads * L / 1e9
    out["sample_runs"] = dict(
        workload=f"{sample_reads} reads x {L} bp from a {sample_reads * L // 30} bp genome (30x, same error profile)",
        oracle_st_s=round(times["st"], 2), oracle_st_gbases_s=round(g / times["st"], 6),
        oracle_omp_s=round(times["omp"], 2), oracle_omp_gbases_s=round(g / times["omp"], 6))
    if not full:
        out.update(value=round(g / times["omp"], 6), sample=out["sample_runs"]["workload"], gpu_matches_oracle=None)
    log(f"  sample: oracle_st {times['st']:.1f} s, oracle_omp {times['omp']:.1f} s")
    return out


def row_hashes(a):
    """A 64-bit hash per row of a uint8 matrix (for multiset comparisons)."""
    n, L = a.shape
    w = (L + 7) // 8
    pad = np.zeros((n, 8 * w), dtype=np.uint8)
    pad[:, :L] = a
    words = pad.view(np.uint64)
    h = np.full(n, 0x9E3779B97F4A7C15, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for k in range(w):
            h = (h ^ words[:, k]) * np.uint64(0xBF58476D1CE4E5B9)
            h ^= h >> np.uint64(31)
    return h


def decode_run(ctx, blk, n_reads, L, hb):
    """curc -d's read rebuild of the same block on the GPU, from its streams
    (order not preserved), twice; the second run is timed.  The rebuilt
    reads must be the input's reads as a multiset."""
    from curc_amd import archive

    bs = archive.BlockStreams.from_block(blk, n_reads, L, False, False)
    ctx.decode(bs, L, False, False)
    ctx.reset_stats()
    ctx.set_timing(True)
    ctx.device_sync()
    t = time.perf_counter()
    out = ctx.decode(bs, L, False, False)
    ms = (time.perf_counter() - t) * 1e3
    ctx.set_timing(False)
    st = {s["name"]: s["ms"] for s in ctx.stats() if s["ms"] > 0}
    kern = st.get("decode", 0.0)
    inp = fastq.sequences(hb.array, L)
    lossless = bool(np.array_equal(np.sort(row_hashes(out)), np.sort(row_hashes(inp))))
    gb = n_reads * L / 1e9
    res = dict(kernel_ms=round(kern, 3), gbases_s=round(gb / (kern / 1e3), 3) if kern else None,
               host_to_host_ms=round(ms, 3), host_to_host_gbases_s=round(gb / (ms / 1e3), 3),
               h2d_ms=round(st.get("decode_h2d", 0.0), 3), d2h_ms=round(st.get("decode_d2h", 0.0), 3),
               reads_match_input=lossless)
    log(f"GPU read rebuild (curc -d): kernels {kern:.1f} ms ({res['gbases_s']} Gbases/s), host to host "
        f"{ms:.1f} ms; reads == input (multiset): {lossless}")
    return res


def repeat_pe_run(reps=2):
    """The repeat-heavy paired-end configuration at its stated size (500 Mbp,
    45% repeats, 100M x 100 bp pairs, order preserved, block_ratio 0.25 at
    the 50 MB chunk granularity -> 4 uneven blocks), as one batch from FASTQ
    bytes in pinned host memory (curc_fastq_batch_run, 4 streams); then each
    block alone, and the 4 blocks streamed in order on one stream."""
    from curc_amd import workloads

    t0 = time.perf_counter()
    _, reads = ca.synth_reads(workloads.config_repeat_pe())
    L = reads.shape[1]
    n = len(reads)
    with ca.Context(0) as ctx:
        f1, f2 = workloads.pinned_fastq(ctx, reads, True)
        del reads
        ranges = fastq.block_ranges(f1.array, f2.array, 0.25)
        jobs = [(f1.array[s:e], f2.array[s:e]) for s, e in ranges]
        log(f"repeat-heavy PE input: {n} reads, 2 x {f1.nbytes / 1e9:.2f} GB FASTQ, {len(ranges)} blocks, "
            f"built in {time.perf_counter() - t0:.1f} s")
        kw = dict(streams=True, preserve_order=True, paired=True)
        for b in ctx.fastq_batch(jobs, L, workers=4, **kw):  # warm-up (scratch, pinned result chunks)
            b.free()
        times = []
        counts = None
        for _ in range(reps):
            ctx.device_sync()
            t = time.perf_counter()
            blocks = ctx.fastq_batch(jobs, L, workers=4, **kw)
            ctx.device_sync()
            times.append(time.perf_counter() - t)
            counts = [b.n_reads for b in blocks]
            for b in blocks:
                b.free()
        single = []
        for j in jobs:
            ctx.device_sync()
            t = time.perf_counter()
            ctx.fastq_block(j[0], j[1], L=L, **kw).free()
            single.append(round((time.perf_counter() - t) * 1e3, 1))
        ctx.device_sync()
        t = time.perf_counter()
        ctx.fastq_stream(jobs, L, lambda i, b: b.free(), **kw)
        stream_s = time.perf_counter() - t
        f1.free()
        f2.free()
    gb = n * L / 1e9
    bs = min(times)
    out = dict(workload="500 Mbp genome, 45% repeats (300 bp-6 kbp, 5-15% divergence), 2 x 100 bp pairs, "
                        "fragment ~N(400,50), 20x, 0.5% subst, preserve_order, block_ratio 0.25, "
                        "from pinned FASTQ bytes, streams built",
               reads=n, gbases=round(gb, 3), blocks=len(ranges), block_reads=counts,
               batch_s=round(bs, 4), gbases_s=round(gb / bs, 3), batch_runs_s=[round(x, 4) for x in times],
               block_alone_ms=single, stream_in_order_s=round(stream_s, 4),
               stream_in_order_gbases_s=round(gb / stream_s, 3))
    log(f"  batch of {len(ranges)} blocks: {bs:.3f} s ({gb / bs:.2f} Gbases/s); blocks alone {single} ms; "
        f"in order on one stream {stream_s:.3f} s")
    return out


def block_bits(ctx, plain_blk, run, n_reads, L, learned=True):
    """bits/base of the block's .curc archive (GPU streams + self-match,
    xz-coded sub-files), plain and with mapping under the trained scorer, with
    the learned archive's reads rebuilt on the GPU and checked against the
    plain one's as a multiset."""
    from curc_amd import archive, scorer, train_scorer

    t = time.perf_counter()
    bs = archive.BlockStreams.from_block(plain_blk, n_reads, L, False, False)
    size = archive.archive_bytes([bs], False, False, L)
    out = dict(plain=round(8.0 * size / (n_reads * L), 5), plain_archive_bytes=size)
    if learned:
        params, thr = train_scorer.load()
        net = scorer.Mlp(ctx, params)
        ctx.set_scorer(net, thr)
        lb = run(streams=True, refmatch=True, learned=True)
        ctx.set_scorer(None)
        net.close()
        lbs = archive.BlockStreams.from_block(lb, n_reads, L, False, False)
        lsize = archive.archive_bytes([lbs], False, False, L)
        a = ctx.decode(lbs, L, False, False)
        b = ctx.decode(bs, L, False, False)
        same = bool(np.array_equal(np.sort(row_hashes(a)), np.sort(row_hashes(b))))
        lb.free()
        out.update(learned=round(8.0 * lsize / (n_reads * L), 5), learned_archive_bytes=lsize,
                   learned_vs_plain=round(lsize / size, 5), learned_lossless=same)
    out["seconds"] = round(time.perf_counter() - t, 1)
    log(f"  bits/base of this block: plain {out['plain']}" +
        (f", learned {out['learned']} ({out['learned_vs_plain']}x, lossless {out['learned_lossless']})"
         if learned else "") + f" ({out['seconds']} s)")
    return out


def make_fastq(ctx, cfg):
    """The workload's reads as FASTQ records in pinned host memory."""
    _, reads = ca.synth_reads(cfg)
    rec = len(fastq.fastq_bytes(reads[:1]))
    hb = ctx.host_buffer(len(reads) * rec)
    step = 1 << 20
    for i in range(0, len(reads), step):
        part = reads[i:i + step]
        hb.array[i * rec:(i + len(part)) * rec] = fastq.fastq_bytes(part, first_id=i)
    return hb


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              f"(blocks shard across GPUs as independent jobs; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    L = 150
    cfg = ca.config_scale(n_reads=a.reads, genome_len=a.genome)
    ctx = ca.Context(0)
    t0 = time.perf_counter()
    hb = make_fastq(ctx, cfg)
    nbytes = hb.nbytes
    log(f"generated {a.reads} reads as {nbytes / 1e9:.2f} GB of FASTQ in pinned memory in "
        f"{time.perf_counter() - t0:.1f} s")
    # the same bytes resident in HBM, for the tuner, the HBM-resident figure
    # and the runs beside the metric
    d_fq = ctx.dev_alloc(nbytes)
    ctx.h2d(d_fq, hb.array)
    ctx.device_sync()

    def run_dev(**kw):
        return ctx.fastq_block(d_fq, L=L, on_device=True, nbytes=(nbytes, 0), **kw)

    tuning = None
    if not a.no_tune:
        # launch shapes only (results do not depend on them), chosen on this
        # block before the timed region
        from curc_amd import autotune

        t0 = time.perf_counter()
        best, table = autotune.tune_block(ctx, lambda: run_dev().free(), reps=2, log=log)
        tuning = dict(best, seconds=round(time.perf_counter() - t0, 1), candidates_ms=table)

    info = {}

    def on_block(i, b):
        if "hq" not in info:
            info["hq"] = len(b.get("hq_rec"))
            info["mapped"] = len(b.get("map_rec"))
            info["lq"] = len(b.get("lq_rec"))
            info["nlq"] = len(b.get("nlq_rec"))
        b.free()

    t = time.perf_counter()
    ctx.fastq_stream([hb] * a.warmup, L, on_block)
    log(f"warmup: {a.warmup} blocks in {time.perf_counter() - t:.3f} s; HQ records {info.get('hq')}, mapped "
        f"{info.get('mapped')}, LQ records {info.get('lq')}, N-LQ records {info.get('nlq')}")
    last = {}

    def on_timed(i, b):
        if i == a.steps - 1:
            last["block"] = b  # kept for the comparison with the oracle
        else:
            b.free()

    ctx.reset_stats()
    ctx.set_timing(True)
    ctx.device_sync()
    t0 = time.perf_counter()
    ctx.fastq_stream([hb] * a.steps, L, on_timed)
    ctx.device_sync()
    dt = time.perf_counter() - t0
    ctx.set_timing(False)
    stats = ctx.stats()
    work = {s["name"]: int(s["items"] / a.steps) for s in stats if s["name"].startswith("map_") and s["ms"] == 0}
    graphs = {s["name"][6:]: int(s["items"]) for s in stats if s["name"].startswith("graph_")}
    stats = [s for s in stats if s["ms"] > 0 or not s["name"].startswith(("map_", "graph_"))]
    ms = dt * 1e3 / a.steps
    gbases = a.reads * L / 1e9
    value = gbases / (ms / 1e3)
    log(f"streamed from pinned FASTQ: {ms:.1f} ms per block, {value:.2f} Gbases/s")
    log("stage breakdown (ms per step):")
    for s in sorted(stats, key=lambda s: -s["ms"]):
        log(f"  {s['name']:<14} {s['ms'] / a.steps:9.2f} ms  ({s['launches'] // max(1, a.steps)} regions/step)")
    log("mapping work per step (both strand passes): " + ", ".join(f"{k[4:]} {v}" for k, v in work.items()))
    default = (a.reads, a.genome) == (20_000_000, 100_000_000)
    roof = dominant_roofline(stats, a.steps, default_workload=default)
    if roof:
        # the measured ceiling beside the vendor peak: a plain copy of 2 GiB
        copy = ctx.stream_copy_gbs()
        roof["measured_copy_GBs"] = round(copy, 1)
        roof["frac_of_measured_copy"] = round(roof["achieved"] / copy, 5)
        if roof.get("traffic"):
            roof["traffic_GBs"] = round(roof["traffic"] / (roof["avg_launch_ms"] * 1e-3) / 1e9, 1)
        log(f"stream copy: {copy:.0f} GB/s measured (vendor peak {HBM_PEAK_GBS:.0f} GB/s)")
    stage_roof = stage_rooflines(stats)

    # the same blocks streamed with the whole FASTQ bytes copied and parsed
    # on the device
    k2 = max(1, min(a.steps, 10))
    host_threads = ctx.get_tuning()["fq_host"]
    ctx.set_tuning(fq_host=0)
    ctx.fastq_stream([hb] * 2, L, lambda i, b: b.free())
    ctx.device_sync()
    t0 = time.perf_counter()
    ctx.fastq_stream([hb] * k2, L, lambda i, b: b.free())
    ctx.device_sync()
    ms_dev_parse = (time.perf_counter() - t0) * 1e3 / k2
    ctx.set_tuning(fq_host=host_threads)
    log(f"streamed, FASTQ bytes copied whole and parsed on the device: {ms_dev_parse:.1f} ms per block "
        f"({gbases / (ms_dev_parse / 1e3):.2f} Gbases/s); host parse threads in the timed run: {host_threads}")

    # the same steps with the FASTQ bytes resident in HBM (no host copy)
    ctx.device_sync()
    t0 = time.perf_counter()
    for _ in range(k2):
        run_dev().free()
    ctx.device_sync()
    ms_hbm = (time.perf_counter() - t0) * 1e3 / k2
    # one block from pinned host FASTQ, nothing overlapped (copy, then compute)
    lat = []
    for _ in range(2):
        ctx.device_sync()
        t0 = time.perf_counter()
        ctx.fastq_block(hb, L=L).free()
        lat.append((time.perf_counter() - t0) * 1e3)
    log(f"FASTQ resident in HBM: {ms_hbm:.1f} ms per block ({gbases / (ms_hbm / 1e3):.2f} Gbases/s); "
        f"one block from pinned host memory, not overlapped: {min(lat):.1f} ms")

    learned = None
    if not a.no_learned:
        learned = learned_run(ctx, a, L, run_dev, ms_hbm, stats, work.get("map_cands", 0))
    full = None
    if not a.no_full_block:
        # the same block once more with the stream builders and the
        # pseudogenome self-match (outside the metric's region)
        ctx.reset_stats()
        ctx.set_timing(True)
        ctx.device_sync()
        t = time.perf_counter()
        b = run_dev(streams=True, refmatch=True)
        ctx.device_sync()
        full_ms = (time.perf_counter() - t) * 1e3
        ctx.set_timing(False)
        fst = [s for s in ctx.stats() if s["ms"] > 0]
        full = {"ms": round(full_ms, 3), "stages_ms": {s["name"]: round(s["ms"], 3) for s in fst},
                "pg_bases": {p: int(len(b.get(p + "_pg"))) for p in ("hq", "lq", "nlq")},
                "pg_bases_after_refmatch": {p: int(len(b.get(f"rm_{p}_text"))) for p in ("hq", "lq", "nlq")}}
        log(f"block with streams + self-match: {full_ms:.1f} ms; " +
            ", ".join(f"{k} {v:.1f}" for k, v in full["stages_ms"].items() if k in ("streams", "refmatch")))
        full["decode"] = decode_run(ctx, b, a.reads, L, hb)
        if not a.no_bits_per_base:
            full["bits_per_base"] = block_bits(ctx, b, run_dev, a.reads, L, learned=not a.no_learned)
        b.free()
    ctx.dev_free(d_fq)
    cpu = None
    if not a.no_cpu_baseline:
        log("timing the CPU oracle on the same FASTQ bytes ...")
        cpu = cpu_baseline(hb.array, L, last["block"], a.cpu_sample_reads, full=not a.no_cpu_full)
        if cpu.get("value"):
            cpu["gpu_over_cpu"] = round(value / cpu["value"], 1)
    last["block"].free()
    hb.free()
    ctx.close()
    bpb = None
    if not a.no_bits_per_base:
        bpb = bits_per_base()
        log(f"bits/base on the 1 Mbp workload: {bpb['value']} (lossless round trip: {bpb['lossless']})")
        if not a.no_learned:
            bl = bits_per_base(learned=True)
            bpb["learned"] = bl["value"]
            bpb["learned_lossless"] = bl["lossless"]
            bpb["learned_vs_plain"] = round(bl["value"] / bpb["value"], 5)
            log(f"  with the learned scorer: {bl['value']} (lossless round trip: {bl['lossless']})")
    bat = None
    if not a.no_batch:
        log("many-input batch ...")
        bat = batched(batched_inputs(), learned=not a.no_learned)
        log(f"  {bat['inputs']} inputs, {bat['gbases']} Gbases: one by one {bat['sequential_s']} s, "
            f"batched {bat['batched_s']} s ({bat['speedup']}x)" +
            (f", learned+batched {bat['learned_batched_s']} s ({bat['learned_batched_vs_plain']}x)"
             if "learned_batched_s" in bat else ""))
    c3 = None
    if not a.no_repeat_pe:
        log("repeat-heavy paired-end configuration ...")
        c3 = repeat_pe_run()
    out = {
        "metric": "Gbases/s (pseudogenome build+mapping, 1 GPU)",
        "value": round(value, 4),
        "unit": "Gbases/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"plain path, {a.genome // 1_000_000} Mbp genome GC 41%, {a.reads} x {L} bp SE, "
                               "subst 0.1%->1.5%, N 0.02%, one block, from FASTQ bytes in pinned host memory",
                   "reads": a.reads, "read_len": L, "genome_bp": a.genome, "fastq_bytes": nbytes},
        "timed_region": "FASTQ bytes in pinned host memory -> HQ/LQ/N-LQ pseudogenomes, records, mismatch "
                        "counts and unmapped lists in host memory; record parse (host threads), host-to-device "
                        "copy of the sequence lines, N split and pack inside; steps streamed (the next block's "
                        "parse and copy beside the current block's compute)",
        "fastq_parse": {"host_threads": host_threads, "value_device_parse": round(gbases / (ms_dev_parse / 1e3), 4),
                        "ms_per_step_device_parse": round(ms_dev_parse, 3)},
        "value_hbm_resident": round(gbases / (ms_hbm / 1e3), 4),
        "ms_per_step_hbm_resident": round(ms_hbm, 3),
        "block_latency_ms_from_pinned_host": round(min(lat), 3),
        "stages_ms": {s["name"]: round(s["ms"] / a.steps, 3) for s in stats},
        "mapping_work": work,
        "offset_loops": graphs,
        "tuning": tuning,
        "roofline": roof,
        "stage_roofline": stage_roof,
        "cpu_baseline": cpu,
        "bits_per_base": bpb,
        "decode_gbases_s": (full or {}).get("decode", {}).get("gbases_s"),
        "full_block": full,
        "batched": bat,
        "c3": c3,
        "learned": learned,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
