"""Benchmark of the CURC pseudogenome path on one GPU.

Workload: 100 Mbp synthetic genome (GC 41%), 20,000,000 single-end 150 bp
reads (3.0 Gbases), 50% reverse complemented, substitutions rising linearly
from 0.1% at base 1 to 1.5% at base 150, N at 0.02% per base, one block.

A step is one block through the whole GPU path: N split + 2-bit pack, dedup,
prefix k-mer index, greedy overlap, prune/partition, exact all-offset
matching, pseudogenome assembly, mapping onto it and the LQ pseudogenome,
starting from the ASCII reads resident in HBM and ending with the
pseudogenomes, map records and unmapped lists in host memory.

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
    p.add_argument("--cpu-reads", type=int, default=2_000_000)
    p.add_argument("--no-full-block", action="store_true")
    p.add_argument("--no-bits-per-base", action="store_true")
    p.add_argument("--no-batch", action="store_true")
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
    """Roofline entry for the stage with the largest event time whose region
    brackets single kernel launches.  traffic (rocprof HBM bytes per launch)
    is given for the default workload only, which the profile covers."""
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


def learned_run(ctx, a, L, d_ascii, plain_ms, plain_stats):
    """The same block with mapping under the trained scorer."""
    from curc_amd import scorer, train_scorer

    params, thr = train_scorer.load()
    net = scorer.Mlp(ctx, params)
    ctx.set_scorer(net, thr)

    def outcome(b):
        mis = b.get("map_mis").astype(np.int64)
        return dict(mapped=int(len(mis)), mismatches=int(mis.sum()), lq_reads=int(len(b.get("lq_rec"))),
                    nlq_reads=int(len(b.get("nlq_rec"))), lq_pg_bases=int(len(b.get("lq_pg"))))

    b = ctx.run_block(None, L=L, device_ptr=d_ascii, n_reads=a.reads)
    plain_out = outcome(b)
    b.free()
    b = ctx.run_block(None, L=L, device_ptr=d_ascii, n_reads=a.reads, learned=True)  # warm-up
    b.free()
    ctx.reset_stats()
    ctx.set_timing(True)
    ctx.device_sync()
    t0 = time.perf_counter()
    for k in range(a.steps):
        b = ctx.run_block(None, L=L, device_ptr=d_ascii, n_reads=a.reads, learned=True)
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
    ctx.set_scorer(None)
    net.close()
    pm = sum(s["ms"] for s in plain_stats if s["name"] == "mapping") / a.steps
    log(f"learned mapping: {ms:.1f} ms per block (plain {plain_ms:.1f}); mapping {st.get('mapping_learned', 0):.1f} ms "
        f"(plain {pm:.1f}); mapped {out['mapped']} (plain {plain_out['mapped']}), mismatches {out['mismatches']} "
        f"(plain {plain_out['mismatches']})")
    return dict(ms_per_step=round(ms, 3), gbases_s=round(a.reads * L / 1e9 / (ms / 1e3), 4),
                speedup_vs_plain=round(plain_ms / ms, 4), mapping_ms=st.get("mapping_learned"),
                mapping_ms_plain=round(pm, 3), threshold=thr, outcome=out, outcome_plain=plain_out,
                stages_ms=st, mapping_work=work)


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
    out = dict(inputs=len(ins), gbases=round(gbases, 4), sequential_s=round(seq, 4), batched_s=round(bat, 4),
               sequential_gbases_s=round(gbases / seq, 4), batched_gbases_s=round(gbases / bat, 4),
               speedup=round(seq / bat, 3), mapped_reads=plain_mapped, workers=workers,
               workers_tuning_s=wtimes,
               workload="64 inputs: genomes 1-50 Mbp, L in {75,100,150,250}, 10k-5M reads log-uniform, "
                        "0.3-1% subst, input in host memory")
    if lb is not None:
        out.update(learned_batched_s=round(lb, 4), learned_batched_gbases_s=round(gbases / lb, 4),
                   learned_batched_vs_plain=round(seq / lb, 3), learned_batched_vs_batched=round(bat / lb, 3),
                   learned_mapped_reads=learned_mapped)
    return out


def cpu_baseline(n_reads, L):
    """Times the OpenMP build of the oracle on a scaled sample of the same
    workload (same coverage, error and N profile) and cross-checks the GPU
    path on it."""
    from oracle import oracle as orc

    genome = max(1_000_000, n_reads * L // 30)
    cfg = ca.config_scale(seed=23, n_reads=n_reads, genome_len=genome)
    _, reads = ca.synth_reads(cfg)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    ob = orc.OracleBlock(reads, stop="lq", kind="omp")
    dt = time.perf_counter() - t0
    with ca.Context(0) as ctx:
        gb = ctx.run_block(reads)
        same = all(np.array_equal(gb.get(f), ob.get(f)) for f in
                   ("hq_pg", "hq_rec", "map_rec", "map_mis", "unmapped", "unmapped_N", "lq_pg", "lq_rec"))
    gbases = n_reads * L / 1e9
    return dict(value=round(gbases / dt, 6), unit="Gbases/s", cores=threads, kind="port",
                sample=f"oracle_omp on {n_reads} reads x {L} bp from a {genome} bp genome "
                       f"(same 30x coverage and error profile), {dt:.1f} s",
                seconds=round(dt, 3), gpu_matches_oracle=bool(same))


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
    t0 = time.perf_counter()
    _, reads = ca.synth_reads(cfg)
    log(f"generated {a.reads} reads in {time.perf_counter() - t0:.1f} s")
    ctx = ca.Context(0)
    nbytes = reads.nbytes
    d_ascii = ctx.dev_alloc(nbytes)
    ctx.h2d(d_ascii, reads)
    del reads
    ctx.device_sync()

    def step():
        b = ctx.run_block(None, L=L, device_ptr=d_ascii, n_reads=a.reads)
        return b

    tuning = None
    if not a.no_tune:
        # launch shapes only (results do not depend on them), chosen on this
        # block before the timed region
        from curc_amd import autotune

        t0 = time.perf_counter()
        best, table = autotune.tune_block(ctx, lambda: step().free(), reps=2, log=log)
        tuning = dict(best, seconds=round(time.perf_counter() - t0, 1), candidates_ms=table)

    for i in range(a.warmup):
        t = time.perf_counter()
        b = step()
        log(f"warmup {i}: {time.perf_counter() - t:.3f} s; HQ records {len(b.get('hq_rec'))}, "
            f"HQ PG ~{len(b.get('hq_pg2')) * 32} bases, mapped {len(b.get('map_rec'))}, "
            f"LQ records {len(b.get('lq_rec'))}, N-LQ records {len(b.get('nlq_rec'))}")
        b.free()
    ctx.reset_stats()
    ctx.set_timing(True)
    ctx.device_sync()
    t0 = time.perf_counter()
    for i in range(a.steps):
        b = step()
        b.free()
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
    log("stage breakdown (ms per step):")
    for s in sorted(stats, key=lambda s: -s["ms"]):
        log(f"  {s['name']:<14} {s['ms'] / a.steps:9.2f} ms  ({s['launches'] // max(1, a.steps)} regions/step)")
    log("mapping work per step (both strand passes): " + ", ".join(f"{k[4:]} {v}" for k, v in work.items()))
    roof = dominant_roofline(stats, a.steps, default_workload=(a.reads, a.genome) == (20_000_000, 100_000_000))
    if roof:
        # the measured ceiling beside the vendor peak: a plain copy of 2 GiB
        copy = ctx.stream_copy_gbs()
        roof["measured_copy_GBs"] = round(copy, 1)
        roof["frac_of_measured_copy"] = round(roof["achieved"] / copy, 5)
        if roof.get("traffic"):
            roof["traffic_GBs"] = round(roof["traffic"] / (roof["avg_launch_ms"] * 1e-3) / 1e9, 1)
        log(f"stream copy: {copy:.0f} GB/s measured (vendor peak {HBM_PEAK_GBS:.0f} GB/s)")
    learned = None
    if not a.no_learned:
        learned = learned_run(ctx, a, L, d_ascii, ms, stats)
    full = None
    if not a.no_full_block:
        # the same block once more with the stream builders and the
        # pseudogenome self-match (outside the metric's region)
        ctx.reset_stats()
        ctx.set_timing(True)
        ctx.device_sync()
        t = time.perf_counter()
        b = ctx.run_block(None, L=L, device_ptr=d_ascii, n_reads=a.reads, streams=True, refmatch=True)
        ctx.device_sync()
        full_ms = (time.perf_counter() - t) * 1e3
        ctx.set_timing(False)
        fst = [s for s in ctx.stats() if s["ms"] > 0]
        full = {"ms": round(full_ms, 3), "stages_ms": {s["name"]: round(s["ms"], 3) for s in fst},
                "pg_bases": {p: int(len(b.get(p + "_pg"))) for p in ("hq", "lq", "nlq")},
                "pg_bases_after_refmatch": {p: int(len(b.get(f"rm_{p}_text"))) for p in ("hq", "lq", "nlq")}}
        log(f"block with streams + self-match: {full_ms:.1f} ms; " +
            ", ".join(f"{k} {v:.1f}" for k, v in full["stages_ms"].items() if k in ("streams", "refmatch")))
        b.free()
    ctx.dev_free(d_ascii)
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
    cpu = None
    if not a.no_cpu_baseline:
        log("timing the CPU oracle on a bounded sample ...")
        cpu = cpu_baseline(a.cpu_reads, L)
        log(f"  {cpu['sample']}: {cpu['value']} Gbases/s, GPU==oracle: {cpu['gpu_matches_oracle']}")
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
                               "subst 0.1%->1.5%, N 0.02%, one block",
                   "reads": a.reads, "read_len": L, "genome_bp": a.genome},
        "stages_ms": {s["name"]: round(s["ms"] / a.steps, 3) for s in stats},
        "mapping_work": work,
        "offset_loops": graphs,
        "tuning": tuning,
        "roofline": roof,
        "cpu_baseline": cpu,
        "bits_per_base": bpb,
        "full_block": full,
        "batched": bat,
        "learned": learned,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
