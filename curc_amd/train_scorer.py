"""Trains the mapping scorer on the GPU and writes curc_amd/models/.

    python -m curc_amd.train_scorer [--records 40000000] [--epochs 3] [--out DIR]

Training data come from traced mapping runs (CURC_RUN_TRACE) over synthetic
blocks whose seeds are disjoint from every workload the benchmark and the
tests evaluate on: at each checkpoint of a strand pass a sampled read
contributes its 32 features and the label "the rest of the pass lowered the
read's best mismatch count", as decided by exact verification.  The network
is trained with the library's own bfloat16 MFMA forward / backward kernels
and Adam.  There are two stop thresholds: for reads that already have a
mapping, the logit that keeps `--recall` of their positive validation
records above it; for reads without one (feature 17 set), the logit that
keeps `--recall-unmapped` of theirs.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import Context, config_scale, synth_config, synth_reads
from .scorer import DeviceArray, Mlp, init_params

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_DIR = os.path.join(HERE, "models")
WEIGHTS = "map_scorer.f32"
META = "map_scorer.json"

# training workloads: seeds 7001.. (evaluation uses seeds below 1000)
TRAIN_SETS = (
    ("100 Mbp-profile sample", lambda s: config_scale(seed=s, n_reads=3_000_000, genome_len=15_000_000)),
    ("1 Mbp-profile", lambda s: synth_config(seed=s, genome_len=1_000_000, gc=0.5, n_reads=133_333, L=150,
                                             rc_prob=0.5, sub_first=0.005, sub_last=0.005, n_rate=0.0001)),
    ("short reads", lambda s: synth_config(seed=s, genome_len=2_000_000, gc=0.45, n_reads=400_000, L=100,
                                           sub_first=0.003, sub_last=0.012, n_rate=0.0002)),
    ("75 bp, uniform 1% errors", lambda s: synth_config(seed=s, genome_len=3_000_000, gc=0.45, n_reads=600_000,
                                                        L=75, sub_first=0.01, sub_last=0.01, n_rate=0.0002)),
    ("250 bp, uniform 0.6% errors", lambda s: synth_config(seed=s, genome_len=8_000_000, gc=0.45,
                                                           n_reads=400_000, L=250, sub_first=0.006,
                                                           sub_last=0.006, n_rate=0.0002)),
    ("150 bp, low coverage", lambda s: synth_config(seed=s, genome_len=20_000_000, gc=0.45, n_reads=300_000,
                                                    L=150, sub_first=0.008, sub_last=0.008, n_rate=0.0002)),
)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def collect(ctx, target, every=4):
    xs, ys = [], []
    total = 0
    seed = 7001
    ctx.set_trace(every=every, max_records=1 << 24)
    while total < target:
        for name, make in TRAIN_SETS:
            _, reads = synth_reads(make(seed))
            blk = ctx.run_block(reads, trace=True)
            x, y = blk.get("trace_x"), blk.get("trace_y")
            blk.free()
            xs.append(x)
            ys.append(y)
            total += len(y)
            log(f"  traced {name} (seed {seed}): {len(y)} records, {y.mean():.3f} positive; total {total}")
            seed += 1
            if total >= target:
                break
    return np.concatenate(xs)[:target], np.concatenate(ys)[:target]


def train(ctx, x, y, epochs, batch, lr, seed=0):
    rng = np.random.default_rng(seed)
    net = Mlp(ctx, init_params(seed))
    n = len(y)
    dx = DeviceArray(ctx, x)
    dy = DeviceArray(ctx, y)
    steps = 0
    t0 = time.perf_counter()
    for ep in range(epochs):
        # batches are contiguous slices of a per-epoch shuffled copy
        perm = rng.permutation(n)
        xs = DeviceArray(ctx, x[perm])
        ys = DeviceArray(ctx, y[perm])
        losses = []
        for b0 in range(0, n - batch + 1, batch):
            want = (steps % 200) == 0
            l = net.train_step_dev(xs.ptr + b0 * 128, ys.ptr + b0 * 4, batch, lr, want_loss=want)
            if want:
                losses.append(l)
            steps += 1
        xs.free()
        ys.free()
        log(f"  epoch {ep}: {steps} steps, sampled loss {np.mean(losses):.4f}")
    ctx.device_sync()
    dt = time.perf_counter() - t0
    dx.free()
    dy.free()
    return net, steps, dt


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--records", type=int, default=40_000_000)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--recall", type=float, default=0.95)
    ap.add_argument("--recall-unmapped", type=float, default=0.9999)
    ap.add_argument("--out", default=MODEL_DIR)
    ap.add_argument("--calibrate", action="store_true",
                    help="keep the trained weights in --out and only choose the threshold for --recall on "
                         "one fresh traced block of each training workload (seeds 7990+)")
    a = ap.parse_args(argv)
    if a.calibrate:
        return calibrate(a.out, a.recall, a.recall_unmapped)
    with Context(0) as ctx:
        log("collecting traced mapping records ...")
        x, y = collect(ctx, a.records)
        rng = np.random.default_rng(1)
        perm = rng.permutation(len(y))
        nval = max(10_000, len(y) // 20)
        xv, yv = x[perm[:nval]], y[perm[:nval]]
        xt, yt = x[perm[nval:]], y[perm[nval:]]
        log(f"training on {len(yt)} records ({yt.mean():.3f} positive), validating on {len(yv)}")
        net, steps, dt = train(ctx, xt, yt, a.epochs, a.batch, a.lr)
        flops = 6.0 * 20865 * len(yt) * a.epochs  # forward + backward per sample
        logit = net.forward(xv)
        thr, thr_un = thresholds(xv, yv, logit, a.recall, a.recall_unmapped)
        stop = float(np.mean(logit < np.where(xv[:, UNMAPPED_FEATURE] > 0.5, thr_un, thr)))
        acc = float(np.mean((logit > 0) == (yv > 0.5)))
        lv = float(np.mean(np.logaddexp(0, logit) - yv * logit))
        params = net.get_params()
        net.close()
    os.makedirs(a.out, exist_ok=True)
    params.astype("<f4").tofile(os.path.join(a.out, WEIGHTS))
    meta = dict(threshold=thr, threshold_unmapped=thr_un, recall=a.recall, recall_unmapped=a.recall_unmapped,
                records=int(len(y)), positive=float(y.mean()),
                val_loss=lv, val_accuracy=acc, val_stop_fraction=stop, epochs=a.epochs, batch=a.batch, lr=a.lr,
                steps=steps, train_seconds=round(dt, 3), train_tflops=round(flops / dt / 1e12, 3),
                features=32, layers=[32, 128, 128, 1],
                data="traced mapping of synthetic blocks, seeds 7001+, disjoint from evaluation")
    with open(os.path.join(a.out, META), "w") as f:
        json.dump(meta, f, indent=1)
    log(json.dumps(meta))
    return 0


# feature 17 is 1 for a read with no mapping within max_mismatch yet
UNMAPPED_FEATURE = 17


def thresholds(x, y, logit, recall, recall_unmapped):
    """(threshold for reads with a mapping, threshold for reads without):
    each keeps the given fraction of its group's positive records."""
    out = []
    un = x[:, UNMAPPED_FEATURE] > 0.5
    for group, r in ((~un, recall), (un, recall_unmapped)):
        pos = np.sort(logit[group & (y > 0.5)])
        out.append(float(pos[int((1 - r) * len(pos))]) if len(pos) else -1e30)
    return tuple(out)


def calibrate(model_dir, recall, recall_unmapped):
    params, _ = load(model_dir)
    xs, ys = [], []
    with Context(0) as ctx:
        ctx.set_trace(every=4, max_records=1 << 24)
        for k, (name, make) in enumerate(TRAIN_SETS):
            _, reads = synth_reads(make(7990 + k))
            blk = ctx.run_block(reads, trace=True)
            xs.append(blk.get("trace_x"))
            ys.append(blk.get("trace_y"))
            blk.free()
        x, y = np.concatenate(xs), np.concatenate(ys)
        net = Mlp(ctx, params)
        logit = net.forward(x)
        net.close()
    thr, thr_un = thresholds(x, y, logit, recall, recall_unmapped)
    with open(os.path.join(model_dir, META)) as f:
        meta = json.load(f)
    stop = logit < np.where(x[:, UNMAPPED_FEATURE] > 0.5, thr_un, thr)
    meta.update(threshold=thr, threshold_unmapped=thr_un, recall=recall, recall_unmapped=recall_unmapped,
                calib_records=int(len(y)), calib_stop_fraction=float(np.mean(stop)))
    with open(os.path.join(model_dir, META), "w") as f:
        json.dump(meta, f, indent=1)
    log(json.dumps(meta))
    return 0


def load(model_dir=MODEL_DIR):
    """(params float32 [20865], (threshold for reads with a mapping,
    threshold for reads without)) of the trained scorer."""
    p = np.fromfile(os.path.join(model_dir, WEIGHTS), dtype="<f4")
    with open(os.path.join(model_dir, META)) as f:
        meta = json.load(f)
    thr = float(meta["threshold"])
    return p, (thr, float(meta.get("threshold_unmapped", thr)))


if __name__ == "__main__":
    sys.exit(main())
