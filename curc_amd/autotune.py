"""Autotuner for the library's launch shapes (curc_tuning).

The knobs change speed only, never results: workgroup caps of the overlap
probe, the exact-matching probe and the mapping kernels, keys per thread of
a radix-sort tile, the bits per prefix of the overlap k-mer filter and the
list bound of the fused exact-matching loop, the form of the overlap
probe and whether exact matching keeps its lists in place, for one block;
the number of
streams for a batch.  `tune_block` times whole blocks of a sample input under candidate
values, one knob at a time with the others fixed (coordinate descent,
best of `reps` runs per candidate), and leaves the fastest setting on the
context.  `tune_workers` does the same for curc_batch_run's stream count
on a sample batch.  Settings can be kept in a JSON file keyed by read
length and block size class (`load` / `save`).
"""
import json
import os
import time

BLOCK_KNOBS = (
    ("map_grid", (4096, 8192, 16384, 32768, 65536, 1 << 20)),
    ("ov_queue", (0, 1, 2, 3)),
    ("map_both", (0, 1, 2)),
    ("ov_grid", (1024, 2048, 4096, 16384)),
    ("fm_grid", (2048, 4096, 16384)),
    ("radix_items", (8, 16, 32)),
    ("filter_bits", (16, 32, 64)),
    ("fm_fused", (0, 1024, 4096)),
    ("fm_static", (0, 1)),
)


def _time(run, reps):
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        run()
        best = min(best, time.perf_counter() - t)
    return best


def tune_block(ctx, run, reps=3, knobs=BLOCK_KNOBS, log=None):
    """run(): one block on ctx (returns after the block is complete).
    Returns (tuning dict, {knob: {value: seconds}})."""
    run()  # grows scratch and warms the caches
    best = ctx.get_tuning()
    table = {}
    for knob, values in knobs:
        times = {}
        for v in values:
            ctx.set_tuning(**dict(best, **{knob: v}))
            times[v] = _time(run, reps)
        pick = min(times, key=times.get)
        best[knob] = pick
        table[knob] = {str(k): round(v * 1e3, 3) for k, v in times.items()}
        if log:
            log(f"  autotune {knob}: " + ", ".join(f"{k} {v * 1e3:.1f} ms" for k, v in times.items()) +
                f" -> {pick}")
    ctx.set_tuning(**best)
    return best, table


def tune_workers(ctx, inputs, values=(1, 2, 3, 4), reps=1, log=None):
    """Stream count of curc_batch_run on a sample batch (list of read arrays)."""

    def run():
        for b in ctx.batch_run(inputs, workers=0):
            b.free()

    times = {}
    for w in values:
        ctx.set_tuning(workers=w)
        run()
        times[w] = _time(run, reps)
    pick = min(times, key=times.get)
    ctx.set_tuning(workers=pick)
    if log:
        log("  autotune workers: " + ", ".join(f"{k} {v:.3f} s" for k, v in times.items()) + f" -> {pick}")
    return pick, {str(k): round(v, 4) for k, v in times.items()}


def size_class(n_reads):
    """Block size class: the power of two at or above the read count."""
    c = 1
    while c < n_reads:
        c <<= 1
    return c


def key(L, n_reads):
    return f"L{L}_n{size_class(n_reads)}"


def load(path, L, n_reads):
    """Saved tuning for this shape class, or None."""
    try:
        with open(path) as f:
            return json.load(f).get(key(L, n_reads))
    except (OSError, ValueError):
        return None


def save(path, L, n_reads, tuning):
    data = {}
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        pass
    data[key(L, n_reads)] = tuning
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
