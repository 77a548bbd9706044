"""Launch shapes (curc_tuning) change speed only: a block gives the same
results under every tuning the autotuner may pick, invalid settings are
refused, and the autotuner leaves a valid setting on the context."""
import numpy as np
import pytest

import curc_amd as ca
from curc_amd import autotune
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("sorted_ids", "next", "prev", "offset", "hq_pg", "hq_rec", "map_rec", "map_mis", "unmapped",
          "unmapped_N", "lq_pg", "lq_rec", "nlq_pg", "nlq_rec")


def reads_of(seed, n, L, genome):
    cfg = ca.synth_config(seed=seed, genome_len=genome, gc=0.45, n_reads=n, L=L, sub_first=0.003, sub_last=0.01,
                          n_rate=0.001)
    return ca.synth_reads(cfg)[1]


TUNINGS = [
    dict(ov_grid=1, fm_grid=1, map_grid=1, radix_items=8, filter_bits=8, fm_fused=0, ov_queue=1, fm_static=1),
    dict(ov_grid=7, fm_grid=3, map_grid=5, radix_items=32, filter_bits=128, fm_fused=1 << 20, ov_queue=0),
    dict(ov_grid=16384, fm_grid=16384, map_grid=32768, radix_items=16, fm_fused=4096, ov_queue=0, fm_static=0),
    dict(ov_grid=4096, fm_grid=4096, map_grid=1 << 20, radix_items=16, fm_fused=0, ov_queue=1, fm_static=0),
    dict(fm_grid=5, fm_fused=0, fm_static=1, filter_bits=8),
    dict(ov_grid=3, filter_bits=16, ov_queue=1),
    dict(ov_grid=1 << 20, filter_bits=16, ov_queue=1),
    dict(ov_grid=16384, filter_bits=8, ov_queue=2),
    dict(ov_grid=5, filter_bits=64, ov_queue=2, map_both=1),
    dict(map_grid=3, map_both=1),
    dict(map_grid=2, map_both=2),
    dict(map_grid=2, map_both=0),
    dict(ov_grid=3, filter_bits=8, ov_queue=3),
    dict(ov_grid=1 << 20, filter_bits=128, ov_queue=3, map_both=1),
]


@pytest.mark.parametrize("L", [100, 150])
def test_results_do_not_depend_on_tuning(L):
    r = reads_of(930 + L, 9_000, L, 70_000)
    ref = orc.OracleBlock(r, stop="all")
    with ca.Context(0) as ctx:
        for t in TUNINGS:
            print(f"tuning {t}", flush=True)
            ctx.set_tuning(**t)
            b = ctx.run_block(r, keep_stages=True)
            for f in FIELDS:
                assert np.array_equal(b.get(f), ref.get(f)), f"{t}: {f} differs from the oracle"
            b.free()


def test_invalid_tuning_is_refused(gpu_ctx):
    before = gpu_ctx.get_tuning()
    for bad in (dict(radix_items=12), dict(map_grid=0), dict(workers=5), dict(filter_bits=24), dict(filter_bits=4),
                dict(ov_queue=4), dict(map_both=3), dict(fq_host=65), dict(fm_static=2)):
        with pytest.raises(ca.CurcError):
            gpu_ctx.set_tuning(**bad)
    assert gpu_ctx.get_tuning() == before


def test_autotune_sets_a_candidate():
    r = reads_of(940, 20_000, 150, 150_000)
    with ca.Context(0) as ctx:
        best, table = autotune.tune_block(ctx, lambda: ctx.run_block(r).free(), reps=1)
        assert ctx.get_tuning() == best
        for knob, values in autotune.BLOCK_KNOBS:
            assert best[knob] in values
            assert len(table[knob]) == len(values)
        pick, times = autotune.tune_workers(ctx, [reads_of(941 + i, 3_000, 100, 30_000) for i in range(4)],
                                            values=(1, 2))
        assert ctx.get_tuning()["workers"] == pick


def test_stream_copy_ceiling(gpu_ctx):
    """The copy kernel behind the measured bandwidth ceiling moves the data
    and reports a plausible rate; misaligned sizes are refused."""
    n = 1 << 24
    src = gpu_ctx.dev_alloc(n)
    dst = gpu_ctx.dev_alloc(n)
    try:
        data = np.random.default_rng(3).integers(0, 256, n, dtype=np.uint8)
        gpu_ctx.h2d(src, data)
        ms = ca.ctypes.c_float()
        assert ca.lib().curc_stream_copy(gpu_ctx._h, dst, src, n, 2, ca.ctypes.byref(ms)) == ca.CURC_OK
        out = np.empty(n, dtype=np.uint8)
        gpu_ctx.d2h(out, dst)
        assert np.array_equal(out, data)
        assert ms.value > 0
        assert ca.lib().curc_stream_copy(gpu_ctx._h, dst, src, n - 8, 1, ca.ctypes.byref(ms)) != ca.CURC_OK
    finally:
        gpu_ctx.dev_free(src)
        gpu_ctx.dev_free(dst)
    assert gpu_ctx.stream_copy_gbs(1 << 28, 2) > 500
