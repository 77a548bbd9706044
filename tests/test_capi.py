"""CPU-side checks of libcurc_gpu.so: it loads, exports every function that
include/*.h declares, derives the same parameters as the oracle, refuses to
run without a device (no CPU fallback), and its synthetic generator is
deterministic."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import curc_amd as ca
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(curc_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ca.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_params_match_oracle():
    for L in range(1, 513):
        try:
            p = ca.make_params(L)
        except ca.CurcError:
            with pytest.raises(ValueError):
                orc.params(L)
            continue
        q = orc.params(L)
        assert (p.W, p.max_off, p.k, p.kip, p.max_mm) == (q.W, q.max_off, q.k, q.kip, q.max_mm)
    with pytest.raises(ca.CurcError):
        ca.make_params(513)
    with pytest.raises(ca.CurcError):
        ca.make_params(0)


@pytest.mark.skipif(ca.device_count() > 0, reason="a HIP device is visible")
def test_no_device_fails_loudly():
    with pytest.raises(ca.CurcError):
        ca.Context(0)


def test_synth_deterministic_and_shaped():
    cfg = ca.synth_config(seed=3, genome_len=200_000, n_reads=20_000, L=150, sub_first=0.001, sub_last=0.015,
                          n_rate=0.001, gc=0.41)
    g1, r1 = ca.synth_reads(cfg)
    g2, r2 = ca.synth_reads(cfg)
    assert np.array_equal(g1, g2) and np.array_equal(r1, r2)
    assert r1.shape == (20_000, 150)
    assert set(np.unique(r1).tolist()) <= set(b"ACGTN")
    gc = np.isin(g1, np.frombuffer(b"GC", dtype=np.uint8)).mean()
    assert abs(gc - 0.41) < 0.01
    nfrac = (r1 == ord("N")).mean()
    assert 0.0005 < nfrac < 0.002
    cfg2 = ca.synth_config(seed=4, genome_len=200_000, n_reads=20_000, L=150)
    assert not np.array_equal(ca.synth_reads(cfg2)[1], r1)


def test_synth_reads_come_from_the_genome():
    cfg = ca.synth_config(seed=8, genome_len=5_000, n_reads=200, L=100, sub_first=0.0, sub_last=0.0,
                          n_rate=0.0, rc_prob=0.5)
    g, reads = ca.synth_reads(cfg)
    gs = g.tobytes().decode()
    comp = str.maketrans("ACGT", "TGCA")
    rc = 0
    for r in reads:
        s = r.tobytes().decode()
        if s in gs:
            continue
        assert s[::-1].translate(comp) in gs
        rc += 1
    assert 60 < rc < 140


def test_synth_paired_end_fragments():
    cfg = ca.synth_config(seed=9, genome_len=20_000, n_reads=400, L=100, paired=1, frag_mean=400,
                          frag_sd=50, sub_first=0.0, sub_last=0.0, n_rate=0.0)
    g, reads = ca.synth_reads(cfg)
    gs = g.tobytes().decode()
    comp = str.maketrans("ACGT", "TGCA")
    for k in range(0, 400, 2):
        a, b = reads[k].tobytes().decode(), reads[k + 1].tobytes().decode()
        fwd = [x for x in (a, b) if x in gs]
        rev = [x for x in (a, b) if x[::-1].translate(comp) in gs]
        assert len(fwd) >= 1 and len(rev) >= 1


def test_scorer_parameter_layout_matches_reference():
    import re

    from curc_amd import scorer
    from oracle import mlp

    hdr = open(os.path.join(ROOT, "include", "curc_gpu.h")).read()
    n = int(re.search(r"#define CURC_MLP_NPARAMS (\d+)u", hdr).group(1))
    assert n == scorer.N_PARAMS == mlp.N_PARAMS == scorer.init_params(0).size


def test_autotune_cache_round_trip(tmp_path):
    from curc_amd import autotune

    assert autotune.size_class(1) == 1 and autotune.size_class(5) == 8 and autotune.size_class(1 << 20) == 1 << 20
    p = str(tmp_path / "t.json")
    assert autotune.load(p, 150, 1000) is None
    t = dict(ov_grid=4096, fm_grid=2048, map_grid=32768, radix_items=16, workers=3, filter_bits=32, fm_fused=4096)
    autotune.save(p, 150, 1000, t)
    autotune.save(p, 100, 1000, dict(t, workers=1))
    assert autotune.load(p, 150, 1024) == t
    assert autotune.load(p, 150, 1025) is None
    assert autotune.load(p, 100, 900)["workers"] == 1
