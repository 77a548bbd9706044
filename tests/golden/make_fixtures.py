"""Writes tests/golden/blocks.npz: small seeded blocks and the oracle's
outputs for them (regression vectors for the oracle itself and, on a GPU,
for the HIP path; the reference has no golden vectors of its own).

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import curc_amd as ca  # noqa: E402
from oracle import oracle as orc  # noqa: E402

FIELDS = ("sorted_ids", "next", "prev", "offset", "hq_pg", "hq_rec", "map_rec", "map_mis", "unmapped", "unmapped_N",
          "lq_pg", "lq_rec", "nlq_pg", "nlq_rec")

CASES = {
    "l150": dict(seed=1201, genome_len=6_000, gc=0.5, n_reads=900, L=150, sub_first=0.002, sub_last=0.01,
                 n_rate=0.002),
    "l75_rc": dict(seed=1202, genome_len=3_000, gc=0.45, n_reads=700, L=75, rc_prob=0.5, sub_first=0.005,
                   sub_last=0.005, n_rate=0.003),
    "l100_repeats": dict(seed=1203, genome_len=8_000, gc=0.45, repeat_frac=0.4, rep_min=300, rep_max=1200,
                         div_min=0.02, div_max=0.05, rep_families=3, n_reads=800, L=100, sub_first=0.003,
                         sub_last=0.003, n_rate=0.001),
}


def main():
    out = {}
    for name, kw in CASES.items():
        _, reads = ca.synth_reads(ca.synth_config(**kw))
        ob = orc.OracleBlock(reads, stop="all")
        out[f"{name}/reads"] = reads
        for f in FIELDS:
            v = ob.get(f)
            out[f"{name}/{f}"] = v.view(np.uint8).reshape(-1) if v.dtype.names else v
    np.savez_compressed(os.path.join(HERE, "blocks.npz"), **out)
    print("wrote", os.path.join(HERE, "blocks.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
