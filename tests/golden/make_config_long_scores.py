"""Writes tests/golden/config_long_scores.json: the oracle's scores of four pairs of the long-pair
configuration (2,000 DNA pairs 10-100 kbp, 15% divergence; global, +1/-4, gap (-6,-1)) -- the
longest pair, the shortest and two more -- with the SHA-256 of the packed inputs, so that the GPU
test can check the wavefront kernel on the configuration's own longest pair without running a
one-thread oracle over 10^10 cells on the GPU host.

    python tests/golden/make_config_long_scores.py

The scores come from the CPU oracle (oracle/gotoh_oracle.c, 64-bit DP), nothing else.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pairwise_aligner_amd import workloads  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config_long_scores.json")


def main():
    lengths = workloads.long_lengths()
    order = np.argsort(lengths, kind="stable")
    select = [int(order[-1]), int(order[0]), 0, 777]
    w = workloads.config_long(select=select)
    cfg = O.OracleConfig(mode=O.GLOBAL, match=1, mismatch=-4, gap_open=-6, gap_extension=-1)
    scores = O.score_packed(w["b1"], w["o1"], w["b2"], w["o2"], cfg, threads=len(select))
    rec = {"config": "long: 2,000 DNA pairs 10-100 kbp (workloads.config_long), global, +1/-4, gap (-6,-1)",
           "pairs": select, "len1": np.diff(w["o1"]).astype(int).tolist(), "len2": np.diff(w["o2"]).astype(int).tolist(),
           "inputs_sha256": workloads.digest(w["b1"], w["o1"], w["b2"], w["o2"]),
           "scores": [int(s) for s in scores]}
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
