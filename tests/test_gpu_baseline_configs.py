"""The five benchmark configurations on the GPU, bit-exact against the oracle.

Each test runs the seeded generator of pairwise_aligner_amd/workloads.py with the
configuration's own scoring and lane choice, as bench.py does, and compares every score with
the CPU oracle -- the reference's differential pattern (test/unit/affine/
alignment_simd_test_template.hpp:195-227) at the benchmark's shapes:

  DNA 150 nt global     all 10,000 pairs, int16 lanes
  protein local         the first 5,000 of the 100,000 pairs, BLOSUM62, int16 + int32 recompute
  uneven batch          the first 20,000 of the 1,000,000 pairs (half global, half local),
                        autotuned, main pass replayed from a hipGraph three times
  long pairs            the configuration's longest pair (99,988 x 100,172) and three more,
                        int32 (as configured) and re-centred int16; the oracle's scores of these
                        pairs are a committed fixture (tests/golden/config_long_scores.json,
                        written by tests/golden/make_config_long_scores.py: one oracle thread
                        takes about a minute for the longest pair)
  mixed stream          the first 2,000 pairs of each half (DNA and protein, local) under fixed
                        int32, int16 speculation and the learned dispatcher
"""
import json
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pairwise_aligner_amd import device, learned, tables, workloads

pytestmark = pytest.mark.gpu

THREADS = max(1, min(os.cpu_count() or 1, 16))
GOLDEN_LONG = os.path.join(os.path.dirname(__file__), "golden", "config_long_scores.json")


def log(msg):
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return device.default_context(0)


def blosum(method, policy=device.LANES_AUTO):
    t = tables.blosum62_standard()
    return device.DeviceConfig(method=method, score_kind=device.SCORE_MATRIX, symbols="".join(s for s, _ in t).encode(),
                               matrix=tuple(v for _, r in t for v in r), gap_open=-11, gap_extension=-1,
                               lane_policy=policy)


def oracle_cfg(dc):
    return O.OracleConfig(mode=dc.method, first_column_free=dc.first_column_free, first_row_free=dc.first_row_free,
                          last_column_free=dc.last_column_free, last_row_free=dc.last_row_free,
                          gap_open=dc.gap_open, gap_extension=dc.gap_extension, score_kind=dc.score_kind,
                          match=dc.match, mismatch=dc.mismatch, symbols=dc.symbols, matrix=dc.matrix)


def assert_oracle(got, b1, o1, b2, o2, dc, pm=None, label=""):
    want = O.score_packed(b1, o1, b2, o2, oracle_cfg(dc), pair_method=pm, threads=THREADS)
    bad = np.nonzero(got.astype(np.int64) != want)[0]
    assert len(bad) == 0, (f"{label}: {len(bad)} mismatches of {len(want)}; first {bad[:5].tolist()}: device "
                           f"{got[bad[:5]].tolist()} oracle {want[bad[:5]].tolist()}")


def test_config_dna150_all_pairs(ctx):
    w = workloads.config_dna150()
    dc = device.DeviceConfig(method=device.METHOD_GLOBAL, match=4, mismatch=-5, gap_open=-10, gap_extension=-1)
    log("config dna150: 10,000 pairs x 150 nt, global, int16")
    bt = device.Batch(ctx, w["b1"], w["o1"], w["b2"], w["o2"], dc)
    bt.run()
    got, flags = bt.results()
    bt.close()
    assert np.all(flags & device.PAIR_INT16)
    assert_oracle(got, w["b1"], w["o1"], w["b2"], w["o2"], dc, label="dna150")


def test_config_protein_first_5000(ctx):
    w = workloads.config_protein()
    b1, o1 = workloads.take(w["b1"], w["o1"], 5000)
    b2, o2 = workloads.take(w["b2"], w["o2"], 5000)
    dc = blosum(device.METHOD_LOCAL)
    log("config protein: first 5,000 of 100,000 pairs, local BLOSUM62")
    bt = device.Batch(ctx, b1, o1, b2, o2, dc)
    bt.run()
    got, flags = bt.results()
    bt.close()
    assert np.all(flags & device.PAIR_INT16)
    assert_oracle(got, b1, o1, b2, o2, dc, label="protein")


def test_config_uneven_first_20000_autotuned_graph(ctx):
    w = workloads.config_uneven(limit=20_000)
    dc = device.DeviceConfig(method=device.METHOD_GLOBAL, match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
    log("config uneven: first 20,000 of 1,000,000 pairs, autotuned, graph replays")
    tuning = device.tune(ctx, dc, w["b1"], w["o1"], w["b2"], w["o2"], w["pair_method"], sample_pairs=2000, iters=1)
    dct = device.DeviceConfig(**{**dc.__dict__, "tuning": device.Tuning(**{**tuning.__dict__, "eager": 0})})
    bt = device.Batch(ctx, w["b1"], w["o1"], w["b2"], w["o2"], dct, w["pair_method"])
    assert bt.stats()["n_launches"] > 2   # several variants: the pass is a captured graph
    want = O.score_packed(w["b1"], w["o1"], w["b2"], w["o2"], oracle_cfg(dc), pair_method=w["pair_method"],
                          threads=THREADS)
    for rep in range(3):
        bt.run()
        got, _ = bt.results()
        bad = np.nonzero(got.astype(np.int64) != want)[0]
        assert len(bad) == 0, f"replay {rep}: {len(bad)} mismatches, first {bad[:5].tolist()}"
    bt.close()


@pytest.mark.parametrize("policy", [device.LANES_INT32, device.LANES_AUTO], ids=["int32", "recentred-int16"])
def test_config_long_pairs(ctx, policy):
    ref = json.load(open(GOLDEN_LONG))
    w = workloads.config_long(select=ref["pairs"])
    assert workloads.digest(w["b1"], w["o1"], w["b2"], w["o2"]) == ref["inputs_sha256"]
    assert max(ref["len1"]) >= 90_000
    dc = device.DeviceConfig(method=device.METHOD_GLOBAL, match=1, mismatch=-4, gap_open=-6, gap_extension=-1,
                             lane_policy=policy)
    log(f"config long: pairs {ref['pairs']} lengths {ref['len1']}, policy {policy}")
    got, flags = ctx.score_packed(w["b1"], w["o1"], w["b2"], w["o2"], dc)
    assert got.tolist() == ref["scores"], (got.tolist(), ref["scores"])
    width = device.PAIR_INT32 if policy == device.LANES_INT32 else device.PAIR_INT16
    assert np.all(flags & width), flags


def test_config_mixed_first_2000_each_half(ctx):
    """Both halves of the mixed stream under fixed int32, int16 speculation and the learned
    dispatcher (trained here on DP labels of separately seeded pairs): identical scores, equal to
    the oracle."""
    w = workloads.config_mixed(limit=2000)
    dna = device.DeviceConfig(method=device.METHOD_LOCAL, match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
    prot = blosum(device.METHOD_LOCAL)
    sets = [(prot, learned.training_pairs("protein", 1001, 1500)), (dna, learned.training_pairs("dna", 1002, 800))]
    x, y = learned.build_dataset(ctx, sets)
    model, _ = learned.train(ctx, x, y, steps=400, batch=1024)
    for kind, dc in (("dna", dna), ("protein", prot)):
        h = w[kind]
        log(f"config mixed: {kind} half, 2,000 pairs")
        runs = {}
        for name, policy, mdl in (("int32", device.LANES_INT32, None), ("int16", device.LANES_INT16, None),
                                  ("learned", device.LANES_LEARNED, model)):
            dcp = device.DeviceConfig(**{**dc.__dict__, "lane_policy": policy, "model": mdl})
            runs[name] = ctx.score_packed(h["b1"], h["o1"], h["b2"], h["o2"], dcp)
        for name in ("int16", "learned"):
            assert np.array_equal(runs[name][0], runs["int32"][0]), f"{kind}: {name} differs from int32"
        assert np.all(runs["int32"][1] & device.PAIR_INT32)
        assert_oracle(runs["int32"][0], h["b1"], h["o1"], h["b2"], h["o2"], dc, label=f"mixed {kind}")
    model.close()
