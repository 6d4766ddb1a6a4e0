"""GPU: seeded random configurations against the oracle, bit-exact.

Each case draws a configuration the way a user could write one -- method (or per-pair methods),
the four end-gap rules, gap open and extension in [-15, 0] and [-5, 0] (zero included), unitary
scores or BLOSUM62 (standard or extended), a lane policy -- and a batch whose lengths mix empty,
short, medium and long sequences (beyond the wavefront kernel's threshold now and then), related
and unrelated pairs.  It runs through the one-to-one entry point, for one case in three the
one-to-many entry point with the first sequence as the query, and for one seed in four the DP
export (final DP column and row of every pair).  The reference's differential tests
(alignment_simd_test_template.hpp:195-227) check SIMD against scalar results the same way.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pairwise_aligner_amd import device, tables, workloads

pytestmark = pytest.mark.gpu

# 64 seeded cases by default; PA_RANDOM_CASES widens the sweep (the seeds of the default run stay
# the first ones).
N_CASES = int(os.environ.get("PA_RANDOM_CASES", "64"))


def log(msg):
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return device.Context(0)


def random_case(seed):
    rng = np.random.default_rng(7000 + seed)
    method = int(rng.integers(0, 2))
    fc, fr, lc, lr = (bool(x) for x in rng.integers(0, 2, size=4))
    go, ge = -int(rng.integers(0, 16)), -int(rng.integers(0, 6))
    policy = [device.LANES_AUTO, device.LANES_INT16, device.LANES_INT32][int(rng.integers(0, 3))]
    kind = int(rng.integers(0, 3))
    if kind == 0:
        alphabet = workloads.DNA
        dc = device.DeviceConfig(method=method, first_column_free=fc, first_row_free=fr, last_column_free=lc,
                                 last_row_free=lr, gap_open=go, gap_extension=ge, match=int(rng.integers(0, 11)),
                                 mismatch=-int(rng.integers(0, 11)), lane_policy=policy)
    else:
        t = tables.blosum62_standard() if kind == 1 else tables.blosum62_extended()
        symbols = "".join(s for s, _ in t).encode()
        alphabet = np.frombuffer(symbols, dtype=np.uint8)
        dc = device.DeviceConfig(method=method, score_kind=device.SCORE_MATRIX, symbols=symbols,
                                 matrix=tuple(v for _, r in t for v in r), first_column_free=fc, first_row_free=fr,
                                 last_column_free=lc, last_row_free=lr, gap_open=go, gap_extension=ge,
                                 lane_policy=policy)
    n = int(rng.integers(50, 400))
    lens = rng.choice([0, 1, 5, 40, 150, 400, 900], size=n, p=[0.04, 0.04, 0.1, 0.3, 0.3, 0.17, 0.05])
    lens = (lens * rng.uniform(0.5, 1.5, size=n)).astype(np.int64)
    if rng.random() < 0.3:
        lens[int(rng.integers(0, n))] = int(rng.integers(3200, 4200))   # a wavefront-kernel pair
    s1 = workloads.unpack(*workloads.random_packed(rng, lens, alphabet))
    related = rng.random(n) < 0.5
    s2 = []
    for k in range(n):
        if related[k] and len(s1[k]):
            b, o = workloads.mutate_packed(rng, *workloads.pack([s1[k]]), np.array([rng.uniform(0.01, 0.3)]), 0.3,
                                           0.6, alphabet)
            s2.append(workloads.unpack(b, o)[0])
        else:
            L = int(max(0, lens[int(rng.integers(0, n))] * rng.uniform(0.5, 1.5)))
            s2.append(workloads.unpack(*workloads.random_packed(rng, np.array([L]), alphabet))[0])
    pm = (rng.random(n) < 0.5).astype(np.uint8) if rng.random() < 0.3 else None
    return dc, s1, s2, pm, bool(rng.random() < 0.34)


def oracle_cfg(dc):
    return O.OracleConfig(mode=dc.method, first_column_free=dc.first_column_free, first_row_free=dc.first_row_free,
                          last_column_free=dc.last_column_free, last_row_free=dc.last_row_free,
                          gap_open=dc.gap_open, gap_extension=dc.gap_extension, score_kind=dc.score_kind,
                          match=dc.match, mismatch=dc.mismatch, symbols=dc.symbols, matrix=dc.matrix)


@pytest.mark.parametrize("seed", range(N_CASES))
def test_random_configuration(ctx, seed):
    dc, s1, s2, pm, one_to_many = random_case(seed)
    log(f"random case {seed}: {len(s1)} pairs, one_to_many={one_to_many}, per-pair methods={pm is not None}, "
        f"cfg={dc}")
    b1, o1 = workloads.pack(s1)
    b2, o2 = workloads.pack(s2)
    want = O.score_packed(b1, o1, b2, o2, oracle_cfg(dc), pair_method=pm, threads=8)
    got, flags = ctx.score_packed(b1, o1, b2, o2, dc, pm)
    bad = np.nonzero(got.astype(np.int64) != want)[0]
    assert len(bad) == 0, (f"case {seed}: {len(bad)} mismatches, first {bad[:5].tolist()}: device "
                           f"{got[bad[:5]].tolist()} oracle {want[bad[:5]].tolist()}")
    if one_to_many:
        q = s1[0]
        rep1, rep_o1 = workloads.pack([q] * len(s2))
        want_q = O.score_packed(rep1, rep_o1, b2, o2, oracle_cfg(dc), pair_method=pm, threads=8)
        got_q, _ = ctx.score_one_to_many(q, b2, o2, dc, pm)
        assert np.array_equal(got_q.astype(np.int64), want_q), f"case {seed}: one-to-many"
    if seed % 4 == 1:
        want_s, want_rows, want_cols, want_rg, want_cg = O.dp_vectors_packed(b1, o1, b2, o2, oracle_cfg(dc),
                                                                             pair_method=pm, threads=8, gaps=True)
        got_s, _, cols, rows, cg, rg = ctx.score_packed_dp(b1, o1, b2, o2, dc, pm, gaps=True)
        assert np.array_equal(got_s.astype(np.int64), want_s), f"case {seed}: DP export scores"
        assert np.array_equal(cols.astype(np.int64), want_cols), f"case {seed}: DP export columns"
        assert np.array_equal(rows.astype(np.int64), want_rows), f"case {seed}: DP export rows"
        assert np.array_equal(cg.astype(np.int64), want_cg), f"case {seed}: DP export column gap states"
        assert np.array_equal(rg.astype(np.int64), want_rg), f"case {seed}: DP export row gap states"
