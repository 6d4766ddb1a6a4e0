"""GPU: one query against many targets (pa_score_one_to_many, pa_batch_create_one_to_many), the
device counterpart of the reference's aligner.compute(seq1, bulk2)
(interface/interface_one_to_many_bulk.hpp:47-92) and its query profile
(score_model/score_model_matrix_simd_1xN.hpp:63-183).

Every result is checked against the CPU oracle with the query repeated as each pair's first
sequence (bit-exact), and against the one-to-one entry point on the same repeated pairs.  The
tests also check which path ran: the query profile for the configurations it covers, the
one-to-one tables where it does not apply (a free last row with a penalised last column puts
the query on the column side; a query whose profile exceeds 64 KiB of LDS; a query long enough
for the wavefront kernel).
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pairwise_aligner_amd import device, tables, workloads

pytestmark = pytest.mark.gpu


def log(msg):
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return device.Context(0)


def oracle_cfg(dc):
    return O.OracleConfig(mode=dc.method, first_column_free=dc.first_column_free, first_row_free=dc.first_row_free,
                          last_column_free=dc.last_column_free, last_row_free=dc.last_row_free,
                          gap_open=dc.gap_open, gap_extension=dc.gap_extension, score_kind=dc.score_kind,
                          match=dc.match, mismatch=dc.mismatch, symbols=dc.symbols, matrix=dc.matrix)


def blosum(method, **kw):
    t = tables.blosum62_standard()
    return device.DeviceConfig(method=method, score_kind=device.SCORE_MATRIX, symbols="".join(s for s, _ in t).encode(),
                               matrix=tuple(v for _, r in t for v in r), gap_open=-11, gap_extension=-1, **kw)


def check(ctx, query: bytes, targets, dc, pair_method=None, expect_profile=None, label=""):
    b2, o2 = workloads.pack(targets)
    n = len(targets)
    log(f"one-to-many {label}: query {len(query)}, {n} targets")
    rep1, rep_o1 = workloads.pack([query] * n)
    want = O.score_packed(rep1, rep_o1, b2, o2, oracle_cfg(dc), pair_method=pair_method, threads=8)
    got, flags = ctx.score_one_to_many(query, b2, o2, dc, pair_method)
    bad = np.nonzero(got.astype(np.int64) != want)[0]
    assert len(bad) == 0, f"{label}: {len(bad)} mismatches, first {bad[:5].tolist()}"
    one, one_flags = ctx.score_packed(rep1, rep_o1, b2, o2, dc, pair_method)
    assert np.array_equal(one, got)
    batch = device.Batch.one_to_many(ctx, query, b2, o2, dc, pair_method)
    batch.run()
    again, _ = batch.results()
    assert np.array_equal(again, got)
    st = batch.stats()
    if expect_profile is True:
        assert st["n_profiled"] > 0, f"{label}: the query profile was not used ({st})"
    elif expect_profile is False:
        assert st["n_profiled"] == 0, f"{label}: the query profile was used ({st})"
    return got, flags, st


def test_protein_local_profile(ctx):
    rng = np.random.default_rng(101)
    q = workloads.protein_pairs(rng, 1, 300, 300)[0][0]
    targets = workloads.protein_pairs(rng, 3000, 20, 600)[1]
    got, flags, st = check(ctx, q, targets, blosum(device.METHOD_LOCAL), expect_profile=True, label="protein local")
    assert st["n_profiled"] == st["n_int16"]


@pytest.mark.parametrize("ends", range(16))
def test_dna_global_end_gaps(ctx, ends):
    fc, fr, lc, lr = [(ends >> k) & 1 == 1 for k in range(4)]
    rng = np.random.default_rng(200 + ends)
    q = bytes(workloads.random_residues(rng, workloads.DNA, 180))
    targets = [bytes(t) for t in workloads.unpack(*workloads.mutate_packed(
        rng, *workloads.pack([q] * 400), rng.uniform(0.02, 0.3, size=400), 0.3, 0.7, workloads.DNA))]
    targets += [bytes(workloads.random_residues(rng, workloads.DNA, int(L))) for L in rng.integers(1, 500, size=200)]
    dc = device.DeviceConfig(match=4, mismatch=-5, gap_open=-10, gap_extension=-1, first_column_free=fc,
                             first_row_free=fr, last_column_free=lc, last_row_free=lr)
    # a free last row with a penalised last column exchanges the sequences: no profile then
    check(ctx, q, targets, dc, expect_profile=not (lr and not lc), label=f"dna global ends {ends:04b}")


@pytest.mark.parametrize("policy", [device.LANES_AUTO, device.LANES_INT16, device.LANES_INT32])
def test_lane_policies_and_saturation(ctx, policy):
    """High-scoring DNA pairs: int16 profile tasks that saturate are recomputed in int32."""
    rng = np.random.default_rng(303)
    q = bytes(workloads.random_residues(rng, workloads.DNA, 1500))
    near = [bytes(t) for t in workloads.unpack(*workloads.mutate_packed(
        rng, *workloads.pack([q] * 60), np.full(60, 0.01), 0.1, 0.7, workloads.DNA))]
    far = [bytes(workloads.random_residues(rng, workloads.DNA, int(L))) for L in rng.integers(50, 3000, size=300)]
    dc = device.DeviceConfig(match=30, mismatch=-5, gap_open=-10, gap_extension=-1, lane_policy=policy)
    got, flags, st = check(ctx, q, near + far, dc, expect_profile=None if policy != device.LANES_INT32 else False,
                           label=f"saturation policy {policy}")
    if policy == device.LANES_INT16:
        assert np.any(flags[:60] & device.PAIR_RECOMPUTED), "expected int16 saturation on the near pairs"
        assert st["n_profiled"] > 0


def test_mixed_methods_per_pair(ctx):
    rng = np.random.default_rng(404)
    q = workloads.protein_pairs(rng, 1, 250, 250)[0][0]
    targets = workloads.protein_pairs(rng, 1500, 10, 700)[1]
    pm = (np.arange(1500) % 2).astype(np.uint8)
    check(ctx, q, targets, blosum(device.METHOD_GLOBAL), pair_method=pm, expect_profile=True, label="mixed methods")


def test_fallbacks_are_exact(ctx):
    rng = np.random.default_rng(505)
    # profile over 64 KiB (21 letters + pad, ~1,600 rows): one-to-one tables
    q = workloads.protein_pairs(rng, 1, 1600, 1600)[0][0]
    targets = workloads.protein_pairs(rng, 300, 50, 400)[1]
    check(ctx, q, targets, blosum(device.METHOD_LOCAL), expect_profile=False, label="long protein query")
    # a query beyond the striped kernel's rows: the wavefront kernel
    qd = bytes(workloads.random_residues(rng, workloads.DNA, 5000))
    td = [bytes(workloads.random_residues(rng, workloads.DNA, int(L))) for L in rng.integers(100, 3000, size=40)]
    check(ctx, qd, td, device.DeviceConfig(match=2, mismatch=-3, gap_open=-5, gap_extension=-2),
          expect_profile=False, label="long DNA query")


def test_edge_cases(ctx):
    rng = np.random.default_rng(606)
    targets = [b"", b"A", b"ACGT", bytes(workloads.random_residues(rng, workloads.DNA, 300))]
    for method in (device.METHOD_GLOBAL, device.METHOD_LOCAL):
        dc = device.DeviceConfig(method=method, match=4, mismatch=-5, gap_open=-10, gap_extension=-1)
        check(ctx, b"ACGTTGCA", targets, dc, label=f"short query method {method}")
        check(ctx, b"", targets, dc, expect_profile=False, label=f"empty query method {method}")
        check(ctx, b"G", targets[:1], dc, label=f"one empty target method {method}")
    b2, o2 = workloads.pack([])
    got, _ = ctx.score_one_to_many(b"ACGT", b2, o2, device.DeviceConfig())
    assert len(got) == 0


def test_learned_policy(ctx):
    """The learned lane dispatcher reads one first sequence per pair: with one query the planner
    repeats it for the feature kernel.  An untrained network makes arbitrary int16 guesses; the
    saturation flag and the int32 recompute keep the scores exact."""
    rng = np.random.default_rng(707)
    q = workloads.protein_pairs(rng, 1, 400, 400)[0][0]
    targets = workloads.protein_pairs(rng, 800, 20, 900)[1]
    model = device.Model(ctx, seed=11)
    check(ctx, q, targets, blosum(device.METHOD_LOCAL, lane_policy=device.LANES_LEARNED, model=model),
          label="learned policy")
