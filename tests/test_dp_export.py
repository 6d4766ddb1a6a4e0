"""GPU: the final DP column and row of every pair, whole cells (score and gap state;
pa_score_batch_dp_cells, and pa_score_batch_dp for the scores alone), the device counterpart of
the reference's result.dp_column() / dp_row() (result/aligner_result.hpp:41-59,
affine/affine_cell.hpp:25-73).

Every exported entry is compared with the oracle's (oracle.dp_vectors_packed, pinned by
enumeration in tests/test_oracle.py), across the end-gap combinations, the local method, per-pair
methods, BLOSUM62 scoring, exchanged orientations (the planner puts the shorter sequence on the
rows where the configuration allows), empty sequences and pairs longer than the wavefront
kernel's threshold.
"""
import itertools
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


def check(ctx, s1, s2, dc, pair_method=None, label=""):
    log(f"dp export {label}: {len(s1)} pairs")
    b1, o1 = workloads.pack(s1)
    b2, o2 = workloads.pack(s2)
    want = O.dp_vectors_packed(b1, o1, b2, o2, oracle_cfg(dc), pair_method=pair_method, threads=8, gaps=True)
    want_s = want[0]
    got = ctx.score_packed_dp(b1, o1, b2, o2, dc, pair_method, gaps=True)
    got_s = got[0]
    assert np.array_equal(got_s.astype(np.int64), want_s), f"{label}: scores"
    # columns, rows, then their gap halves (pa_score_batch_dp_cells); the oracle returns rows first
    _, want_rows, want_cols, want_row_gaps, want_col_gaps = want
    for name, g, w in zip(("column", "row", "column gap", "row gap"), got[2:],
                          (want_cols, want_rows, want_col_gaps, want_row_gaps)):
        bad = np.nonzero(g.astype(np.int64) != w)[0]
        assert len(bad) == 0, f"{label}: {len(bad)} {name} entries differ, first at {bad[:5].tolist()}"
    # the scores-only entry point writes the same score halves
    _, _, cols, rows = ctx.score_packed_dp(b1, o1, b2, o2, dc, pair_method)
    assert np.array_equal(cols, got[2]) and np.array_equal(rows, got[3])
    plain, _ = ctx.score_packed(b1, o1, b2, o2, dc, pair_method)
    assert np.array_equal(plain, got_s)


def dna_batch(rng, count, hi):
    s1, s2 = workloads.dna_pairs(rng, count, 0, hi, divergence=0.15, indel_fraction=0.4)
    # unrelated lengths too, so that both orientations occur
    s2 = [b if k % 3 else bytes(workloads.random_residues(rng, workloads.DNA, int(rng.integers(0, hi))))
          for k, b in enumerate(s2)]
    return s1, s2


@pytest.mark.parametrize("flags", list(itertools.product([False, True], repeat=4)),
                         ids=lambda f: "".join("F" if x else "p" for x in f))
def test_end_gaps(ctx, flags):
    fc, fr, lc, lr = flags
    rng = np.random.default_rng(900 + sum(int(f) << k for k, f in enumerate(flags)))
    s1, s2 = dna_batch(rng, 150, 260)
    dc = device.DeviceConfig(match=4, mismatch=-5, gap_open=-10, gap_extension=-1, first_column_free=fc,
                             first_row_free=fr, last_column_free=lc, last_row_free=lr)
    check(ctx, s1, s2, dc, label=f"ends {flags}")


def test_local_and_mixed_methods(ctx):
    rng = np.random.default_rng(950)
    s1, s2 = dna_batch(rng, 300, 400)
    dc = device.DeviceConfig(method=device.METHOD_LOCAL, match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
    check(ctx, s1, s2, dc, label="local")
    pm = (np.arange(len(s1)) % 2).astype(np.uint8)
    check(ctx, s1, s2, device.DeviceConfig(match=2, mismatch=-3, gap_open=-5, gap_extension=-2), pair_method=pm,
          label="mixed methods")


def test_blosum62(ctx):
    rng = np.random.default_rng(960)
    s1, s2 = workloads.protein_pairs(rng, 200, 0, 300)
    t = tables.blosum62_standard()
    for method in (device.METHOD_GLOBAL, device.METHOD_LOCAL):
        dc = device.DeviceConfig(method=method, score_kind=device.SCORE_MATRIX,
                                 symbols="".join(s for s, _ in t).encode(), matrix=tuple(v for _, r in t for v in r),
                                 gap_open=-11, gap_extension=-1, last_row_free=True)
        check(ctx, s1, s2, dc, label=f"blosum method {method}")


def test_long_and_empty(ctx):
    rng = np.random.default_rng(970)
    s1, s2 = workloads.dna_pairs(rng, 4, 3500, 5000, divergence=0.1)
    s1 += [b"", b"ACGT", b"", b"G"]
    s2 += [b"", b"", b"ACGTA", b"G"]
    for lr in (False, True):
        dc = device.DeviceConfig(match=1, mismatch=-4, gap_open=-6, gap_extension=-1, last_row_free=lr)
        check(ctx, s1, s2, dc, label=f"long and empty, last row free {lr}")


def test_configured_aligner_bulk_and_one_to_many_vectors():
    """The four-argument compute of the configured aligner (cfg mirror of the reference's
    interface_one_to_one_bulk.hpp:73-100 and interface_one_to_many_bulk.hpp:67-77): every bulk
    result carries its own pair's final DP column and row, equal to the oracle's; chains start
    from the method rule and from cfg.initial."""
    from pairwise_aligner_amd import cfg
    rng = np.random.default_rng(77)
    s1, s2 = dna_batch(rng, 60, 90)
    rule = cfg.score_model_unitary(cfg.gap_model_affine(
        cfg.method_global(cfg.leading_end_gap(first_row=cfg.end_gap.free),
                          cfg.trailing_end_gap(last_column=cfg.end_gap.free)), -10, -1), 4, -5)
    a = cfg.configure_aligner(rule)
    dc = a.config
    for firsts, results in ((s1, a.compute(s1, s2, [], [])), ([s1[0]] * len(s2), a.compute(s1[0], s2, [], []))):
        b1, o1 = workloads.pack(firsts)
        b2, o2 = workloads.pack(s2)
        want_s, want_rows, want_cols, want_rg, want_cg = O.dp_vectors_packed(b1, o1, b2, o2, oracle_cfg(dc),
                                                                             threads=8, gaps=True)
        for p, r in enumerate(results):
            assert r.score() == want_s[p]
            c = slice(int(o1[p]) + p, int(o1[p]) + p + len(firsts[p]) + 1)
            w = slice(int(o2[p]) + p, int(o2[p]) + p + len(s2[p]) + 1)
            assert r.dp_column() == list(zip(want_cols[c].tolist(), want_cg[c].tolist())), p
            assert r.dp_row() == list(zip(want_rows[w].tolist(), want_rg[w].tolist())), p
            assert [x.score() for x in r.dp_row()] == want_rows[w].tolist()
    loc = cfg.configure_aligner(cfg.method_local(cfg.gap_model_affine(cfg.score_model_unitary(cfg.initial, 4, -5),
                                                                      -10, -1)))
    one = loc.compute(s1[3], s2[3], [], [])
    want_s, want_rows, want_cols, want_rg, want_cg = O.dp_vectors_packed(
        *workloads.pack([s1[3]]), *workloads.pack([s2[3]]), oracle_cfg(loc.config), gaps=True)
    assert one.score() == want_s[0]
    assert one.dp_row() == list(zip(want_rows.tolist(), want_rg.tolist()))
    assert one.dp_column() == list(zip(want_cols.tolist(), want_cg.tolist()))
