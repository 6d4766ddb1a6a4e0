"""Pins the CPU oracle before anything is compared against it.

* every known-answer vector of the reference's own tests (tests/golden/known_answers.json);
* exhaustive enumeration of alignments on tiny inputs, for every end-gap combination,
  the local method, unitary and BLOSUM62 scoring, and empty sequences;
* invariants that the mathematics fixes (transposition symmetry, closed forms);
* planted mistakes: a deliberately broken copy of the recurrence must fail these checks.
"""
import itertools
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from pairwise_aligner_amd import tables

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")


def oracle_config(c):
    score = c["score"]
    kw = dict(mode=O.LOCAL if c["mode"] == "local" else O.GLOBAL,
              first_column_free=c["first_column_free"], first_row_free=c["first_row_free"],
              last_column_free=c["last_column_free"], last_row_free=c["last_row_free"],
              gap_open=c["gap_open"], gap_extension=c["gap_extension"])
    if score["kind"] == "unitary":
        return O.OracleConfig(score_kind=O.UNITARY, match=score["match"], mismatch=score["mismatch"], **kw)
    table = getattr(tables, score["matrix"])()
    return O.OracleConfig(score_kind=O.MATRIX, symbols="".join(s for s, _ in table).encode(),
                          matrix=tuple(v for _, row in table for v in row), **kw)


def load_golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("vec", load_golden(), ids=lambda v: v["name"])
def test_known_answers(vec):
    cfg = oracle_config(vec["config"])
    assert O.score(vec["seq1"].encode(), vec["seq2"].encode(), cfg) == vec["expected"], vec["source"]


def test_golden_count():
    # 25 scalar vectors of test/unit/affine plus the two configuration smoke tests.
    assert len(load_golden()) == 27


def _random_seq(rng, alphabet, lo, hi):
    n = int(rng.integers(lo, hi + 1))
    return rng.choice(np.frombuffer(alphabet.encode(), dtype=np.uint8), size=n).tobytes() if n else b""


FLAG_COMBOS = list(itertools.product([False, True], repeat=4))


@pytest.mark.parametrize("flags", FLAG_COMBOS, ids=lambda f: "".join("F" if x else "p" for x in f))
def test_brute_force_global_unitary(flags):
    rng = np.random.default_rng(hash(flags) & 0xFFFF)
    fc, fr, lc, lr = flags
    for trial in range(25):
        go = int(rng.integers(-6, 1))
        ge = int(rng.integers(-3, 0))
        cfg = O.OracleConfig(mode=O.GLOBAL, first_column_free=fc, first_row_free=fr, last_column_free=lc,
                             last_row_free=lr, gap_open=go, gap_extension=ge, match=int(rng.integers(1, 5)),
                             mismatch=int(rng.integers(-5, 0)))
        a = _random_seq(rng, "ACG", 0, 5)
        b = _random_seq(rng, "ACG", 0, 5)
        assert O.score(a, b, cfg) == O.brute_force_score(a, b, cfg), (a, b, cfg)


def test_brute_force_local_unitary():
    rng = np.random.default_rng(7)
    for trial in range(60):
        cfg = O.OracleConfig(mode=O.LOCAL, gap_open=int(rng.integers(-6, 0)), gap_extension=int(rng.integers(-3, 0)),
                             match=int(rng.integers(1, 5)), mismatch=int(rng.integers(-5, 0)))
        a = _random_seq(rng, "ACGT", 0, 4)
        b = _random_seq(rng, "ACGT", 0, 4)
        assert O.score(a, b, cfg) == O.brute_force_score(a, b, cfg), (a, b, cfg)


@pytest.mark.parametrize("mode", [O.GLOBAL, O.LOCAL])
def test_brute_force_blosum(mode):
    table = tables.blosum62_standard()
    syms = "".join(s for s, _ in table)
    rng = np.random.default_rng(11 + mode)
    for trial in range(30):
        cfg = O.OracleConfig(mode=mode, gap_open=-11, gap_extension=-1, score_kind=O.MATRIX,
                             symbols=syms.encode(), matrix=tuple(v for _, r in table for v in r),
                             first_row_free=bool(trial % 2), last_column_free=bool(trial % 3 == 0))
        a = _random_seq(rng, syms, 0, 4)
        b = _random_seq(rng, syms, 0, 4)
        assert O.score(a, b, cfg) == O.brute_force_score(a, b, cfg), (a, b)


def test_transposition_symmetry():
    """Swapping seq1/seq2 together with first_row<->first_column and last_row<->last_column
    leaves every score unchanged."""
    rng = np.random.default_rng(3)
    for trial in range(200):
        fc, fr, lc, lr = (bool(x) for x in rng.integers(0, 2, 4))
        kw = dict(gap_open=int(rng.integers(-12, 0)), gap_extension=int(rng.integers(-4, 0)),
                  match=int(rng.integers(1, 6)), mismatch=int(rng.integers(-6, 0)))
        mode = O.LOCAL if trial % 5 == 0 else O.GLOBAL
        c1 = O.OracleConfig(mode=mode, first_column_free=fc, first_row_free=fr, last_column_free=lc, last_row_free=lr, **kw)
        c2 = O.OracleConfig(mode=mode, first_column_free=fr, first_row_free=fc, last_column_free=lr, last_row_free=lc, **kw)
        a = _random_seq(rng, "ACGT", 0, 40)
        b = _random_seq(rng, "ACGT", 0, 40)
        assert O.score(a, b, c1) == O.score(b, a, c2)


def test_closed_forms():
    cfg = O.OracleConfig(gap_open=-10, gap_extension=-1, match=4, mismatch=-5)
    rng = np.random.default_rng(5)
    for n in [0, 1, 2, 17, 300]:
        s = _random_seq(rng, "ACGT", n, n)
        assert O.score(s, s, cfg) == 4 * n                       # identical sequences
        assert O.score(s, b"", cfg) == (0 if n == 0 else -10 - n)  # one empty sequence
        assert O.score(b"", s, cfg) == (0 if n == 0 else -10 - n)
    free = O.OracleConfig(gap_open=-10, gap_extension=-1, first_column_free=True)
    assert O.score(b"ACGT", b"", free) == 0
    loc = O.OracleConfig(mode=O.LOCAL)
    assert O.score(b"", b"ACGT", loc) == O.INT32_MIN and O.score(b"", b"", loc) == O.INT32_MIN
    # no symbol in common, equal length n: k mismatches plus two gaps of n-k is linear in k, so
    # the optimum is at k = 0 or k = n: max(n*mismatch, 2*(go + ge*n))
    for n in [1, 3, 10, 40]:
        assert O.score(b"A" * n, b"T" * n, cfg) == max(-5 * n, 2 * (-10 - n))
    # local score of non-empty input is never negative, and never below any global variant
    for trial in range(50):
        a = _random_seq(rng, "ACGT", 1, 30)
        b = _random_seq(rng, "ACGT", 1, 30)
        lo = O.score(a, b, O.OracleConfig(mode=O.LOCAL))
        assert lo >= 0
        assert lo >= O.score(a, b, O.OracleConfig(mode=O.GLOBAL, first_column_free=True, first_row_free=True,
                                                  last_column_free=True, last_row_free=True))


def test_batch_threads_agree():
    rng = np.random.default_rng(9)
    s1 = [_random_seq(rng, "ACGT", 0, 60) for _ in range(300)]
    s2 = [_random_seq(rng, "ACGT", 0, 60) for _ in range(300)]
    cfg = O.OracleConfig()
    single = O.score_pairs(s1, s2, cfg, threads=1)
    multi = O.score_pairs(s1, s2, cfg, threads=4)
    assert np.array_equal(single, multi)
    assert all(single[i] == O.score(s1[i], s2[i], cfg) for i in range(0, 300, 37))


def test_invalid_matrix_symbol_raises():
    table = tables.blosum62_standard()
    cfg = O.OracleConfig(score_kind=O.MATRIX, symbols="".join(s for s, _ in table).encode(),
                         matrix=tuple(v for _, r in table for v in r))
    with pytest.raises(ValueError):
        O.score(b"ACDZ", b"ACD", cfg)


# ---------------------------------------------------------------------------------------------
# Planted mistakes: a pure-Python copy of the recurrence with one switchable bug each.  The
# checks above (known answers + brute force) must reject every variant, which shows they can
# tell a plausible mistake from the right recurrence.
# ---------------------------------------------------------------------------------------------

def _py_dp(a, b, cfg, bug=None):
    go, ge = cfg.gap_open, cfg.gap_extension
    m, n = len(a), len(b)
    local = cfg.mode == O.LOCAL
    fc = local or cfg.first_column_free
    fr = local or cfg.first_row_free
    NEG = -10 ** 9
    H = [[0] * (n + 1) for _ in range(m + 1)]
    V = [[NEG] * (n + 1) for _ in range(m + 1)]
    W = [[NEG] * (n + 1) for _ in range(m + 1)]
    for i in range(1, m + 1):
        H[i][0] = 0 if fc else go + ge * i
    for j in range(1, n + 1):
        H[0][j] = 0 if fr else go + (ge * j if bug != "init_drops_extension" else 0)
    best = O.INT32_MIN
    for i in range(1, m + 1):
        for j in range(1, n + 1):
            s = cfg.match if a[i - 1] == b[j - 1] else cfg.mismatch
            if bug == "transposed_index":
                s = cfg.match if a[min(j, m) - 1] == b[min(i, n) - 1] else cfg.mismatch
            d = H[i - 1][j - 1] + s
            if local:
                d = max(d, 0)
            ext = ge if bug != "extension_sign" else -ge
            V[i][j] = max(V[i - 1][j] + ext, H[i - 1][j] + go + ge)
            opn = go if bug != "open_without_extension" else 0
            W[i][j] = max(W[i][j - 1] + ge, H[i][j - 1] + opn + ge)
            H[i][j] = max(d, V[i][j], W[i][j])
            best = max(best, H[i][j])
    if local:
        return best
    sc = H[m][n]
    if cfg.last_row_free:
        sc = max([sc] + H[m][(0 if bug != "row_max_skips_first" else 1):])
    if cfg.last_column_free:
        sc = max([sc] + [H[i][n] for i in range(m + 1)])
    return sc


@pytest.mark.parametrize("bug", [None, "init_drops_extension", "transposed_index", "extension_sign",
                                 "open_without_extension", "row_max_skips_first"])
def test_planted_mistakes_are_caught(bug):
    golden = [v for v in load_golden() if v["config"]["score"]["kind"] == "unitary"]
    rng = np.random.default_rng(21)
    caught = False
    for v in golden:
        cfg = oracle_config(v["config"])
        if _py_dp(v["seq1"].encode(), v["seq2"].encode(), cfg, bug) != v["expected"]:
            caught = True
    for trial in range(300):
        fc, fr, lc, lr = (bool(x) for x in rng.integers(0, 2, 4))
        cfg = O.OracleConfig(first_column_free=fc, first_row_free=fr, last_column_free=lc, last_row_free=lr,
                             gap_open=int(rng.integers(-6, 0)), gap_extension=int(rng.integers(-3, 0)),
                             match=2, mismatch=-3)
        a = _random_seq(rng, "ACG", 0, 4)
        b = _random_seq(rng, "ACG", 0, 4)
        if _py_dp(a, b, cfg, bug) != O.brute_force_score(a, b, cfg):
            caught = True
            break
    assert caught == (bug is not None)


# ---------------------------------------------------------------------------------------------
# The CPU baseline of bench.py (oracle/simd_port.cpp) against the scalar oracle.
# ---------------------------------------------------------------------------------------------

def _blosum_cfg(mode):
    from pairwise_aligner_amd import tables
    t = tables.blosum62_standard()
    return O.OracleConfig(mode=mode, gap_open=-11, gap_extension=-1, score_kind=O.MATRIX,
                          symbols="".join(s for s, _ in t).encode(), matrix=tuple(v for _, r in t for v in r))


def test_simd_baseline_matches_oracle_all_end_gaps():
    import itertools
    from pairwise_aligner_amd import workloads
    if O.simd_isa() == 0:
        pytest.skip("no AVX2 / AVX-512 on this CPU")
    rng = np.random.default_rng(71)
    s1, s2 = workloads.dna_pairs(rng, 300, 0, 200, divergence=0.2)
    b1, o1 = workloads.pack(s1)
    b2, o2 = workloads.pack(s2)
    for fl in itertools.product([False, True], repeat=4):
        for mode in (O.GLOBAL, O.LOCAL):
            cfg = O.OracleConfig(mode=mode, first_column_free=fl[0], first_row_free=fl[1], last_column_free=fl[2],
                                 last_row_free=fl[3], gap_open=-10, gap_extension=-1, match=4, mismatch=-5)
            want = O.score_packed(b1, o1, b2, o2, cfg, threads=4)
            got = O.simd_score_packed(b1, o1, b2, o2, cfg, threads=4)
            assert np.array_equal(want, got.astype(np.int64)), (fl, mode)


def test_simd_baseline_matches_oracle_blosum_mixed_methods_and_int32_bulks():
    from pairwise_aligner_amd import workloads
    if O.simd_isa() == 0:
        pytest.skip("no AVX2 / AVX-512 on this CPU")
    rng = np.random.default_rng(73)
    s1, s2 = workloads.protein_pairs(rng, 300, 0, 300)
    b1, o1 = workloads.pack(s1)
    b2, o2 = workloads.pack(s2)
    pm = (np.arange(300) % 3 == 0).astype(np.uint8)
    cfg = _blosum_cfg(O.GLOBAL)
    want = O.score_packed(b1, o1, b2, o2, cfg, pair_method=pm, threads=4)
    got = O.simd_score_packed(b1, o1, b2, o2, cfg, pair_method=pm, threads=4)
    assert np.array_equal(want, got.astype(np.int64))
    # bulks whose bound needs int32 lanes (scores beyond int16)
    s1, s2 = workloads.dna_pairs(rng, 20, 9000, 10000, divergence=0.02)
    b1, o1 = workloads.pack(s1)
    b2, o2 = workloads.pack(s2)
    cfg = O.OracleConfig(mode=O.GLOBAL, gap_open=-10, gap_extension=-1, match=4, mismatch=-5)
    want = O.score_packed(b1, o1, b2, o2, cfg, threads=4)
    got = O.simd_score_packed(b1, o1, b2, o2, cfg, threads=4)
    assert np.array_equal(want, got.astype(np.int64))
    assert want.max() > 32767


# ---- the final DP row and column (oracle.dp_vectors_packed) ------------------------------------
# H(i, j) of the global method is the best score of aligning s1[:i] with s2[:j] under the leading
# end-gap rules and penalised trailing gaps, so each exported entry is pinned by enumeration.  For
# the local method H(i, j) = max(0, best fully penalised alignment of s1[k:i] with s2[l:j] over all
# k <= i, l <= j): every cell can start an alignment at value 0 (the zero clamp and the free
# leading gaps).

def _prefix_global(a, b, cfg):
    pen = O.OracleConfig(mode=O.GLOBAL, first_column_free=cfg.first_column_free,
                         first_row_free=cfg.first_row_free, last_column_free=False, last_row_free=False,
                         gap_open=cfg.gap_open, gap_extension=cfg.gap_extension, match=cfg.match,
                         mismatch=cfg.mismatch)
    return O.brute_force_score(a, b, pen)


def _cell_local(a, b, cfg):
    pen = O.OracleConfig(mode=O.GLOBAL, gap_open=cfg.gap_open, gap_extension=cfg.gap_extension,
                         match=cfg.match, mismatch=cfg.mismatch)
    best = 0
    for k in range(len(a) + 1):
        for l in range(len(b) + 1):
            if k < len(a) or l < len(b):
                best = max(best, O.brute_force_score(a[k:], b[l:], pen))
    return best


def _dp_vectors(a, b, cfg):
    b1, o1 = O.pack([a])
    b2, o2 = O.pack([b])
    s, rows, cols = O.dp_vectors_packed(b1, o1, b2, o2, cfg)
    return int(s[0]), rows.tolist(), cols.tolist()


@pytest.mark.parametrize("flags", FLAG_COMBOS, ids=lambda f: "".join("F" if x else "p" for x in f))
def test_dp_vectors_global_by_enumeration(flags):
    rng = np.random.default_rng(1000 + (hash(flags) & 0xFFF))
    fc, fr, lc, lr = flags
    for trial in range(8):
        cfg = O.OracleConfig(mode=O.GLOBAL, first_column_free=fc, first_row_free=fr, last_column_free=lc,
                             last_row_free=lr, gap_open=int(rng.integers(-6, 1)), gap_extension=int(rng.integers(-3, 0)),
                             match=int(rng.integers(1, 5)), mismatch=int(rng.integers(-5, 0)))
        a = _random_seq(rng, "ACG", 0, 4)
        b = _random_seq(rng, "ACG", 0, 4)
        score, rows, cols = _dp_vectors(a, b, cfg)
        assert rows == [_prefix_global(a, b[:j], cfg) for j in range(len(b) + 1)], (a, b, cfg)
        assert cols == [_prefix_global(a[:i], b, cfg) for i in range(len(a) + 1)], (a, b, cfg)
        # the tracker's reading of the two vectors (tracker_global_scalar.hpp:36-61)
        want = rows[-1]
        if lr:
            want = max([want] + rows)
        if lc:
            want = max([want] + cols)
        assert score == want


def test_dp_vectors_local_by_enumeration():
    rng = np.random.default_rng(77)
    for trial in range(12):
        cfg = O.OracleConfig(mode=O.LOCAL, gap_open=int(rng.integers(-6, 0)), gap_extension=int(rng.integers(-3, 0)),
                             match=int(rng.integers(1, 5)), mismatch=int(rng.integers(-5, 0)))
        a = _random_seq(rng, "ACGT", 0, 4)
        b = _random_seq(rng, "ACGT", 0, 4)
        _, rows, cols = _dp_vectors(a, b, cfg)
        assert rows == [_cell_local(a, b[:j], cfg) for j in range(len(b) + 1)], (a, b, cfg)
        assert cols == [_cell_local(a[:i], b, cfg) for i in range(len(a) + 1)], (a, b, cfg)


# The gap-state halves of the same cells.  Unrolling V(i+1, j) = max(V(i, j) + ge, H(i, j) + go + ge)
# from V(1, j) = H(0, j) + go + ge gives the closed form
#   V(m+1, j) = max over i = 0..m of H(i, j) + go + ge * (m + 1 - i)
# (the best vertical gap into row m+1, over the row it opens below), and the same along a row for
# W(i, n+1); entry 0 of each vector keeps its initial go + ge.  H(i, j) comes from enumeration, so
# the gap vectors are pinned by a formula that does not restate the recurrence.

def _gap_vectors_closed_form(a, b, cfg, cell):
    go, ge = cfg.gap_open, cfg.gap_extension
    m, n = len(a), len(b)
    H = [[cell(a[:i], b[:j], cfg) for j in range(n + 1)] for i in range(m + 1)]
    row_gaps = [go + ge] + [max(H[i][j] + go + ge * (m + 1 - i) for i in range(m + 1)) for j in range(1, n + 1)]
    col_gaps = [go + ge] + [max(H[i][j] + go + ge * (n + 1 - j) for j in range(n + 1)) for i in range(1, m + 1)]
    return row_gaps, col_gaps


def _gap_vectors(a, b, cfg):
    b1, o1 = O.pack([a])
    b2, o2 = O.pack([b])
    _, _, _, rg, cg = O.dp_vectors_packed(b1, o1, b2, o2, cfg, gaps=True)
    return rg.tolist(), cg.tolist()


@pytest.mark.parametrize("flags", FLAG_COMBOS, ids=lambda f: "".join("F" if x else "p" for x in f))
def test_dp_gap_vectors_global_by_closed_form(flags):
    rng = np.random.default_rng(2000 + (hash(flags) & 0xFFF))
    fc, fr, lc, lr = flags
    for trial in range(6):
        cfg = O.OracleConfig(mode=O.GLOBAL, first_column_free=fc, first_row_free=fr, last_column_free=lc,
                             last_row_free=lr, gap_open=int(rng.integers(-6, 1)), gap_extension=int(rng.integers(-3, 0)),
                             match=int(rng.integers(1, 5)), mismatch=int(rng.integers(-5, 0)))
        a = _random_seq(rng, "ACG", 0, 4)
        b = _random_seq(rng, "ACG", 0, 4)
        assert _gap_vectors(a, b, cfg) == _gap_vectors_closed_form(a, b, cfg, _prefix_global), (a, b, cfg)


def test_dp_gap_vectors_local_by_closed_form():
    rng = np.random.default_rng(79)
    for trial in range(10):
        cfg = O.OracleConfig(mode=O.LOCAL, gap_open=int(rng.integers(-6, 0)), gap_extension=int(rng.integers(-3, 0)),
                             match=int(rng.integers(1, 5)), mismatch=int(rng.integers(-5, 0)))
        a = _random_seq(rng, "ACGT", 0, 4)
        b = _random_seq(rng, "ACGT", 0, 4)
        assert _gap_vectors(a, b, cfg) == _gap_vectors_closed_form(a, b, cfg, _cell_local), (a, b, cfg)


@pytest.mark.parametrize("bug", ["state_before_last_row", "extension_dropped", "entry0_updated"])
def test_dp_gap_vectors_planted_mistakes_are_caught(bug):
    """Plausible mistakes in the gap halves (exporting the state one row early, dropping the
    extension, updating entry 0 like the others) must fail the closed form."""
    rng = np.random.default_rng(80)
    caught = False
    for trial in range(60):
        cfg = O.OracleConfig(mode=O.GLOBAL, first_column_free=bool(rng.integers(0, 2)),
                             first_row_free=bool(rng.integers(0, 2)), gap_open=int(rng.integers(-6, 0)),
                             gap_extension=int(rng.integers(-3, 0)), match=2, mismatch=-3)
        a = _random_seq(rng, "ACG", 1, 4)
        b = _random_seq(rng, "ACG", 1, 4)
        rg, cg = _gap_vectors(a, b, cfg)
        if bug == "state_before_last_row":
            rg = _gap_vectors(a[:-1], b, cfg)[0]
        elif bug == "extension_dropped":
            rg = [v - cfg.gap_extension if j else v for j, v in enumerate(rg)]
        else:
            h0 = 0 if cfg.first_row_free else cfg.gap_open + cfg.gap_extension * len(b)
            cg = [h0 + cfg.gap_open + cfg.gap_extension] + cg[1:]
        if (rg, cg) != _gap_vectors_closed_form(a, b, cfg, _prefix_global):
            caught = True
            break
    assert caught


def test_dp_vectors_transpose_and_batch_layout():
    rng = np.random.default_rng(78)
    seqs1 = [_random_seq(rng, "ACGT", 0, 30) for _ in range(40)]
    seqs2 = [_random_seq(rng, "ACGT", 0, 30) for _ in range(40)]
    cfg = O.OracleConfig(mode=O.GLOBAL, gap_open=-5, gap_extension=-2, match=2, mismatch=-3)
    b1, o1 = O.pack(seqs1)
    b2, o2 = O.pack(seqs2)
    s, rows, cols, rg, cg = O.dp_vectors_packed(b1, o1, b2, o2, cfg, threads=4, gaps=True)
    st, rows_t, cols_t, rg_t, cg_t = O.dp_vectors_packed(b2, o2, b1, o1, cfg, threads=4, gaps=True)
    assert np.array_equal(s, st) and np.array_equal(rows, cols_t) and np.array_equal(cols, rows_t)
    assert np.array_equal(rg, cg_t) and np.array_equal(cg, rg_t)
    # the batch writes each pair's gap halves where the pair alone puts them
    for p in (0, 7, 39):
        one_rg, one_cg = _gap_vectors(seqs1[p], seqs2[p], cfg)
        r0, c0 = int(o2[p]) + p, int(o1[p]) + p
        assert rg[r0:r0 + len(seqs2[p]) + 1].tolist() == one_rg and cg[c0:c0 + len(seqs1[p]) + 1].tolist() == one_cg
    for p in range(40):
        r0 = int(o2[p]) + p
        c0 = int(o1[p]) + p
        assert rows[r0 + len(seqs2[p])] == s[p] == cols[c0 + len(seqs1[p])]
