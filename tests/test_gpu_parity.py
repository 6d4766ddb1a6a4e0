"""GPU parity: the device library against the CPU oracle, bit-exact, through the C ABI.

Mirrors the reference's own test strategy: the known-answer vectors of
test/unit/affine/*_scalar*_test.cpp through the configuration API, and seeded differential
tests (alignment_simd_test_template.hpp:195-227) over every method, end-gap combination, score
model and lane width, with equal and variable lengths, empty sequences and saturating pairs.
"""
import dataclasses
import itertools
import json
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from pairwise_aligner_amd import cfg, device, tables, workloads

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")


def log(msg):
    sys.stdout.write(msg + "\n")
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    return device.Context(0)


def oracle_cfg(dc: device.DeviceConfig, method=None) -> O.OracleConfig:
    return O.OracleConfig(mode=dc.method if method is None else method, first_column_free=dc.first_column_free,
                          first_row_free=dc.first_row_free, last_column_free=dc.last_column_free,
                          last_row_free=dc.last_row_free, gap_open=dc.gap_open, gap_extension=dc.gap_extension,
                          score_kind=dc.score_kind, match=dc.match, mismatch=dc.mismatch, symbols=dc.symbols,
                          matrix=dc.matrix)


def check(ctx, s1, s2, dc, pair_method=None, label=""):
    log(f"case {label}: {len(s1)} pairs, cfg={dc}")
    got, flags = ctx.score(s1, s2, dc, pair_method)
    if pair_method is None:
        want = O.score_pairs(s1, s2, oracle_cfg(dc), threads=8)
    else:
        want = O.score_pairs(s1, s2, oracle_cfg(dc), pair_method=pair_method, threads=8)
    bad = np.nonzero(got.astype(np.int64) != want)[0]
    assert len(bad) == 0, (f"{label}: {len(bad)} mismatches; first {bad[:5].tolist()}: device "
                           f"{got[bad[:5]].tolist()} oracle {want[bad[:5]].tolist()} "
                           f"lens {[(len(s1[i]), len(s2[i])) for i in bad[:5]]}")
    return got, flags


def test_known_answers_through_cfg_api():
    vecs = json.load(open(GOLDEN))
    for v in vecs:
        c = v["config"]
        lead = cfg.leading_end_gap(cfg.end_gap(c["first_column_free"]), cfg.end_gap(c["first_row_free"]))
        trail = cfg.trailing_end_gap(cfg.end_gap(c["last_column_free"]), cfg.end_gap(c["last_row_free"]))
        gap = cfg.gap_model_affine(c["gap_open"], c["gap_extension"])
        method = cfg.method_local(gap) if c["mode"] == "local" else cfg.method_global(gap, lead, trail)
        sc = c["score"]
        rule = (cfg.score_model_unitary(method, sc["match"], sc["mismatch"]) if sc["kind"] == "unitary"
                else cfg.score_model_matrix(method, getattr(tables, sc["matrix"])()))
        aligner = cfg.configure_aligner(rule)
        log(f"known answer {v['name']}")
        assert aligner.compute(v["seq1"], v["seq2"]).score() == v["expected"], v["source"]


@pytest.mark.parametrize("method", [device.METHOD_GLOBAL, device.METHOD_LOCAL])
@pytest.mark.parametrize("lengths", [(1, 20), (140, 160), (10, 300), (300, 700)])
def test_dna_unitary(ctx, method, lengths):
    rng = np.random.default_rng(100 + lengths[0] + method)
    s1, s2 = workloads.dna_pairs(rng, 2000, lengths[0], lengths[1], divergence=0.15)
    dc = device.DeviceConfig(method=method, match=4, mismatch=-5, gap_open=-10, gap_extension=-1)
    check(ctx, s1, s2, dc, label=f"dna method={method} lengths={lengths}")


@pytest.mark.parametrize("flags", list(itertools.product([False, True], repeat=4)),
                         ids=lambda f: "".join("F" if x else "p" for x in f))
def test_end_gaps(ctx, flags):
    fc, fr, lc, lr = flags
    rng = np.random.default_rng(7)
    s1, s2 = workloads.dna_pairs(rng, 600, 0, 90, divergence=0.3)
    # unrelated pairs too, so that free ends matter
    u1, u2 = workloads.dna_pairs(rng, 300, 0, 90, divergence=0.0)
    s2 = s2[:300] + u2[::-1]
    dc = device.DeviceConfig(first_column_free=fc, first_row_free=fr, last_column_free=lc, last_row_free=lr,
                             match=4, mismatch=-5, gap_open=-10, gap_extension=-1)
    check(ctx, s1, s2, dc, label=f"end gaps {flags}")


@pytest.mark.parametrize("policy", [device.LANES_AUTO, device.LANES_INT16, device.LANES_INT32])
def test_blosum62_end_gaps_and_lane_policies(ctx, policy):
    """Every end-gap combination with the matrix score (pair-row table in int16, per-entry table
    in int32), including the per-row maxima of a free last row and the tracked running maximum
    of speculative int16 pairs; rows of every parity so that pair rows end in a pad row."""
    rng = np.random.default_rng(67 + policy)
    s1, s2 = workloads.protein_pairs(rng, 400, 0, 330)
    t = tables.blosum62_standard()
    for flags in itertools.product([False, True], repeat=4):
        fc, fr, lc, lr = flags
        dc = device.DeviceConfig(first_column_free=fc, first_row_free=fr, last_column_free=lc, last_row_free=lr,
                                 score_kind=device.SCORE_MATRIX, symbols="".join(x for x, _ in t).encode(),
                                 matrix=tuple(v for _, r in t for v in r), gap_open=-11, gap_extension=-1,
                                 lane_policy=policy)
        check(ctx, s1, s2, dc, label=f"blosum62 end gaps {flags} policy {policy}")


@pytest.mark.parametrize("method", [device.METHOD_GLOBAL, device.METHOD_LOCAL])
def test_blosum62(ctx, method):
    rng = np.random.default_rng(55 + method)
    s1, s2 = workloads.protein_pairs(rng, 1500, 0, 420)
    t = tables.blosum62_standard()
    dc = device.DeviceConfig(method=method, score_kind=device.SCORE_MATRIX, symbols="".join(s for s, _ in t).encode(),
                             matrix=tuple(v for _, r in t for v in r), gap_open=-11, gap_extension=-1)
    check(ctx, s1, s2, dc, label=f"blosum62 method={method}")


@pytest.mark.parametrize("method", [device.METHOD_GLOBAL, device.METHOD_LOCAL])
def test_blosum62_extended_large_alphabet(ctx, method):
    """27 symbols + pad: too many for the pair-row LDS table, so int16 tasks take the replicated
    per-entry table; odd and even lengths, residues from the whole extended alphabet."""
    rng = np.random.default_rng(61 + method)
    t = tables.blosum62_extended()
    syms = np.frombuffer("".join(s for s, _ in t).encode(), dtype=np.uint8)
    lens1 = rng.integers(0, 400, size=600)
    lens2 = rng.integers(0, 400, size=600)
    s1 = [syms[rng.integers(0, len(syms), size=n)].tobytes() for n in lens1]
    s2 = [syms[rng.integers(0, len(syms), size=n)].tobytes() for n in lens2]
    dc = device.DeviceConfig(method=method, score_kind=device.SCORE_MATRIX, symbols=syms.tobytes(),
                             matrix=tuple(v for _, r in t for v in r), gap_open=-11, gap_extension=-1)
    check(ctx, s1, s2, dc, label=f"blosum62 extended method={method}")


def test_unitary_many_symbols(ctx):
    """More than four distinct symbols (e.g. N): unitary equality over an LDS table."""
    rng = np.random.default_rng(9)
    s1, s2 = workloads.dna_pairs(rng, 800, 0, 200, divergence=0.2)
    s1 = [s.replace(b"A", b"N", 3) for s in s1]
    for method in (device.METHOD_GLOBAL, device.METHOD_LOCAL):
        dc = device.DeviceConfig(method=method, match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
        check(ctx, s1, s2, dc, label=f"unitary with N method={method}")


@pytest.mark.parametrize("policy", [device.LANES_INT32, device.LANES_INT16, device.LANES_AUTO])
def test_lane_policies(ctx, policy):
    rng = np.random.default_rng(13)
    s1, s2 = workloads.dna_pairs(rng, 1000, 50, 400, divergence=0.1)
    pm = (np.arange(1000) % 2).astype(np.uint8)
    dc = device.DeviceConfig(match=2, mismatch=-3, gap_open=-5, gap_extension=-2, lane_policy=policy)
    got, flags = check(ctx, s1, s2, dc, pair_method=pm, label=f"policy {policy}")
    if policy == device.LANES_INT32:
        assert np.all(flags & device.PAIR_INT32)


@pytest.mark.parametrize("length, match", [(9000, 4), (2500, 15)], ids=["wavefront", "striped"])
def test_saturation_recompute(ctx, length, match):
    """Scores beyond int16: identical sequences (9 kbp at +4, 36,000, in the wavefront kernel;
    2.5 kbp at +15, 37,500, in the striped kernel) in int16 lanes must be flagged, recomputed in
    int32 and come out exact, for the global method (saturation tracking) and the local one
    (running maximum in the column drift's frames)."""
    rng = np.random.default_rng(21)
    base = workloads.unpack(*workloads.random_packed(rng, np.full(6, length), workloads.DNA))
    short1, short2 = workloads.dna_pairs(rng, 40, 100, 200, divergence=0.05)
    s1 = base + short1
    s2 = list(base) + short2
    for method in (device.METHOD_GLOBAL, device.METHOD_LOCAL):
        dc = device.DeviceConfig(method=method, match=match, mismatch=-5, gap_open=-10, gap_extension=-1,
                                 lane_policy=device.LANES_INT16)
        got, flags = check(ctx, s1, s2, dc, label=f"saturation method={method} length={length}")
        assert np.all(flags[:6] & device.PAIR_RECOMPUTED), flags[:6]
        assert np.all(got[:6] == length * match)
        assert not np.any(flags[6:] & device.PAIR_RECOMPUTED)


INT16_TOP = 31744   # the largest value the int16 three-way-maximum lanes hold exactly


@pytest.mark.parametrize("kernel", ["striped", "wavefront"])
def test_int16_top_local(ctx, kernel):
    """Local scores on both sides of the int16 lanes' top value, 31,744 (pa_plan.hpp, kInt16Top),
    in the striped and in the wavefront kernel: identical DNA sequences at +4 (31,000 to 32,760)
    and identical protein sequences under BLOSUM62.  Under int16 speculation every score is exact,
    every pair scoring above the top is flagged and recomputed in int32 (such scores were exact
    in int16 lanes when the top was 32,767), and pairs far below it are not; under the length
    bound nothing is recomputed and the pairs above the top run in int32."""
    rng = np.random.default_rng(31744)
    lengths = [7750, 7900, 7936, 7950, 8100, 8190]
    dna = workloads.unpack(*workloads.random_packed(rng, np.array(lengths), workloads.DNA))
    short1, short2 = workloads.dna_pairs(rng, 30, 100, 400, divergence=0.1)
    tuning = device.Tuning(wavefront_min_rows=1 << 20) if kernel == "striped" else device.Tuning(wavefront_tail=-1.0)
    t = tables.blosum62_standard()
    prot = workloads.unpack(*workloads.random_packed(rng, np.array([5600, 6000, 6150, 6300]), workloads.AMINO,
                                                     workloads.AMINO_FREQ))
    cases = [("dna", device.DeviceConfig(method=device.METHOD_LOCAL, match=4, mismatch=-5, gap_open=-10,
                                         gap_extension=-1), dna + short1, dna + short2),
             ("protein", device.DeviceConfig(method=device.METHOD_LOCAL, score_kind=device.SCORE_MATRIX,
                                             symbols="".join(s for s, _ in t).encode(),
                                             matrix=tuple(v for _, r in t for v in r), gap_open=-11,
                                             gap_extension=-1), prot, list(prot))]
    for name, base, s1, s2 in cases:
        for policy in (device.LANES_INT16, device.LANES_AUTO):
            dc = dataclasses.replace(base, lane_policy=policy, tuning=tuning)
            got, flags = check(ctx, s1, s2, dc, label=f"int16 top {name} {kernel} policy={policy}")
            long_ = np.arange(len(s1)) < (len(dna) if name == "dna" else len(prot))
            log(f"  scores {got[long_].tolist()} flags {flags[long_].tolist()}")
            if policy == device.LANES_INT16:
                assert np.all(flags[got > INT16_TOP] & device.PAIR_RECOMPUTED), flags
                assert not np.any(flags[got < 30000] & device.PAIR_RECOMPUTED), flags
            else:   # the bound policy proves int16 or runs int32: nothing to recompute
                assert not np.any(flags & device.PAIR_RECOMPUTED), flags
                assert np.all(flags[got > INT16_TOP] & device.PAIR_INT32), flags
        assert np.any(got > INT16_TOP) and np.any((got < INT16_TOP) & long_)


def test_short_pair_alone_int16(ctx):
    """A batch of one short pair is a task of fewer columns (N) than the 16-step first chunk of the
    striped kernel: its lanes pass column N-1 inside that chunk and sweep pad columns from there
    on, so that no lane's state falls behind the column drift -- a stale state would sink below
    the range of the int16 lanes' three-way maximum, and an identical 3-residue pair then scored 0.
    Identical and unrelated pairs of 1-20 residues, each alone in its batch, local and global,
    DNA and BLOSUM62, int16 lanes."""
    rng = np.random.default_rng(316)
    t = tables.blosum62_standard()
    blosum = dict(score_kind=device.SCORE_MATRIX, symbols="".join(s for s, _ in t).encode(),
                  matrix=tuple(v for _, r in t for v in r), gap_open=-11, gap_extension=-1)
    for method in (device.METHOD_LOCAL, device.METHOD_GLOBAL):
        for alpha, extra in ((workloads.DNA, dict(match=2, mismatch=-3, gap_open=-5, gap_extension=-2)),
                             (workloads.AMINO, blosum)):
            dc = device.DeviceConfig(method=method, lane_policy=device.LANES_INT16, **extra)
            got_all, want_all = [], []
            for n in range(1, 21):
                a = bytes(workloads.random_residues(rng, alpha, n))
                b = bytes(workloads.random_residues(rng, alpha, n))
                for s1, s2 in (([a], [a]), ([a], [b])):
                    got, flags = ctx.score(s1, s2, dc)
                    assert flags[0] & device.PAIR_INT16, flags
                    got_all.append(int(got[0]))
                    want_all.append(int(O.score_pairs(s1, s2, oracle_cfg(dc))[0]))
            log(f"  short pairs alone method={method} alphabet={len(alpha)}: {got_all}")
            assert got_all == want_all, (method, len(alpha), got_all, want_all)


def test_int16_bias_window_global(ctx):
    """Global pairs on both sides of the int16 window of the three-way-maximum lanes (their lower
    bound lifted to 0 by a per-pair bias, their upper bound at most 31,744): with +4/-5 and gap
    (-10, -1), m = n up to about 5,270 fit (biased lanes), 5,300 and 5,500 do not (absolute int16
    lanes with the running maximum).  Identical sequences score near the top, unrelated ones near
    the bottom of the window; every score is exact and none is recomputed, in the striped kernel,
    for every end-gap combination."""
    rng = np.random.default_rng(5270)
    lengths = [4000, 5000, 5200, 5300, 5500]
    a = workloads.unpack(*workloads.random_packed(rng, np.array(lengths * 2), workloads.DNA))
    b = workloads.unpack(*workloads.random_packed(rng, np.array(lengths), workloads.DNA))
    s1 = a
    s2 = a[:len(lengths)] + b            # identical pairs, then unrelated pairs of the same lengths
    tuning = device.Tuning(wavefront_min_rows=1 << 20)
    for fc, fr, lc, lr in itertools.product([False, True], repeat=4):
        if (fc, fr, lc, lr) not in ((False, False, False, False), (True, True, True, True), (True, False, False, True),
                                    (False, True, True, False)):
            continue
        dc = device.DeviceConfig(method=device.METHOD_GLOBAL, match=4, mismatch=-5, gap_open=-10, gap_extension=-1,
                                 first_column_free=fc, first_row_free=fr, last_column_free=lc, last_row_free=lr,
                                 lane_policy=device.LANES_AUTO, tuning=tuning)
        got, flags = check(ctx, s1, s2, dc, label=f"global bias window ends={(fc, fr, lc, lr)}")
        assert np.all(flags & device.PAIR_INT16) and not np.any(flags & device.PAIR_RECOMPUTED), flags


@pytest.mark.parametrize("ends", [(False, False), (True, False), (False, True), (True, True)],
                         ids=["penalised", "last-column-free", "last-row-free", "both-free"])
def test_tracked_int16_global_scores(ctx, ends):
    """int16 global pairs whose bound does not rule out saturation (tracked for the int32
    recompute) but that stay below it: the score is H(m, n) (or the free-end maximum), never the
    maximum of the last row that the tracking keeps for its saturation test."""
    lc, lr = ends
    rng = np.random.default_rng(22)
    s1, s2 = workloads.dna_pairs(rng, 120, 1100, 2200, divergence=0.4, indel_fraction=0.3)
    dc = device.DeviceConfig(match=30, mismatch=-5, gap_open=-10, gap_extension=-1, last_column_free=lc,
                             last_row_free=lr, lane_policy=device.LANES_INT16)
    got, flags = check(ctx, s1, s2, dc, label=f"tracked int16 ends={ends}")
    stayed = ((flags & device.PAIR_INT16) != 0) & ((flags & device.PAIR_RECOMPUTED) == 0)
    assert np.count_nonzero(stayed) > 30   # scored by the tracked int16 kernels


@pytest.mark.parametrize("ge", [-600, -3000])
def test_large_gap_extension_runs_int32(ctx, ge):
    """A gap extension so large that the int16 kernels' drift frames (|ge| per row and per
    column step) would leave int16 by themselves: the planner runs every pair in int32, also under
    int16 speculation, and the scores equal the oracle's (short pairs and one pair longer than
    3,072 rows, for the wavefront kernel)."""
    rng = np.random.default_rng(90 - ge)
    s1, s2 = workloads.dna_pairs(rng, 80, 20, 300, divergence=0.2)
    l1, l2 = workloads.dna_pairs(rng, 1, 3500, 3600, divergence=0.1)
    s1, s2 = s1 + l1, s2 + l2
    for method in (device.METHOD_GLOBAL, device.METHOD_LOCAL):
        for policy in (device.LANES_INT16, device.LANES_AUTO):
            dc = device.DeviceConfig(method=method, match=4, mismatch=-5, gap_open=-10, gap_extension=ge,
                                     lane_policy=policy)
            got, flags = check(ctx, s1, s2, dc, label=f"ge={ge} method={method} policy={policy}")
            assert np.all(flags & device.PAIR_INT32), flags


def test_empty_and_tiny(ctx):
    s1 = [b"", b"A", b"", b"ACGT", b"G", b"", b"AC"]
    s2 = [b"", b"", b"A", b"", b"G", b"ACGTT", b"CA"]
    for flags in itertools.product([False, True], repeat=4):
        fc, fr, lc, lr = flags
        for method in (device.METHOD_GLOBAL, device.METHOD_LOCAL):
            dc = device.DeviceConfig(method=method, first_column_free=fc, first_row_free=fr, last_column_free=lc,
                                     last_row_free=lr)
            check(ctx, s1, s2, dc, label=f"empty {flags} {method}")
    # a batch without pairs launches nothing and reports no error, however often it is made
    for _ in range(4):
        got, _ = ctx.score([], [], device.DeviceConfig())
        assert len(got) == 0


def test_invalid_symbol_is_an_error(ctx):
    t = tables.blosum62_standard()
    dc = device.DeviceConfig(score_kind=device.SCORE_MATRIX, symbols="".join(s for s, _ in t).encode(),
                             matrix=tuple(v for _, r in t for v in r))
    with pytest.raises(device.PaError) as e:
        ctx.score([b"ACDZ"], [b"ACD"], dc)
    assert e.value.status == 2


def test_positive_gap_penalties_are_unsupported(ctx):
    for go, ge in ((1, -2), (0, 1)):
        dc = device.DeviceConfig(gap_open=go, gap_extension=ge)
        with pytest.raises(device.PaError) as e:
            ctx.score([b"ACGT"], [b"ACG"], dc)
        assert e.value.status == 3


def test_batch_repeat_runs(ctx):
    rng = np.random.default_rng(31)
    s1, s2 = workloads.dna_pairs(rng, 3000, 100, 200)
    b1, o1 = workloads.pack(s1)
    b2, o2 = workloads.pack(s2)
    dc = device.DeviceConfig()
    bt = device.Batch(ctx, b1, o1, b2, o2, dc)
    bt.run()
    first, _ = bt.results()
    bt.run()
    second, _ = bt.results()
    assert np.array_equal(first, second)
    want = O.score_pairs(s1, s2, oracle_cfg(dc), threads=8)
    assert np.array_equal(first.astype(np.int64), want)
    st = bt.stats()
    assert st["n_pairs"] == 3000 and st["cells"] == sum(len(a) * len(b) for a, b in zip(s1, s2))


@pytest.mark.parametrize("policy", [device.LANES_INT32, device.LANES_AUTO])
@pytest.mark.parametrize("method", [device.METHOD_GLOBAL, device.METHOD_LOCAL])
def test_long_pairs_wavefront(ctx, policy, method):
    """Pairs longer than 3072 rows go to the wavefront kernel (8 waves per task, LDS rings
    between waves, HBM between rounds); 9,000-12,000 rows make 9-12 stripes of 1,024 rows,
    i.e. two rounds.  Short pairs in the same batch go to the striped kernel."""
    rng = np.random.default_rng(41 + policy + method)
    l1, l2 = workloads.dna_pairs(rng, 6, 3100, 12000, divergence=0.15, indel_fraction=0.6, geom_p=0.5)
    s1_, s2_ = workloads.dna_pairs(rng, 200, 50, 400, divergence=0.1)
    s1 = l1 + s1_
    s2 = l2 + s2_
    dc = device.DeviceConfig(method=method, match=1, mismatch=-4, gap_open=-6, gap_extension=-1, lane_policy=policy)
    check(ctx, s1, s2, dc, label=f"long policy={policy} method={method}")


@pytest.mark.parametrize("policy", [device.LANES_INT32, device.LANES_AUTO])
@pytest.mark.parametrize("method", [device.METHOD_GLOBAL, device.METHOD_LOCAL])
def test_long_pairs_striped(ctx, policy, method):
    """The same kind of pairs with the wavefront kernel's tail rule set so high that no long pair
    counts as the batch's tail: every pair runs in the striped kernel, the longest (12,000 rows)
    as 47 stripes of 256 rows with the boundary row passed through HBM between stripes."""
    rng = np.random.default_rng(43 + policy + method)
    s1, s2 = workloads.dna_pairs(rng, 6, 3100, 12000, divergence=0.15, indel_fraction=0.6, geom_p=0.5)
    s1_, s2_ = workloads.dna_pairs(rng, 100, 50, 400, divergence=0.1)
    dc = device.DeviceConfig(method=method, match=1, mismatch=-4, gap_open=-6, gap_extension=-1, lane_policy=policy,
                             tuning=device.Tuning(wavefront_tail=1e9))
    check(ctx, s1 + s1_, s2 + s2_, dc, label=f"long striped policy={policy} method={method}")


def test_long_protein_local_and_free_ends_striped(ctx):
    rng = np.random.default_rng(92)
    s1, s2 = workloads.protein_pairs(rng, 4, 3200, 6000)
    t = tables.blosum62_standard()
    never = device.Tuning(wavefront_tail=1e9)
    dc = device.DeviceConfig(method=device.METHOD_LOCAL, score_kind=device.SCORE_MATRIX,
                             symbols="".join(s for s, _ in t).encode(), matrix=tuple(v for _, r in t for v in r),
                             gap_open=-11, gap_extension=-1, tuning=never)
    check(ctx, s1, s2, dc, label="long protein local striped")
    d1, d2 = workloads.dna_pairs(rng, 4, 3500, 9000, divergence=0.3)
    d2 = [d2[0], d2[1][: len(d2[1]) // 2], d2[2], d2[3]]
    dc = device.DeviceConfig(first_column_free=True, first_row_free=False, last_column_free=True, last_row_free=False,
                             match=2, mismatch=-3, gap_open=-5, gap_extension=-2, tuning=never)
    check(ctx, d1, d2, dc, label="long free ends striped")


@pytest.mark.parametrize("flags", [(True, False, True, False), (False, True, False, True), (True, True, True, True)])
def test_long_pairs_free_ends(ctx, flags):
    fc, fr, lc, lr = flags
    rng = np.random.default_rng(77)
    s1, s2 = workloads.dna_pairs(rng, 4, 3500, 9000, divergence=0.3)
    s2 = [s2[0], s2[1][: len(s2[1]) // 2], s2[2], s2[3]]
    dc = device.DeviceConfig(first_column_free=fc, first_row_free=fr, last_column_free=lc, last_row_free=lr,
                             match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
    check(ctx, s1, s2, dc, label=f"long free ends {flags}")


def test_long_protein_local(ctx):
    rng = np.random.default_rng(91)
    s1, s2 = workloads.protein_pairs(rng, 4, 3200, 6000)
    t = tables.blosum62_standard()
    dc = device.DeviceConfig(method=device.METHOD_LOCAL, score_kind=device.SCORE_MATRIX,
                             symbols="".join(s for s, _ in t).encode(), matrix=tuple(v for _, r in t for v in r),
                             gap_open=-11, gap_extension=-1)
    check(ctx, s1, s2, dc, label="long protein local")


def test_large_one_shot_batch(ctx):
    """A one-shot call on 528,609 pairs: scores equal the oracle's and flags equal those of the
    same pairs run as a device-resident batch, with per-pair methods, empty sequences and int16
    pairs that saturate and are recomputed, spread over the batch."""
    rng = np.random.default_rng(41)
    n = 2 * 262144 + 4321
    lens1 = rng.integers(0, 48, size=n)
    lens2 = rng.integers(0, 48, size=n)
    sat = [5, 300000, n - 7]            # identical 1,200-nt pairs: 36,000 at match +30
    for p in sat:
        lens1[p] = lens2[p] = 1200
    b1, o1 = workloads.random_packed(rng, lens1, workloads.DNA)
    b2, o2 = workloads.random_packed(rng, lens2, workloads.DNA)
    b2 = b2.copy()
    for p in sat:
        b2[o2[p]:o2[p + 1]] = b1[o1[p]:o1[p + 1]]
    pm = (rng.random(n) < 0.5).astype(np.uint8)
    dc = device.DeviceConfig(match=30, mismatch=-5, gap_open=-10, gap_extension=-1, lane_policy=device.LANES_INT16)
    log(f"case large one-shot batch: {n} pairs")
    got, flags = ctx.score_packed(b1, o1, b2, o2, dc, pm)
    want = O.score_packed(b1, o1, b2, o2, oracle_cfg(dc), pair_method=pm, threads=8)
    bad = np.nonzero(got.astype(np.int64) != want)[0]
    assert len(bad) == 0, f"{len(bad)} mismatches, first {bad[:5].tolist()}"
    assert np.all(flags[sat] & device.PAIR_RECOMPUTED) and np.all(got[sat] == 36000)
    one = device.Batch(ctx, b1, o1, b2, o2, dc, pm)
    one.run()
    same, same_flags = one.results()
    one.close()
    assert np.array_equal(same, got) and np.array_equal(same_flags, flags)


@pytest.mark.parametrize("one_query", [False, True])
def test_batch_does_not_read_host_inputs_after_creation(ctx, one_query):
    """A device-resident batch keeps no host copy of the residues: after creation the caller's
    buffers are overwritten, and repeated runs -- including the int32 re-plan and recompute of
    saturated int16 pairs, which takes the alphabet from the main plan -- still give the oracle's
    scores."""
    rng = np.random.default_rng(51)
    s1, s2 = workloads.dna_pairs(rng, 600, 50, 400, divergence=0.2, indel_fraction=0.3)
    if one_query:
        s1 = [b"ACGT" * 400] * len(s2)          # one 1,600-nt query
        s2[:4] = [b"ACGT" * 400] * 4               # identical targets saturate int16 at match +30
    else:
        s1[:4] = [b"ACGT" * 400] * 4
        s2[:4] = list(s1[:4])
    dc = device.DeviceConfig(match=30, mismatch=-5, gap_open=-10, gap_extension=-1, lane_policy=device.LANES_INT16)
    want = O.score_pairs(s1, s2, oracle_cfg(dc), threads=8)
    b2, o2 = workloads.pack(s2)
    if one_query:
        q = np.frombuffer(s1[0], dtype=np.uint8).copy()
        batch = device.Batch.one_to_many(ctx, q.tobytes(), b2, o2, dc)
        batch._keep[0][:] = ord("T")   # the query buffer the library was given
    else:
        b1, o1 = workloads.pack(s1)
        batch = device.Batch(ctx, b1, o1, b2, o2, dc)
        b1[:] = ord("T")
    b2[:] = ord("G")
    for rep in range(3):
        log(f"host-input independence, one_query={one_query}, run {rep}")
        batch.run()
        got, flags = batch.results()
        assert np.array_equal(got.astype(np.int64), want), f"run {rep}"
        assert np.all(flags[:4] & device.PAIR_RECOMPUTED)
    batch.close()


@pytest.mark.parametrize("ends", [(False, False, False), (True, True, False), (False, False, True)],
                         ids=["penalised", "leading-free", "last-column-free"])
def test_recentred_int16_long_pairs(ctx, ends):
    """Long global pairs in re-centred int16 lanes (wavefront kernel): scores far outside the
    int16 range (40,000 for identical 40 kbp sequences, about -100,000 along a penalised first
    row), exact and without a recompute; several rounds of stripes (the HBM boundary row), two
    pairs of different lengths per task."""
    fc, fr, lc = ends
    rng = np.random.default_rng(61)
    same = bytes(workloads.random_residues(rng, workloads.DNA, 40000))
    s1 = [same]
    s2 = [same]
    a, b = workloads.dna_pairs(rng, 3, 12000, 30000, divergence=0.15, indel_fraction=0.6)
    s1 += a
    s2 += b
    s1.append(bytes(workloads.random_residues(rng, workloads.DNA, 20000)))
    s2.append(bytes(workloads.random_residues(rng, workloads.DNA, 4000)))
    dc = device.DeviceConfig(match=1, mismatch=-4, gap_open=-6, gap_extension=-1, first_column_free=fc,
                             first_row_free=fr, last_column_free=lc)
    got, flags = check(ctx, s1, s2, dc, label=f"re-centred int16 long pairs ends={ends}")
    assert np.all(flags & device.PAIR_INT16) and not np.any(flags & device.PAIR_RECOMPUTED), flags
    if not (fc or fr or lc):
        assert got[0] == 40000
