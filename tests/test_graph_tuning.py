"""GPU: hipGraph replay of the main pass, planner tunings and the autotuner, all bit-exact
against the CPU oracle through the C ABI.

The replay test follows the rule for captured graphs: it replays the graph several times with
eager work and synchronisations of the program between the replays, and checks every replay
against the eagerly launched result.
"""
import ctypes
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


def uneven_batch(seed, count, hi=4000):
    """DNA pairs with log-normal lengths (median 250), half global half local."""
    rng = np.random.default_rng(seed)
    L = np.clip(np.round(rng.lognormal(np.log(250.0), 0.8, size=count)), 1, hi).astype(np.int64)
    b1, o1 = workloads.random_packed(rng, L, workloads.DNA)
    b2, o2 = workloads.mutate_packed(rng, b1, o1, rng.uniform(0.01, 0.2, size=count), 0.1, 0.7, workloads.DNA)
    pm = (np.arange(count) % 2).astype(np.uint8)
    return b1, o1, b2, o2, pm


def oracle_scores(b1, o1, b2, o2, dc, pm):
    return O.score_packed(b1, o1, b2, o2, oracle_cfg(dc), pair_method=pm, threads=8)


def test_graph_replays_match_eager(ctx):
    b1, o1, b2, o2, pm = uneven_batch(41, 3000)
    base = dict(match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
    eager = device.DeviceConfig(**base, tuning=device.Tuning(eager=1))
    graph = device.DeviceConfig(**base)
    want = oracle_scores(b1, o1, b2, o2, graph, pm)
    be = device.Batch(ctx, b1, o1, b2, o2, eager, pm)
    be.run()
    ref, ref_flags = be.results()
    assert np.array_equal(ref.astype(np.int64), want)
    bg = device.Batch(ctx, b1, o1, b2, o2, graph, pm)
    for rep in range(4):
        log(f"graph replay {rep}")
        bg.run()
        got, flags = bg.results()
        assert np.array_equal(got, ref), f"replay {rep}: {np.count_nonzero(got != ref)} mismatches"
        assert np.array_equal(flags, ref_flags)
        # eager work of the program between replays
        be.run()
        again, _ = be.results()
        assert np.array_equal(again, ref)
    bg.time(3)
    got, _ = bg.results()
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("form", [0, 1, 2, 3, 4], ids=["default", "queue_per_launch", "fused", "unfused", "global_fused"])
@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
def test_persistent_work_queue_launches(ctx, graph, form):
    """The launch forms of pa_tuning.persistent: the default (pair-row-table local launches fused
    into one launch with one work queue), resident blocks taking tasks from a work queue per launch
    (1), with every int16 local launch that differs only in R fused (2), no fusion (3), and the
    global method's launches fused too (4): exact on every replay (the last block resets the
    queue for the next pass), with eager work in between, int16 saturation recomputed, a matrix
    score model and a batch larger than the grid."""
    b1, o1, b2, o2, pm = uneven_batch(47, 30000, hi=3000)
    base = dict(match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
    dc = device.DeviceConfig(**base, tuning=device.Tuning(persistent=form, eager=0 if graph else 1))
    want = oracle_scores(b1, o1, b2, o2, dc, pm)
    bt = device.Batch(ctx, b1, o1, b2, o2, dc, pm)
    if form == 2:   # the local launches (one per R) are one launch
        unfused = device.Batch(ctx, b1, o1, b2, o2, device.DeviceConfig(**base, tuning=device.Tuning(persistent=1)), pm)
        assert bt.stats()["n_launches"] < unfused.stats()["n_launches"], (bt.stats(), unfused.stats())
        unfused.close()
    if form == 4:   # the global launches too, in two launches by rows per lane
        local_only = device.Batch(ctx, b1, o1, b2, o2, device.DeviceConfig(**base, tuning=device.Tuning(persistent=0)), pm)
        assert bt.stats()["n_launches"] < local_only.stats()["n_launches"], (bt.stats(), local_only.stats())
        local_only.close()
    be = device.Batch(ctx, b1, o1, b2, o2, device.DeviceConfig(**base, tuning=device.Tuning(eager=1)), pm)
    for rep in range(4):
        log(f"persistent pass {rep}")
        bt.run()
        got, _ = bt.results()
        assert np.array_equal(got.astype(np.int64), want), f"pass {rep}: {np.count_nonzero(got != want)} mismatches"
        be.run()
    bt.time(3)
    got, _ = bt.results()
    assert np.array_equal(got.astype(np.int64), want)
    bt.close()
    be.close()
    rng = np.random.default_rng(48)
    s1, s2 = workloads.protein_pairs(rng, 3000, 0, 600)
    t = tables.blosum62_standard()
    prot = device.DeviceConfig(method=device.METHOD_LOCAL, score_kind=device.SCORE_MATRIX,
                               symbols="".join(x for x, _ in t).encode(), matrix=tuple(v for _, r in t for v in r),
                               gap_open=-11, gap_extension=-1,
                               tuning=device.Tuning(persistent=form, eager=0 if graph else 1))
    want = O.score_pairs(s1, s2, oracle_cfg(prot), threads=8)
    for rep in range(2):
        got, _ = ctx.score(s1, s2, prot)
        assert np.array_equal(got.astype(np.int64), want), f"protein pass {rep}"
    # int16 saturation (identical 2.5 kbp sequences at +15 in the striped kernel) and its int32
    # recomputation, also as a work-queue launch
    d1, d2 = workloads.dna_pairs(rng, 400, 50, 400, divergence=0.1)
    same = workloads.unpack(*workloads.random_packed(rng, np.full(3, 2500), workloads.DNA))
    d1, d2 = d1 + same, d2 + list(same)
    p1, q1 = workloads.pack(d1)
    p2, q2 = workloads.pack(d2)
    for method in (device.METHOD_GLOBAL, device.METHOD_LOCAL):
        dna = device.DeviceConfig(method=method, match=15, mismatch=-5, gap_open=-10, gap_extension=-1,
                                  lane_policy=device.LANES_INT16,
                                  tuning=device.Tuning(persistent=form, eager=0 if graph else 1))
        want = O.score_pairs(d1, d2, oracle_cfg(dna), threads=8)
        bt = device.Batch(ctx, p1, q1, p2, q2, dna)
        for rep in range(3):
            bt.run()
            got, flags = bt.results()
            assert np.array_equal(got.astype(np.int64), want), f"saturating pass {rep}, method {method}"
        assert np.all(flags[-3:] & device.PAIR_RECOMPUTED) and np.all(got[-3:] == 2500 * 15)
        bt.close()


def test_batch_keeps_its_own_tuning(ctx):
    """A batch copies the caller's pa_tuning at creation: the int32 re-plan of its saturated pairs,
    made at its first run, follows the tuning the batch was created with even after the caller
    has changed (or freed) its struct.  Three identical 6 kbp pairs at +15 saturate int16 and are
    recomputed in int32 through the wavefront kernel (wavefront_tail < 0); planned with the
    caller's changed struct instead (R = 4 only, no wavefront kernel) the recomputation would be
    a 94-stripe striped sweep, about ten times slower."""
    rng = np.random.default_rng(71)
    same = workloads.unpack(*workloads.random_packed(rng, np.full(3, 6000), workloads.DNA))
    d1, d2 = workloads.dna_pairs(rng, 300, 50, 400, divergence=0.1)
    d1, d2 = d1 + same, d2 + list(same)
    dc = device.DeviceConfig(method=device.METHOD_LOCAL, match=15, mismatch=-5, gap_open=-10, gap_extension=-1,
                             lane_policy=device.LANES_INT16, tuning=device.Tuning(wavefront_tail=-1.0))
    want = O.score_pairs(d1, d2, oracle_cfg(dc), threads=8)
    b1, o1 = workloads.pack(d1)
    b2, o2 = workloads.pack(d2)
    b1, o1, b2, o2, _ = device.checked_batch(b1, o1, b2, o2, None)
    holder = device._CConfigHolder(dc)
    h = ctypes.c_void_p()
    ctx.check(ctx._lib.pa_batch_create(ctx.handle, ctypes.byref(holder.c), b1.ctypes.data, o1.ctypes.data,
                                       b2.ctypes.data, o2.ctypes.data, len(o1) - 1, None, ctypes.byref(h)))
    bt = device.Batch.__new__(device.Batch)
    bt.ctx, bt.handle, bt.n, bt._keep = ctx, h, len(o1) - 1, (b1, o1, b2, o2)
    # the caller's struct changes after the batch exists
    holder.tun.rows_mask = 1
    holder.tun.wavefront_min_rows = 60000
    holder.tun.wavefront_tail = 1e9
    for rep in range(2):
        bt.run()
        got, flags = bt.results()
        assert np.array_equal(got.astype(np.int64), want), f"pass {rep}"
    assert np.all(flags[-3:] & device.PAIR_RECOMPUTED) and np.all(got[-3:] == 6000 * 15)
    ms, _ = bt.time(2)
    log(f"pass with the int32 recomputation: {ms / 2:.1f} ms")
    assert ms / 2 < 60.0, ms / 2
    bt.close()


def test_graph_with_saturation_recompute(ctx):
    """The recompute pass (int32, eager) after a replayed int16 pass."""
    rng = np.random.default_rng(43)
    base = workloads.unpack(*workloads.random_packed(rng, np.full(4, 9000), workloads.DNA))
    short1, short2 = workloads.dna_pairs(rng, 200, 50, 300, divergence=0.05)
    s1, s2 = base + short1, list(base) + short2
    b1, o1 = workloads.pack(s1)
    b2, o2 = workloads.pack(s2)
    dc = device.DeviceConfig(match=4, mismatch=-5, gap_open=-10, gap_extension=-1, lane_policy=device.LANES_INT16)
    want = O.score_packed(b1, o1, b2, o2, oracle_cfg(dc), threads=8)
    bt = device.Batch(ctx, b1, o1, b2, o2, dc)
    for rep in range(3):
        log(f"saturating replay {rep}")
        bt.run()
        got, flags = bt.results()
        assert np.array_equal(got.astype(np.int64), want)
        assert np.all(flags[:4] & device.PAIR_RECOMPUTED)


TUNINGS = [
    device.Tuning(rows_mask=0x55),                      # R = 4, 8, 12, 16 only
    device.Tuning(rows_mask=0x2A, streams=1),           # R = 6, 10, 14 on one stream
    device.Tuning(wavefront_min_rows=512, streams=2, wavefront_tail=-1),   # medium pairs in the wavefront kernel
    device.Tuning(step_cost=64.0, eager=1),
    device.Tuning(cell_cost=4.0, step_cost=8.0, rows_mask=0x01),   # R = 4 only
]


@pytest.mark.parametrize("tuning", TUNINGS, ids=lambda t: f"{t}")
def test_tunings_are_exact(ctx, tuning):
    b1, o1, b2, o2, pm = uneven_batch(47, 1500, hi=3000)
    dc = device.DeviceConfig(match=2, mismatch=-3, gap_open=-5, gap_extension=-2, tuning=tuning)
    want = oracle_scores(b1, o1, b2, o2, dc, pm)
    got, _ = ctx.score_packed(b1, o1, b2, o2, dc, pm)
    bad = np.nonzero(got.astype(np.int64) != want)[0]
    assert len(bad) == 0, f"{tuning}: {len(bad)} mismatches, first {bad[:5].tolist()}"


def test_tunings_blosum_local(ctx):
    rng = np.random.default_rng(53)
    s1, s2 = workloads.protein_pairs(rng, 1200, 20, 600)
    t = tables.blosum62_standard()
    b1, o1 = workloads.pack(s1)
    b2, o2 = workloads.pack(s2)
    for tuning in TUNINGS[:3]:
        dc = device.DeviceConfig(method=device.METHOD_LOCAL, score_kind=device.SCORE_MATRIX,
                                 symbols="".join(s for s, _ in t).encode(), matrix=tuple(v for _, r in t for v in r),
                                 gap_open=-11, gap_extension=-1, tuning=tuning)
        want = O.score_packed(b1, o1, b2, o2, oracle_cfg(dc), threads=8)
        got, _ = ctx.score_packed(b1, o1, b2, o2, dc)
        assert np.array_equal(got.astype(np.int64), want), f"{tuning}"


def test_autotuner_result_is_exact_and_not_slower(ctx):
    b1, o1, b2, o2, pm = uneven_batch(59, 20000)
    dc = device.DeviceConfig(match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
    tuned = device.tune(ctx, dc, b1, o1, b2, o2, pm, sample_pairs=4000, iters=2)
    log(f"tuned: {tuned}")
    assert 1 <= tuned.streams <= 4 and tuned.rows_mask & 0x7F
    dct = device.DeviceConfig(match=2, mismatch=-3, gap_open=-5, gap_extension=-2, tuning=tuned)
    want = oracle_scores(b1, o1, b2, o2, dc, pm)
    bd = device.Batch(ctx, b1, o1, b2, o2, dc, pm)
    bt = device.Batch(ctx, b1, o1, b2, o2, dct, pm)
    bd.run()
    bt.run()
    assert np.array_equal(bt.results()[0].astype(np.int64), want)
    t_default = min(bd.time(3)[0] for _ in range(2))
    t_tuned = min(bt.time(3)[0] for _ in range(2))
    log(f"default {t_default / 3:.3f} ms, tuned {t_tuned / 3:.3f} ms per pass")
    # the tuner measured on a sample; on the whole batch it must not lose much
    assert t_tuned <= t_default * 1.15
