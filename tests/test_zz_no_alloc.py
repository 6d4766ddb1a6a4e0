"""pa_score_batch_noalloc (every device buffer in the library's static per-device scratch) against
pa_score_batch on the same inputs: identical scores and flags, on calls in a row of decreasing
size (the scratch keeps the previous call's contents), through the striped and the wavefront
kernels, int16 saturation with the int32 re-plan inside the scratch, per-pair methods, a matrix
score model; a batch too large for the scratch is refused before anything runs, and the scratch
still serves the next call."""
import numpy as np
import pytest

from pairwise_aligner_amd import device, tables, workloads

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


@pytest.fixture(scope="module")
def ctx():
    return device.default_context(0)


def dna_batch(seed, n, lo, hi, n_long=0, n_saturating=0):
    rng = np.random.default_rng(seed)
    s1, s2 = workloads.dna_pairs(rng, n, lo, hi, divergence=0.15)
    if n_long:
        l1, l2 = workloads.dna_pairs(rng, n_long, 3300, 4200, divergence=0.1)
        s1, s2 = s1 + l1, s2 + l2
    if n_saturating:
        same = workloads.unpack(*workloads.random_packed(rng, np.full(n_saturating, 9000), workloads.DNA))
        s1, s2 = s1 + same, s2 + list(same)
    b1, o1 = workloads.pack(s1)
    b2, o2 = workloads.pack(s2)
    return b1, o1, b2, o2


def same_as_score_batch(ctx, b1, o1, b2, o2, dc, pm=None):
    want, want_flags = ctx.score_packed(b1, o1, b2, o2, dc, pm)
    got, flags = ctx.score_packed_noalloc(b1, o1, b2, o2, dc, pm)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, f"{len(bad)} score differences, first {bad[:5].tolist()}: {got[bad[:5]]} vs {want[bad[:5]]}"
    assert np.array_equal(flags, want_flags)
    return got, flags


def test_noalloc_matches_score_batch_on_calls_in_a_row(ctx):
    dc = device.DeviceConfig(match=4, mismatch=-5, gap_open=-10, gap_extension=-1, lane_policy=device.LANES_INT16)
    # first call: striped and wavefront tasks, per-pair methods, four pairs that saturate int16
    b1, o1, b2, o2 = dna_batch(1, 3000, 50, 400, n_long=3, n_saturating=4)
    pm = (np.arange(len(o1) - 1) % 2).astype(np.uint8)
    got, flags = same_as_score_batch(ctx, b1, o1, b2, o2, dc, pm)
    assert np.count_nonzero(flags & device.PAIR_RECOMPUTED) >= 2
    # second call, smaller, same configuration: nothing of the first call leaks into it
    b1, o1, b2, o2 = dna_batch(2, 500, 20, 300)
    same_as_score_batch(ctx, b1, o1, b2, o2, dc)
    # third call, smaller again, a matrix score model (LDS tables) and the local method
    rng = np.random.default_rng(3)
    s1, s2 = workloads.protein_pairs(rng, 200, 0, 450)
    p1, q1 = workloads.pack(s1)
    p2, q2 = workloads.pack(s2)
    t = tables.blosum62_standard()
    prot = device.DeviceConfig(method=device.METHOD_LOCAL, score_kind=device.SCORE_MATRIX,
                               symbols="".join(s for s, _ in t).encode(), matrix=tuple(v for _, r in t for v in r),
                               gap_open=-11, gap_extension=-1)
    same_as_score_batch(ctx, p1, q1, p2, q2, prot)
    # an empty batch
    got, _ = ctx.score_packed_noalloc(np.zeros(1, np.uint8), np.zeros(1, np.uint64), np.zeros(1, np.uint8),
                                      np.zeros(1, np.uint64), dc)
    assert len(got) == 0


def test_noalloc_refuses_what_does_not_fit_and_keeps_working(ctx):
    dc = device.DeviceConfig(match=4, mismatch=-5, gap_open=-10, gap_extension=-1)
    n = SCRATCH_BYTES // 300 + 1000   # the residues alone exceed the scratch
    b = np.full(n * 150, ord("A"), dtype=np.uint8)
    o = (np.arange(n + 1, dtype=np.uint64) * np.uint64(150))
    with pytest.raises(device.PaError) as e:
        ctx.score_packed_noalloc(b, o, b, o, dc)
    assert e.value.status == 5 and "scratch" in str(e.value)
    with pytest.raises(device.PaError) as e:
        ctx.score_packed_noalloc(b[:150], o[:2], b[:150], o[:2],
                                 device.DeviceConfig(lane_policy=device.LANES_LEARNED))
    assert e.value.status == 3
    b1, o1, b2, o2 = dna_batch(4, 800, 100, 200)
    same_as_score_batch(ctx, b1, o1, b2, o2, dc)
