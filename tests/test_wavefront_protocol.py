"""CPU model of the wavefront kernel's boundary protocol (pairwise_aligner_amd/csrc/pa_kernels.hip,
wavefront_kernel): the waves of a workgroup sweep the stripes of one pair round robin and hand
the bottom row of each stripe to the next wave through an LDS ring of 256 columns (from the last
wave to the first through an HBM row), in chunks of 64 columns with progress counters.

The model replays each wave's chunk-end events in the kernel's order with an adversarial random
interleaving and checks that no interleaving deadlocks, that every read of the ring finds the
position it expects (no slot overwritten before it was consumed, none read before it was
written), and that the HBM row is read before the next round overwrites it.  Planted protocol
mistakes are rejected.
"""
import random

import pytest

G, RING = 64, 256


class ProtocolError(AssertionError):
    pass


def simulate(S, N, NW=4, seed=0, ring_slack=0, lookahead=3):
    """ring_slack > 0 or lookahead < 3 plant a mistake (a producer that overwrites too early, a
    consumer that reads too early)."""
    rng = random.Random(seed)
    prod = [0] * NW
    cons = [0] * NW
    ring = [[None] * RING for _ in range(NW)]
    gbnd = [None] * N
    gread = set()
    nsteps = ((N + 2 * G - 2) // G) * G

    def wave_events(w):
        s = w
        while s < S:
            r = s // NW
            pos0 = r * N
            src = 0 if s == 0 else (1 if w > 0 else 2)
            dst = 0 if s + 1 >= S else (1 if w < NW - 1 else 2)
            src_w = w - 1 if src == 1 else NW - 1
            src_pos0 = pos0 if src == 1 else pos0 - N

            def read_cols(lo, hi):
                for c in range(max(lo, 0), min(hi, N)):
                    if src == 1:
                        got = ring[w - 1][(pos0 + c) % RING]
                        if got != pos0 + c:
                            raise ProtocolError(f"wave {w} stripe {s} column {c}: slot holds {got}")
                    else:
                        if gbnd[c] != r - 1:
                            raise ProtocolError(f"wave 0 round {r} column {c}: HBM row of round {gbnd[c]}")
                        gread.add((r, c))

            if src:
                need = src_pos0 + min(2 * G, N)
                yield lambda need=need, src_w=src_w: prod[src_w] >= need
                read_cols(0, 2 * G)
                if src == 1:
                    cons[w] = pos0 + min(2 * G, N)
            for k0 in range(0, nsteps, G):
                yield None   # the chunk's steps
                if src:
                    need = src_pos0 + min(k0 + lookahead * G, N)
                    yield lambda need=need, src_w=src_w: prod[src_w] >= need
                    read_cols(k0 + 2 * G, k0 + 3 * G)
                    if src == 1:
                        cons[w] = pos0 + min(k0 + 3 * G, N)
                if dst:
                    hi = min(k0 + 1, N)
                    cols = range(max(0, k0 - (G - 1)), hi)
                    if dst == 1:
                        need = pos0 + hi - RING - ring_slack
                        if need > 0:
                            yield lambda need=need: cons[w + 1] >= need
                        for c in cols:
                            ring[w][(pos0 + c) % RING] = pos0 + c
                    else:
                        for c in cols:
                            # round r+1's wave 0 reads this row; round r's must have been read first
                            if r >= 1 and gbnd[c] == r - 1 and (r, c) not in gread:
                                raise ProtocolError(f"round {r} overwrote HBM column {c} before it was read")
                            gbnd[c] = r
                    if hi > k0 + 1 - G:
                        prod[w] = pos0 + hi
            s += NW

    gens = [wave_events(w) for w in range(NW)]
    pending = [None] * NW
    fetched = [False] * NW
    done = [False] * NW
    while not all(done):
        progressed = False
        order = list(range(NW))
        rng.shuffle(order)
        for w in order:
            for _ in range(rng.randint(1, 3)):
                if done[w]:
                    break
                if not fetched[w]:
                    try:
                        pending[w] = next(gens[w])
                    except StopIteration:
                        done[w] = True
                        progressed = True
                        break
                    fetched[w] = True
                    progressed = True
                cond = pending[w]
                if cond is not None and not cond():
                    break
                fetched[w] = False
                progressed = True
        if not progressed:
            raise ProtocolError(f"deadlock S={S} N={N}: prod {prod} cons {cons}")


@pytest.mark.parametrize("nw", [4, 8])
def test_protocol_has_no_deadlock_or_race(nw):
    cases = 0
    for S in list(range(1, 11)) + [13, 17]:
        for N in [1, 2, 63, 64, 65, 128, 129, 192, 255, 256, 257, 513, 900]:
            simulate(S, N, NW=nw, seed=S * 1000 + N)
            cases += 1
    assert cases == 12 * 13


@pytest.mark.parametrize("mistake", [dict(ring_slack=128), dict(lookahead=2)])
def test_planted_protocol_mistakes_are_caught(mistake):
    with pytest.raises(ProtocolError):
        for S in (2, 3, 5):
            for N in (300, 700):
                for seed in range(4):
                    simulate(S, N, NW=4, seed=seed, **mistake)
