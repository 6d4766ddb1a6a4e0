"""Seeded synthetic workloads: random targets and mutated queries, generated with numpy's PCG64.

All generators are vectorised over the whole batch and return (seqs1, seqs2) as lists of bytes
or, for the large configurations, packed buffers (uint8 data + uint64 offsets).

Configurations (sizes as stated; names are used in bench.py):
  dna150     10,000 pairs, target 150 nt uniform ACGT; query = target with 10% substitutions and
             1% indel events (length geometric p=0.7); +4/-5, gap (-10,-1), global.
  protein    100,000 pairs, lengths uniform 50-500 drawn independently, residues i.i.d. from the
             background amino-acid frequencies below; BLOSUM62, gap (-11,-1), local.
  uneven     1,000,000 DNA pairs, length lognormal(median 250, sigma 0.8) clipped to 50-10,000;
             query at divergence uniform 1-20% (90% substitutions, 10% indels, geometric p=0.7);
             alternate pairs global / local; +2/-3, gap (-5,-2).
  long       2,000 DNA pairs, length uniform 10-100 kbp; 15% divergence, 60% indels (geometric
             p=0.5), 40% substitutions; +1/-4, gap (-6,-1), global.
  mixed      500,000 pairs, half DNA half protein, lengths uniform 100-5,000, identity uniform
             50-100%.
  saturating 20,000 protein pairs, lengths uniform 4,800-7,800, identity uniform 80-100%: about
             a quarter score beyond the int16 lanes' exact range, 31,744 (one in six beyond
             32,767; local BLOSUM62, gap (-11,-1)).
"""
from __future__ import annotations

import hashlib
import os
from concurrent.futures import ThreadPoolExecutor
from typing import Tuple

import numpy as np

DNA = np.frombuffer(b"ACGT", dtype=np.uint8)
AMINO = np.frombuffer(b"ARNDCQEGHILKMFPSTWYV", dtype=np.uint8)
# Robinson & Robinson (1991) background amino-acid frequencies, in the order of AMINO.
AMINO_FREQ = np.array([0.07805, 0.05129, 0.04487, 0.05364, 0.01925, 0.04264, 0.06295, 0.07377, 0.02199,
                       0.05142, 0.09019, 0.05744, 0.02243, 0.03856, 0.05203, 0.07120, 0.05841, 0.01330,
                       0.03216, 0.06441])
AMINO_FREQ = AMINO_FREQ / AMINO_FREQ.sum()


def pack(seqs) -> Tuple[np.ndarray, np.ndarray]:
    lens = np.fromiter((len(s) for s in seqs), dtype=np.uint64, count=len(seqs))
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    joined = b"".join(seqs)
    buf = np.frombuffer(joined, dtype=np.uint8).copy() if joined else np.zeros(1, dtype=np.uint8)
    return buf, off


def unpack(buf: np.ndarray, off: np.ndarray):
    b = buf.tobytes()
    return [b[int(off[i]):int(off[i + 1])] for i in range(len(off) - 1)]


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def random_residues(rng, alphabet: np.ndarray, n: int, p=None) -> np.ndarray:
    if p is None:
        return alphabet[rng.integers(0, len(alphabet), size=n)]
    # inverse CDF through a 65536-entry table (frequencies exact to 1/65536)
    cdf = np.cumsum(p)
    cdf[-1] = 1.0
    table = np.searchsorted(cdf, (np.arange(65536) + 0.5) / 65536.0, side="right").clip(0, len(alphabet) - 1)
    return alphabet[table[rng.integers(0, 65536, size=n, dtype=np.uint32)]]


def mutate_packed(rng, buf: np.ndarray, off: np.ndarray, event_rate, indel_fraction: float, geom_p: float,
                  alphabet: np.ndarray, p=None) -> Tuple[np.ndarray, np.ndarray]:
    """Mutates every sequence of a packed batch.  Per position, with probability event_rate (a
    scalar or one rate per sequence) an event happens: an indel with probability indel_fraction
    (insertion or deletion with equal odds, length ~ Geometric(geom_p)), else a substitution by a
    different residue.  Returns the packed mutated batch."""
    n_seq = len(off) - 1
    total = int(off[-1])
    lens = np.diff(off).astype(np.int64)
    rate = np.asarray(event_rate, dtype=np.float64)
    u = rng.random(total)
    r = np.repeat(np.broadcast_to(rate, (n_seq,)), lens) if rate.ndim else rate
    is_sub = (u < r) & (u >= r * indel_fraction)
    indel_pos = np.nonzero(u < r * indel_fraction)[0]
    del u, r
    ins_flag = rng.random(len(indel_pos)) < 0.5
    lengths = rng.geometric(geom_p, size=len(indel_pos))

    out = buf[:total].copy()
    # substitutions: shift by 1..k-1 within the alphabet (always a different residue)
    idx = np.nonzero(is_sub)[0]
    del is_sub
    if len(idx):
        lut = np.full(256, 0, dtype=np.int64)
        lut[alphabet] = np.arange(len(alphabet))
        shift = rng.integers(1, len(alphabet), size=len(idx))
        out[idx] = alphabet[(lut[out[idx]] + shift) % len(alphabet)]
    # deletions: drop `length` positions starting here, never across a sequence end
    keep = np.ones(total, dtype=bool)
    starts = indel_pos[~ins_flag]
    if len(starts):
        seq_end = np.repeat(off[1:].astype(np.int64), lens)
        ends = np.minimum(starts + lengths[~ins_flag], seq_end[starts])
        del seq_end
        cover = np.bincount(starts, minlength=total + 1) - np.bincount(ends, minlength=total + 1)
        keep = np.cumsum(cover[:-1]) == 0
    counts = keep.astype(np.int64)
    counts[indel_pos[ins_flag]] += lengths[ins_flag]
    new_total = int(counts.sum())
    res = np.empty(new_total, dtype=np.uint8)
    # position of each source position's first output slot
    starts_out = np.cumsum(counts) - counts
    kept = np.nonzero(keep)[0]
    res[starts_out[kept]] = out[kept]
    ins_mask = np.ones(new_total, dtype=bool)
    ins_mask[starts_out[kept]] = False
    n_ins = int(ins_mask.sum())
    if n_ins:
        res[ins_mask] = random_residues(rng, alphabet, n_ins, p)
    # new offsets
    csum = np.concatenate([[0], np.cumsum(counts)])
    new_off = csum[off.astype(np.int64)].astype(np.uint64)
    return res, new_off


def random_packed(rng, lengths: np.ndarray, alphabet: np.ndarray, p=None):
    lengths = np.asarray(lengths, dtype=np.int64)
    off = np.zeros(len(lengths) + 1, dtype=np.uint64)
    np.cumsum(lengths, out=off[1:])
    buf = random_residues(rng, alphabet, int(off[-1]), p)
    if buf.size == 0:
        buf = np.zeros(1, dtype=np.uint8)
    return buf, off


def dna_pairs(rng, count: int, lo: int, hi: int, divergence: float = 0.1, indel_fraction: float = 0.1,
              geom_p: float = 0.7):
    """Small helper for tests: list-of-bytes DNA pairs (targets uniform length lo..hi)."""
    b1, o1 = random_packed(rng, rng.integers(lo, hi + 1, size=count), DNA)
    b2, o2 = mutate_packed(rng, b1, o1, divergence, indel_fraction, geom_p, DNA)
    return unpack(b1, o1), unpack(b2, o2)


def protein_pairs(rng, count: int, lo: int, hi: int):
    """Independent protein pairs, lengths uniform lo..hi, background frequencies."""
    b1, o1 = random_packed(rng, rng.integers(lo, hi + 1, size=count), AMINO, AMINO_FREQ)
    b2, o2 = random_packed(rng, rng.integers(lo, hi + 1, size=count), AMINO, AMINO_FREQ)
    return unpack(b1, o1), unpack(b2, o2)


# ------------------------------------------------------------------------------------------
# The five benchmark configurations (packed form).
# ------------------------------------------------------------------------------------------

def config_dna150(seed: int = 1, count: int = 10_000):
    rng = np.random.default_rng(seed)
    b1, o1 = random_packed(rng, np.full(count, 150), DNA)
    # 10% substitutions + 1% indel events: event rate 0.11, indel fraction 1/11
    b2, o2 = mutate_packed(rng, b1, o1, 0.11, 1.0 / 11.0, 0.7, DNA)
    return dict(name="dna150", b1=b1, o1=o1, b2=b2, o2=o2, pair_method=None,
                scoring=dict(kind="unitary", match=4, mismatch=-5), gap=(-10, -1), method="global")


def config_protein(seed: int = 2, count: int = 100_000):
    rng = np.random.default_rng(seed)
    b1, o1 = random_packed(rng, rng.integers(50, 501, size=count), AMINO, AMINO_FREQ)
    b2, o2 = random_packed(rng, rng.integers(50, 501, size=count), AMINO, AMINO_FREQ)
    return dict(name="protein", b1=b1, o1=o1, b2=b2, o2=o2, pair_method=None,
                scoring=dict(kind="blosum62"), gap=(-11, -1), method="local")


def config_protein_1xn(seed: int = 6, count: int = 200_000, query_len: int = 300):
    """One protein query against `count` targets (a database search): lengths of the targets
    uniform 50..500, background residue frequencies.  b1/o1 repeat the query for the one-to-one
    path and the CPU baseline."""
    rng = np.random.default_rng(seed)
    q = random_residues(rng, AMINO, query_len, AMINO_FREQ)
    b2, o2 = random_packed(rng, rng.integers(50, 501, size=count), AMINO, AMINO_FREQ)
    b1 = np.tile(q, count) if count else np.zeros(1, dtype=np.uint8)
    o1 = (np.arange(count + 1, dtype=np.uint64) * np.uint64(query_len))
    return dict(name="protein_1xn", query=q.tobytes(), b1=b1, o1=o1, b2=b2, o2=o2, pair_method=None,
                scoring=dict(kind="blosum62"), gap=(-11, -1), method="local")


def _chunks(count: int, chunk: int):
    return [(c, start, min(chunk, count - start)) for c, start in enumerate(range(0, count, chunk))]


def _map_chunks(fn, items):
    """fn over independent chunks on a few threads (numpy releases the GIL in its bulk work);
    results in item order."""
    workers = max(1, min(len(items), 8, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1))
    if workers == 1:
        return [fn(it) for it in items]
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(fn, items))


def take(buf: np.ndarray, off: np.ndarray, n: int):
    """The first n sequences of a packed batch."""
    off = off[:n + 1]
    end = int(off[-1])
    return (buf[:end] if end else np.zeros(1, dtype=np.uint8)), off


def pick(buf: np.ndarray, off: np.ndarray, idx):
    """Sequences idx (in that order) of a packed batch, packed."""
    return pack([buf[int(off[i]):int(off[i + 1])].tobytes() for i in idx])


def config_uneven(seed: int = 3, count: int = 1_000_000, limit: int = None, chunk: int = 50_000):
    """Chunk c of `chunk` pairs comes from its own stream (seed, c), so the first `limit` pairs
    are the same pairs as in the whole batch."""
    n = count if limit is None else min(limit, count)

    def make(item):
        c, start, k = item
        rng = np.random.default_rng([seed, c])
        L = np.clip(np.round(rng.lognormal(np.log(250.0), 0.8, size=k)), 50, 10_000).astype(np.int64)
        b1, o1 = random_packed(rng, L, DNA)
        div = rng.uniform(0.01, 0.20, size=k)
        b2, o2 = mutate_packed(rng, b1, o1, div, 0.1, 0.7, DNA)
        return (b1[:int(o1[-1])], o1, b2[:int(o2[-1])], o2)

    parts = _map_chunks(make, [it for it in _chunks(count, chunk) if it[1] < n])
    b1, o1, b2, o2 = concat_packed(parts)
    b1, o1 = take(b1, o1, n)
    b2, o2 = take(b2, o2, n)
    pm = (np.arange(n) % 2).astype(np.uint8)   # 0 global, 1 local
    return dict(name="uneven", b1=b1, o1=o1, b2=b2, o2=o2, pair_method=pm,
                scoring=dict(kind="unitary", match=2, mismatch=-3), gap=(-5, -2), method="mixed")


def config_long(seed: int = 4, count: int = 2_000, select=None, chunk: int = 50):
    """Target lengths come from stream (seed, 0); the residues of chunk c of `chunk` pairs from
    stream (seed, 1, c), so any selection of pairs (`select`: pair indices, in that order) is
    generated without the rest.  `long_lengths` gives the target lengths alone."""
    lengths = long_lengths(seed, count)
    want = range(count) if select is None else [int(i) for i in select]
    need = sorted({i // chunk for i in want})

    def make(c):
        start = c * chunk
        k = min(chunk, count - start)
        rng = np.random.default_rng([seed, 1, c])
        b1, o1 = random_packed(rng, lengths[start:start + k], DNA)
        b2, o2 = mutate_packed(rng, b1, o1, 0.15, 0.6, 0.5, DNA)
        return (b1, o1, b2, o2)

    made = dict(zip(need, _map_chunks(make, need)))
    if select is None:
        b1, o1, b2, o2 = concat_packed([(m[0][:int(m[1][-1])], m[1], m[2][:int(m[3][-1])], m[3])
                                        for _, m in sorted(made.items())])
    else:
        s1, s2 = [], []
        for i in want:
            mb1, mo1, mb2, mo2 = made[i // chunk]
            j = i % chunk
            s1.append(mb1[int(mo1[j]):int(mo1[j + 1])].tobytes())
            s2.append(mb2[int(mo2[j]):int(mo2[j + 1])].tobytes())
        b1, o1 = pack(s1)
        b2, o2 = pack(s2)
    return dict(name="long", b1=b1, o1=o1, b2=b2, o2=o2, pair_method=None,
                scoring=dict(kind="unitary", match=1, mismatch=-4), gap=(-6, -1), method="global")


def long_lengths(seed: int = 4, count: int = 2_000) -> np.ndarray:
    return np.random.default_rng([seed, 0]).integers(10_000, 100_001, size=count)


def config_mixed(seed: int = 5, count: int = 500_000, limit: int = None, chunk: int = 25_000):
    """Half DNA, half protein pairs; lengths uniform 100-5,000; identity uniform 50-100%
    (divergence 0-50% of positions, 90% of events substitutions).  Returns the DNA and protein
    halves separately because they use different scoring.  Chunk c of each half comes from its
    own stream (seed, half, c), so `limit` (pairs per half) gives the first pairs of each half of
    the whole batch."""
    half = count // 2
    out = {}
    for h, (kind, alpha, p, n) in enumerate((("dna", DNA, None, half), ("protein", AMINO, AMINO_FREQ, count - half))):
        keep = n if limit is None else min(limit, n)

        def make(item, h=h, alpha=alpha, p=p):
            c, start, k = item
            rng = np.random.default_rng([seed, h, c])
            b1, o1 = random_packed(rng, rng.integers(100, 5_001, size=k), alpha, p)
            div = 1.0 - rng.uniform(0.5, 1.0, size=k)
            b2, o2 = mutate_packed(rng, b1, o1, div, 0.1, 0.7, alpha, p)
            return (b1[:int(o1[-1])], o1, b2[:int(o2[-1])], o2)

        parts = _map_chunks(make, [it for it in _chunks(n, chunk) if it[1] < keep])
        b1, o1, b2, o2 = concat_packed(parts)
        b1, o1 = take(b1, o1, keep)
        b2, o2 = take(b2, o2, keep)
        out[kind] = dict(b1=b1, o1=o1, b2=b2, o2=o2)
    out["name"] = "mixed"
    return out


def config_saturating(seed: int = 8, count: int = 20_000, limit: int = None, chunk: int = 2_000):
    """Protein pairs whose local scores straddle the int16 limit: lengths uniform 4,800-7,800,
    identity uniform 80-100% (90% of events substitutions), background residue frequencies.
    About a quarter score above 31,744 (one in six 32,767 or more) under BLOSUM62 gap (-11,-1) -- the slice where running
    every pair in int16 lanes costs a recomputation.  Chunk c from stream (seed, c)."""
    n = count if limit is None else min(limit, count)

    def make(item):
        c, start, k = item
        rng = np.random.default_rng([seed, c])
        b1, o1 = random_packed(rng, rng.integers(4_800, 7_801, size=k), AMINO, AMINO_FREQ)
        div = 1.0 - rng.uniform(0.8, 1.0, size=k)
        b2, o2 = mutate_packed(rng, b1, o1, div, 0.1, 0.7, AMINO, AMINO_FREQ)
        return (b1[:int(o1[-1])], o1, b2[:int(o2[-1])], o2)

    parts = _map_chunks(make, [it for it in _chunks(count, chunk) if it[1] < n])
    b1, o1, b2, o2 = concat_packed(parts)
    b1, o1 = take(b1, o1, n)
    b2, o2 = take(b2, o2, n)
    return dict(name="saturating", b1=b1, o1=o1, b2=b2, o2=o2, pair_method=None,
                scoring=dict(kind="blosum62"), gap=(-11, -1), method="local")


def concat_packed(parts):
    """Concatenates packed chunks [(b1, o1, b2, o2), ...] into one packed batch."""
    res = []
    for side in (0, 2):
        bufs = [p[side] for p in parts]
        offs = [p[side + 1] for p in parts]
        total = 0
        cat_off = [np.zeros(1, dtype=np.uint64)]
        for o in offs:
            cat_off.append(o[1:] + np.uint64(total))
            total += int(o[-1])
        buf = np.concatenate(bufs) if total else np.zeros(1, dtype=np.uint8)
        res += [buf, np.concatenate(cat_off)]
    return res
