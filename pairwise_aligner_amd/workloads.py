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
"""
from __future__ import annotations

import hashlib
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


def config_uneven(seed: int = 3, count: int = 1_000_000):
    rng = np.random.default_rng(seed)
    L = np.clip(np.round(rng.lognormal(np.log(250.0), 0.8, size=count)), 50, 10_000).astype(np.int64)
    b1, o1 = random_packed(rng, L, DNA)
    div = rng.uniform(0.01, 0.20, size=count)
    b2, o2 = mutate_packed(rng, b1, o1, div, 0.1, 0.7, DNA)
    pm = (np.arange(count) % 2).astype(np.uint8)   # 0 global, 1 local
    return dict(name="uneven", b1=b1, o1=o1, b2=b2, o2=o2, pair_method=pm,
                scoring=dict(kind="unitary", match=2, mismatch=-3), gap=(-5, -2), method="mixed")


def config_long(seed: int = 4, count: int = 2_000):
    rng = np.random.default_rng(seed)
    b1, o1 = random_packed(rng, rng.integers(10_000, 100_001, size=count), DNA)
    b2, o2 = mutate_packed(rng, b1, o1, 0.15, 0.6, 0.5, DNA)
    return dict(name="long", b1=b1, o1=o1, b2=b2, o2=o2, pair_method=None,
                scoring=dict(kind="unitary", match=1, mismatch=-4), gap=(-6, -1), method="global")


def config_mixed(seed: int = 5, count: int = 500_000, chunk: int = 25_000):
    """Half DNA, half protein pairs; lengths uniform 100-5,000; identity uniform 50-100%
    (divergence 0-50% of positions, 90% of events substitutions).  Returns the DNA and protein
    halves separately because they use different scoring.  Generated in chunks of `chunk` pairs
    to bound the temporaries."""
    rng = np.random.default_rng(seed)
    half = count // 2
    out = {}
    for kind, alpha, p, n in (("dna", DNA, None, half), ("protein", AMINO, AMINO_FREQ, count - half)):
        parts = []
        for start in range(0, n, chunk):
            k = min(chunk, n - start)
            b1, o1 = random_packed(rng, rng.integers(100, 5_001, size=k), alpha, p)
            div = 1.0 - rng.uniform(0.5, 1.0, size=k)
            b2, o2 = mutate_packed(rng, b1, o1, div, 0.1, 0.7, alpha, p)
            parts.append((b1[:int(o1[-1])], o1, b2[:int(o2[-1])], o2))
        out[kind] = dict(zip(("b1", "o1", "b2", "o2"), concat_packed(parts)))
    out["name"] = "mixed"
    return out


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
