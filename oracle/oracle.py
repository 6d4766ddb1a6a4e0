"""CPU oracle for batched affine-gap (Gotoh) score-only alignment -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import this
module.  It is the checker: the device library in ``pairwise_aligner_amd`` never imports it.

Two independent restatements of the reference's scalar semantics live here:

* ``score_pairs`` -- ctypes front end of ``gotoh_oracle.c`` (64-bit DP, one pass per pair),
  which follows ``affine/affine_dp_algorithm.hpp:139-156``,
  ``affine/affine_initialisation_strategy.hpp:38-51``,
  ``tracker/tracker_global_scalar.hpp:36-61`` and ``tracker/tracker_local_scalar.hpp:38-49``;
* ``brute_force_score`` -- enumerates every alignment of two short sequences and scores it
  with the affine gap cost go + k*ge per maximal gap run, free leading/trailing runs as the
  end-gap flags say, and for the local method every pair of substrings.  It shares no code
  with the DP and pins it on tiny inputs.

Parameters are plain values (no product types) so that a translation bug in the product's
configuration layer cannot leak into the checker.
"""
from __future__ import annotations

import ctypes
import functools
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

GLOBAL, LOCAL = 0, 1
UNITARY, MATRIX = 0, 1
INT32_MIN = -(2 ** 31)


@dataclass(frozen=True)
class OracleConfig:
    mode: int = GLOBAL
    first_column_free: bool = False
    first_row_free: bool = False
    last_column_free: bool = False
    last_row_free: bool = False
    gap_open: int = -10
    gap_extension: int = -1
    score_kind: int = UNITARY
    match: int = 4
    mismatch: int = -5
    symbols: bytes = b""           # matrix alphabet, rank = position
    matrix: Optional[tuple] = None  # flattened dim*dim, [rank1][rank2]


class _CConfig(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int32),
        ("first_column_free", ctypes.c_int32),
        ("first_row_free", ctypes.c_int32),
        ("last_column_free", ctypes.c_int32),
        ("last_row_free", ctypes.c_int32),
        ("gap_open", ctypes.c_int32),
        ("gap_extension", ctypes.c_int32),
        ("score_kind", ctypes.c_int32),
        ("match", ctypes.c_int32),
        ("mismatch", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("matrix", ctypes.POINTER(ctypes.c_int32)),
        ("symbols", ctypes.POINTER(ctypes.c_uint8)),
    ]


@functools.lru_cache(maxsize=1)
def _lib():
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    lib.pao_dp_batch.restype = ctypes.c_int
    lib.pao_dp_batch.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_int64, ctypes.POINTER(_CConfig), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.pao_score_batch.restype = ctypes.c_int
    lib.pao_score_batch.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ctypes.c_int64, ctypes.POINTER(_CConfig), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    return lib


def pack(seqs: Sequence[bytes]):
    """Concatenate byte strings; return (uint8 buffer, uint64 offsets of length n+1)."""
    lens = np.fromiter((len(s) for s in seqs), dtype=np.uint64, count=len(seqs))
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    buf = np.frombuffer(b"".join(seqs), dtype=np.uint8) if len(seqs) and off[-1] else np.zeros(1, np.uint8)
    return np.ascontiguousarray(buf), off


def score_packed(buf1, off1, buf2, off2, cfg: OracleConfig, pair_method=None, threads: int = 1) -> np.ndarray:
    """Score pairs given as packed buffers + offsets.  Returns int64 scores."""
    n = len(off1) - 1
    scores = np.zeros(max(n, 1), dtype=np.int64)
    mat = sym = None
    c = _CConfig(cfg.mode, int(cfg.first_column_free), int(cfg.first_row_free),
                 int(cfg.last_column_free), int(cfg.last_row_free), cfg.gap_open, cfg.gap_extension,
                 cfg.score_kind, cfg.match, cfg.mismatch, len(cfg.symbols), None, None)
    if cfg.score_kind == MATRIX:
        mat = (ctypes.c_int32 * len(cfg.matrix))(*cfg.matrix)
        sym = (ctypes.c_uint8 * len(cfg.symbols))(*cfg.symbols)
        c.matrix = mat
        c.symbols = sym
    pm = None
    if pair_method is not None:
        pm = np.ascontiguousarray(pair_method, dtype=np.uint8)
    buf1 = np.ascontiguousarray(buf1, dtype=np.uint8)
    buf2 = np.ascontiguousarray(buf2, dtype=np.uint8)
    off1 = np.ascontiguousarray(off1, dtype=np.uint64)
    off2 = np.ascontiguousarray(off2, dtype=np.uint64)
    rc = _lib().pao_score_batch(buf1.ctypes.data, off1.ctypes.data, buf2.ctypes.data, off2.ctypes.data,
                                n, ctypes.byref(c), None if pm is None else pm.ctypes.data,
                                scores.ctypes.data, threads)
    if rc != 0:
        raise ValueError("symbol outside the substitution matrix alphabet")
    return scores[:n]


def dp_vectors_packed(buf1, off1, buf2, off2, cfg: OracleConfig, pair_method=None, threads: int = 1,
                      gaps: bool = False):
    """Scores plus the final DP row and column of every pair (int64): rows[(off2[p]-off2[0]) + p + j]
    = H(m_p, j), j = 0..n_p, and cols[(off1[p]-off1[0]) + p + i] = H(i, n_p), i = 0..m_p.  With
    gaps, also the gap-state halves of the same cells in the same layout (row_gaps: V(m_p+1, j),
    col_gaps: W(i, n_p+1), entry 0 of each go + ge): returns (scores, rows, cols, row_gaps, col_gaps)."""
    n = len(off1) - 1
    scores = np.zeros(max(n, 1), dtype=np.int64)
    off1 = np.ascontiguousarray(off1, dtype=np.uint64)
    off2 = np.ascontiguousarray(off2, dtype=np.uint64)
    rows = np.zeros(max(int(off2[-1] - off2[0]) + n, 1), dtype=np.int64)
    cols = np.zeros(max(int(off1[-1] - off1[0]) + n, 1), dtype=np.int64)
    row_gaps = np.zeros_like(rows) if gaps else None
    col_gaps = np.zeros_like(cols) if gaps else None
    c = _CConfig(cfg.mode, int(cfg.first_column_free), int(cfg.first_row_free),
                 int(cfg.last_column_free), int(cfg.last_row_free), cfg.gap_open, cfg.gap_extension,
                 cfg.score_kind, cfg.match, cfg.mismatch, len(cfg.symbols), None, None)
    mat = sym = None
    if cfg.score_kind == MATRIX:
        mat = (ctypes.c_int32 * len(cfg.matrix))(*cfg.matrix)
        sym = (ctypes.c_uint8 * len(cfg.symbols))(*cfg.symbols)
        c.matrix = mat
        c.symbols = sym
    pm = None if pair_method is None else np.ascontiguousarray(pair_method, dtype=np.uint8)
    buf1 = np.ascontiguousarray(buf1, dtype=np.uint8)
    buf2 = np.ascontiguousarray(buf2, dtype=np.uint8)
    rc = _lib().pao_dp_batch(buf1.ctypes.data, off1.ctypes.data, buf2.ctypes.data, off2.ctypes.data, n,
                             ctypes.byref(c), None if pm is None else pm.ctypes.data, scores.ctypes.data,
                             rows.ctypes.data, cols.ctypes.data,
                             None if row_gaps is None else row_gaps.ctypes.data,
                             None if col_gaps is None else col_gaps.ctypes.data, threads)
    if rc != 0:
        raise ValueError("symbol outside the substitution matrix alphabet")
    nr, nc = int(off2[-1] - off2[0]) + n, int(off1[-1] - off1[0]) + n
    if gaps:
        return scores[:n], rows[:nr], cols[:nc], row_gaps[:nr], col_gaps[:nc]
    return scores[:n], rows[:nr], cols[:nc]


SIMD_LIB_PATH = os.path.join(_HERE, "libsimdport.so")


@functools.lru_cache(maxsize=1)
def _simd_lib():
    if not os.path.exists(SIMD_LIB_PATH):
        raise RuntimeError(f"{SIMD_LIB_PATH} missing: run `make -C oracle`")
    L = ctypes.CDLL(SIMD_LIB_PATH)
    L.pas_isa.restype = ctypes.c_int
    L.pas_score_batch.argtypes = [ctypes.POINTER(_CConfig), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.pas_score_batch.restype = ctypes.c_int
    return L


def simd_isa() -> int:
    """Vector ISA of the CPU baseline on this host: 512 (AVX-512BW), 256 (AVX2) or 0."""
    return int(_simd_lib().pas_isa())


def simd_score_packed(buf1, off1, buf2, off2, cfg: OracleConfig, pair_method=None, threads: int = 1) -> np.ndarray:
    """The CPU baseline (oracle/simd_port.cpp): inter-sequence SIMD lanes, int16 where the bulk's
    bound allows, int32 otherwise.  Returns int32 scores; checked against score_packed in tests."""
    n = len(off1) - 1
    scores = np.zeros(max(n, 1), dtype=np.int32)
    c = _CConfig(cfg.mode, int(cfg.first_column_free), int(cfg.first_row_free),
                 int(cfg.last_column_free), int(cfg.last_row_free), cfg.gap_open, cfg.gap_extension,
                 cfg.score_kind, cfg.match, cfg.mismatch, len(cfg.symbols), None, None)
    mat = sym = None
    if cfg.score_kind == MATRIX:
        mat = (ctypes.c_int32 * len(cfg.matrix))(*cfg.matrix)
        sym = (ctypes.c_uint8 * len(cfg.symbols))(*cfg.symbols)
        c.matrix = mat
        c.symbols = sym
    pm = None if pair_method is None else np.ascontiguousarray(pair_method, dtype=np.uint8)
    buf1 = np.ascontiguousarray(buf1, dtype=np.uint8)
    buf2 = np.ascontiguousarray(buf2, dtype=np.uint8)
    off1 = np.ascontiguousarray(off1, dtype=np.uint64)
    off2 = np.ascontiguousarray(off2, dtype=np.uint64)
    rc = _simd_lib().pas_score_batch(ctypes.byref(c), buf1.ctypes.data, off1.ctypes.data, buf2.ctypes.data,
                                     off2.ctypes.data, n, None if pm is None else pm.ctypes.data, threads,
                                     scores.ctypes.data)
    if rc != 0:
        raise ValueError({-1: "no AVX2/AVX-512 on this CPU", -2: "symbol outside the substitution matrix alphabet",
                          -3: "score range beyond int32"}.get(rc, f"error {rc}"))
    return scores[:n]


def score_pairs(seqs1: Sequence[bytes], seqs2: Sequence[bytes], cfg: OracleConfig,
                pair_method=None, threads: int = 1) -> np.ndarray:
    b1, o1 = pack(seqs1)
    b2, o2 = pack(seqs2)
    return score_packed(b1, o1, b2, o2, cfg, pair_method, threads)


def score(s1: bytes, s2: bytes, cfg: OracleConfig) -> int:
    return int(score_pairs([s1], [s2], cfg)[0])


# --------------------------------------------------------------------------------------------
# Brute force: enumerate alignments (independent of the DP).
# --------------------------------------------------------------------------------------------

def _subst(cfg: OracleConfig, a: int, b: int) -> int:
    if cfg.score_kind == MATRIX:
        d = len(cfg.symbols)
        return cfg.matrix[cfg.symbols.index(a) * d + cfg.symbols.index(b)]
    return cfg.match if a == b else cfg.mismatch


def _alignments(m: int, n: int):
    """Yield every alignment of lengths (m, n) as a tuple of moves: 'M' (diagonal),
    'V' (consume seq1 only), 'H' (consume seq2 only)."""
    if m == 0 and n == 0:
        yield ()
        return
    if m > 0 and n > 0:
        for rest in _alignments(m - 1, n - 1):
            yield rest + ("M",)
    if m > 0:
        for rest in _alignments(m - 1, n):
            yield rest + ("V",)
    if n > 0:
        for rest in _alignments(m, n - 1):
            yield rest + ("H",)


def _score_alignment(moves, s1: bytes, s2: bytes, cfg: OracleConfig, free_lead_v, free_lead_h,
                     free_trail_v, free_trail_h) -> int:
    total = 0
    i = j = 0
    k = 0
    nmoves = len(moves)
    while k < nmoves:
        mv = moves[k]
        if mv == "M":
            total += _subst(cfg, s1[i], s2[j])
            i += 1
            j += 1
            k += 1
            continue
        run = k
        while run < nmoves and moves[run] == mv:
            run += 1
        length = run - k
        leading = k == 0
        trailing = run == nmoves
        free = (mv == "V" and ((leading and free_lead_v) or (trailing and free_trail_v))) or \
               (mv == "H" and ((leading and free_lead_h) or (trailing and free_trail_h)))
        if not free:
            total += cfg.gap_open + cfg.gap_extension * length
        if mv == "V":
            i += length
        else:
            j += length
        k = run
    return total


def brute_force_score(s1: bytes, s2: bytes, cfg: OracleConfig) -> int:
    """Best score over all alignments; intended for len(s1), len(s2) <= 6."""
    if cfg.mode == LOCAL:
        best = INT32_MIN
        m, n = len(s1), len(s2)
        if m == 0 or n == 0:
            return INT32_MIN
        best = 0  # the zero clamp: an empty local alignment
        for i0 in range(m):
            for i1 in range(i0 + 1, m + 1):
                for j0 in range(n):
                    for j1 in range(j0 + 1, n + 1):
                        a, b = s1[i0:i1], s2[j0:j1]
                        for mv in _alignments(len(a), len(b)):
                            best = max(best, _score_alignment(mv, a, b, cfg, False, False, False, False))
        return best
    best = None
    for mv in _alignments(len(s1), len(s2)):
        sc = _score_alignment(mv, s1, s2, cfg, cfg.first_column_free, cfg.first_row_free,
                              cfg.last_column_free, cfg.last_row_free)
        best = sc if best is None else max(best, sc)
    return best
