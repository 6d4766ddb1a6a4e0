"""Python mirror of the reference's configuration API (namespace ``seqan::pairwise_aligner::cfg``).

Every rule function takes an optional predecessor rule as its first argument, exactly like the
reference's CPOs (``rule_base.hpp:41-57``), so both spellings the reference's tests use work:

    cfg.score_model_unitary(cfg.method_global(cfg.gap_model_affine(-10, -1), leading, trailing), 4, -5)
    cfg.method_global(cfg.gap_model_affine(cfg.score_model_unitary(4, -5), -10, -1), leading, trailing)

``configure_aligner`` walks the chain like ``configure_aligner.hpp:165-185``: it requires exactly
one score-model rule and one gap-model rule (the reference ``static_assert``s this, here a
``TypeError``); the method rule is optional and defaults to global with penalised end gaps
(``initial.hpp:49-65``); ``method_local`` makes all end gaps free (``method_local.hpp:41-49``).
The resulting aligner's ``compute`` mirrors ``interface_one_to_one_single.hpp:40-67``,
``interface_one_to_one_bulk.hpp:47-100`` and ``interface_one_to_many_bulk.hpp:47-92`` and runs on
the GPU through the C ABI (pairwise_aligner_amd.device).
"""
from __future__ import annotations

import enum
from dataclasses import dataclass, field
from typing import Any, List, Optional, Sequence, Union

from . import device as _dev


class end_gap(enum.Enum):       # end_gap_policy.hpp:20-24
    penalised = False
    free = True


@dataclass(frozen=True)
class leading_end_gap:          # end_gap_policy.hpp:26-30
    first_column: end_gap = end_gap.penalised
    first_row: end_gap = end_gap.penalised


@dataclass(frozen=True)
class trailing_end_gap:         # end_gap_policy.hpp:32-36
    last_column: end_gap = end_gap.penalised
    last_row: end_gap = end_gap.penalised


@dataclass(frozen=True)
class _Rule:
    category: str               # "score", "gap", "method"
    params: dict
    predecessor: Optional["_Rule"] = None

    def chain(self):
        r = self
        while r is not None:
            yield r
            r = r.predecessor


class _Initial:
    """initial.hpp:74: the rule a chain starts from (passing it is the same as no predecessor)."""

    def __repr__(self):
        return "cfg.initial"


initial = _Initial()


def _split(args):
    if args and isinstance(args[0], _Rule):
        return args[0], args[1:]
    if args and isinstance(args[0], _Initial):
        return None, args[1:]
    return None, args


def _as_table(matrix):
    """Accepts a reference-shaped table ((symbol, row), ...) and returns (symbols, flat values)."""
    symbols = "".join(s for s, _ in matrix)
    values = tuple(int(v) for _, row in matrix for v in row)
    if len(values) != len(symbols) ** 2:
        raise ValueError("substitution matrix must be square")
    return symbols.encode(), values


def score_model_unitary(*args, lanes=_dev.LANES_AUTO):
    """score_model_unitary.hpp:52-59 (scalar int32) and its configuration rule (:209)."""
    pred, (match, mismatch) = _split(args)
    return _Rule("score", dict(kind=_dev.SCORE_UNITARY, match=int(match), mismatch=int(mismatch), lanes=lanes), pred)


def score_model_unitary_simd(*args):
    """score_model_unitary_simd.hpp:51-58; runs in packed int16 lanes, exact by int32 recompute."""
    return score_model_unitary(*args, lanes=_dev.LANES_INT16)


def score_model_unitary_simd_saturated(*args):
    """score_model_unitary_simd_saturated.hpp:47-191 (int8 tiles in the reference); here the
    same scores come from int16 lanes with int32 recompute."""
    return score_model_unitary(*args, lanes=_dev.LANES_INT16)


def score_model_matrix(*args, lanes=_dev.LANES_AUTO):
    """score_model_matrix.hpp:52-64: diag + M[rank(a)][rank(b)]."""
    pred, (matrix,) = _split(args)
    symbols, values = _as_table(matrix)
    return _Rule("score", dict(kind=_dev.SCORE_MATRIX, symbols=symbols, matrix=values, lanes=lanes), pred)


def score_model_matrix_simd_NxN(*args):
    """score_model_matrix_simd_NxN.hpp:26-147."""
    return score_model_matrix(*args, lanes=_dev.LANES_INT16)


def score_model_matrix_simd_1xN(*args):
    """score_model_matrix_simd_1xN.hpp:37-183 (one query against many)."""
    return score_model_matrix(*args, lanes=_dev.LANES_INT16)


def gap_model_affine(*args):
    """gap_model_affine.hpp:37-62: a gap of length k scores open + k * extension."""
    pred, (open_, extension) = _split(args)
    return _Rule("gap", dict(open=int(open_), extension=int(extension)), pred)


def method_global(*args):
    """method_global.hpp:35-53 with the leading/trailing end-gap policies."""
    pred, rest = _split(args)
    lead = next((a for a in rest if isinstance(a, leading_end_gap)), leading_end_gap())
    trail = next((a for a in rest if isinstance(a, trailing_end_gap)), trailing_end_gap())
    return _Rule("method", dict(local=False, leading=lead, trailing=trail), pred)


def method_local(*args):
    """method_local.hpp:41-49: all end gaps free, zero clamp, maximum over all cells."""
    pred, _ = _split(args)
    return _Rule("method", dict(local=True, leading=leading_end_gap(end_gap.free, end_gap.free),
                                trailing=trailing_end_gap(end_gap.free, end_gap.free)), pred)


def lane_policy(*args):
    """Extension rule: choose the lane policy (device.LANES_*) of the device path."""
    pred, (policy,) = _split(args)
    return _Rule("lanes", dict(policy=policy), pred)


INT32_MIN = -(2 ** 31)


class dp_cell(tuple):
    """One entry of the final DP column or row, (score, gap state), as the reference's affine cell
    (affine/affine_cell.hpp:25-73: a pair whose score() is the first member).  The column's gap
    state is the horizontal one right of the last column, the row's the vertical one below the last
    row; entry 0 of either holds go + ge."""

    __slots__ = ()

    def __new__(cls, score, gap):
        return tuple.__new__(cls, (int(score), int(gap)))

    def score(self) -> int:
        return self[0]

    @property
    def first(self) -> int:
        return self[0]

    @property
    def second(self) -> int:
        return self[1]


class aligner_result:
    """aligner_result.hpp:33-105: the two sequences, the score and, after the four-argument
    compute, the cells of the final DP column (H(i, n), i = 0..m) and row (H(m, j), j = 0..n),
    empty otherwise."""

    __slots__ = ("_s1", "_s2", "_score", "_flags", "_dp_column", "_dp_row")

    def __init__(self, s1, s2, score, flags=0, dp_column=(), dp_row=()):
        self._s1, self._s2, self._score, self._flags = s1, s2, int(score), int(flags)
        self._dp_column, self._dp_row = list(dp_column), list(dp_row)

    def dp_column(self) -> list:
        return self._dp_column

    def dp_row(self) -> list:
        return self._dp_row

    def score(self) -> int:
        return self._score

    def sequence1(self):
        return self._s1

    def sequence2(self):
        return self._s2

    def flags(self) -> int:
        return self._flags


def _as_bytes(s) -> bytes:
    if isinstance(s, (bytes, bytearray)):
        return bytes(s)
    if isinstance(s, str):
        return s.encode()
    return bytes(s)


@dataclass
class Aligner:
    config: _dev.DeviceConfig
    device: int = 0

    def _ctx(self):
        return _dev.default_context(self.device)

    def compute(self, seq1, seq2, first_dp_column=None, first_dp_row=None):
        """compute(s1, s2) -> one result; compute(bulk1, bulk2) -> list; compute(s1, bulk2) -> list
        (one against many).  With the initial DP column and row (interface_one_to_one_single.hpp:52-67,
        interface_one_to_one_bulk.hpp:73-100, interface_one_to_many_bulk.hpp:67-77) every result also
        carries its final DP column and row; like the reference, the boundary itself comes from the
        configuration."""
        single1 = isinstance(seq1, (str, bytes, bytearray))
        single2 = isinstance(seq2, (str, bytes, bytearray))
        if first_dp_column is not None or first_dp_row is not None:
            return self._compute_dp(seq1, seq2, single1, single2)
        if single1 and single2:
            return self.compute([seq1], [seq2])[0]
        if single1:
            # one against many: the device plans the query once (pa_score_one_to_many)
            q = _as_bytes(seq1)
            b2, o2 = _dev.pack_sequences([_as_bytes(s) for s in seq2])
            scores, flags = self._ctx().score_one_to_many(q, b2, o2, self.config)
            return [aligner_result(seq1, b, sc, fl) for b, sc, fl in zip(seq2, scores, flags)]
        s1 = [_as_bytes(s) for s in seq1]
        s2 = [_as_bytes(s) for s in seq2]
        if len(s1) != len(s2):
            raise ValueError("bulk compute needs the same number of sequences on both sides")
        scores, flags = self._ctx().score(s1, s2, self.config)
        return [aligner_result(a, b, sc, fl) for a, b, sc, fl in zip(seq1, seq2, scores, flags)]


    def _compute_dp(self, seq1, seq2, single1, single2):
        if single1 and single2:
            return self._compute_dp([seq1], [seq2], False, False)[0]
        firsts = [seq1] * len(seq2) if single1 else list(seq1)
        s1 = [_as_bytes(s) for s in firsts]
        s2 = [_as_bytes(s) for s in seq2]
        if len(s1) != len(s2):
            raise ValueError("bulk compute needs the same number of sequences on both sides")
        b1, o1 = _dev.pack_sequences(s1)
        b2, o2 = _dev.pack_sequences(s2)
        scores, flags, col, row, col_gap, row_gap = self._ctx().score_packed_dp(b1, o1, b2, o2, self.config,
                                                                                gaps=True)
        out = []
        for p, (a, b) in enumerate(zip(firsts, seq2)):
            c = slice(int(o1[p]) + p, int(o1[p]) + p + len(s1[p]) + 1)
            r = slice(int(o2[p]) + p, int(o2[p]) + p + len(s2[p]) + 1)
            out.append(aligner_result(a, b, scores[p], flags[p], map(dp_cell, col[c], col_gap[c]),
                                      map(dp_cell, row[r], row_gap[r])))
        return out


def configure_aligner(rule: _Rule, device: int = 0) -> Aligner:
    """configure_aligner.hpp:194-206."""
    rules = list(rule.chain())
    score = [r for r in rules if r.category == "score"]
    gap = [r for r in rules if r.category == "gap"]
    method = [r for r in rules if r.category == "method"]
    lanes = [r for r in rules if r.category == "lanes"]
    if len(score) != 1:
        raise TypeError("configure_aligner needs exactly one score model rule")
    if len(gap) != 1:
        raise TypeError("configure_aligner needs exactly one gap model rule")
    if len(method) > 1:
        raise TypeError("configure_aligner takes at most one method rule")
    sp, gp = score[0].params, gap[0].params
    mp = method[0].params if method else dict(local=False, leading=leading_end_gap(), trailing=trailing_end_gap())
    policy = lanes[0].params["policy"] if lanes else sp["lanes"]
    kw = dict(method=_dev.METHOD_LOCAL if mp["local"] else _dev.METHOD_GLOBAL,
              first_column_free=mp["leading"].first_column.value, first_row_free=mp["leading"].first_row.value,
              last_column_free=mp["trailing"].last_column.value, last_row_free=mp["trailing"].last_row.value,
              gap_open=gp["open"], gap_extension=gp["extension"], lane_policy=policy)
    if sp["kind"] == _dev.SCORE_UNITARY:
        dc = _dev.DeviceConfig(score_kind=_dev.SCORE_UNITARY, match=sp["match"], mismatch=sp["mismatch"], **kw)
    else:
        dc = _dev.DeviceConfig(score_kind=_dev.SCORE_MATRIX, symbols=sp["symbols"], matrix=sp["matrix"], **kw)
    return Aligner(dc, device)
