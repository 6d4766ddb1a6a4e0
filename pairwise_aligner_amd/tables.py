"""Substitution tables for the Python mirror of the configuration API.

The BLOSUM62 values are read from the one copy kept in
``include/pairwise_aligner/score_model/substitution_matrix.hpp`` (NCBI text block), and
ordered like the reference's ``blosum62_standard`` / ``blosum62_extended``
(include/pairwise_aligner/score_model/substitution_matrix.hpp:28-87 of rrahn/pairwise_aligner).
A table is a tuple of (symbol, row) pairs, the same shape as the reference's
``std::array<std::pair<char, std::array<score_t, N>>, N>``.
"""
from __future__ import annotations

import functools
import os
import re

_HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "include", "pairwise_aligner", "score_model", "substitution_matrix.hpp")

STANDARD_ORDER = "ACDEFGHIKLMNPQRSTVWXY"
EXTENDED_ORDER = "ABCDEFGHIJKLMNOPQRSTUVWXYZ*"


@functools.lru_cache(maxsize=1)
def _ncbi():
    text = open(_HEADER).read()
    blocks = re.findall(r'R"BLOSUM\((.*?)\)BLOSUM"', text, re.S)
    lines = [l for l in blocks[0].strip("\n").split("\n") if l.strip()]
    header = lines[0].split()
    values = {}
    for line in lines[1:]:
        parts = line.split()
        values[parts[0]] = dict(zip(header, map(int, parts[1:])))
    jparts = blocks[1].split()
    jrow = dict(zip(header, map(int, jparts[1:1 + len(header)])))
    jself = int(jparts[1 + len(header)])
    return values, jrow, jself


def blosum62_value(a: str, b: str) -> int:
    values, jrow, jself = _ncbi()
    a = "X" if a in "OU" else a
    b = "X" if b in "OU" else b
    if a == "J" and b == "J":
        return jself
    if a == "J":
        return jrow[b]
    if b == "J":
        return jrow[a]
    return values[a][b]


def _table(order: str):
    return tuple((s, tuple(blosum62_value(s, t) for t in order)) for s in order)


@functools.lru_cache(maxsize=None)
def blosum62_standard():
    return _table(STANDARD_ORDER)


@functools.lru_cache(maxsize=None)
def blosum62_extended():
    return _table(EXTENDED_ORDER)
