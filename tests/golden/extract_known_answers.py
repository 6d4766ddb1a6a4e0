"""Writes tests/golden/known_answers.json from the reference's scalar known-answer tests.

Run once in a container where the reference checkout is mounted:

    python tests/golden/extract_known_answers.py /root/reference

It reads the DEFINE_TEST_VALUES blocks of test/unit/affine/*_scalar*_test.cpp (inputs and
expected scores only) and the two configuration smoke tests, and stores them as data:
configuration parameters, the two sequences, the expected score and the file:line each
vector came from.  Nothing from the reference is executed.
"""
from __future__ import annotations

import json
import os
import re
import sys

INT32_MIN = -(2 ** 31)

SCALAR_FILES = [
    "test/unit/affine/global_standard_affine_scalar_test.cpp",
    "test/unit/affine/global_overlap_affine_scalar_test.cpp",
    "test/unit/affine/global_semi_first_affine_scalar_test.cpp",
    "test/unit/affine/global_semi_second_affine_scalar_test.cpp",
    "test/unit/affine/global_standard_affine_scalar_matrix_test.cpp",
    "test/unit/affine/local_affine_scalar_test.cpp",
]


def _end_gaps(text: str, kind: str):
    """Parse leading_end_gap{...} / trailing_end_gap{...} into two booleans (free?)."""
    names = ("first_column", "first_row") if kind == "leading" else ("last_column", "last_row")
    m = re.search(kind + r"_end_gap\{([^}]*)\}", text)
    flags = {names[0]: False, names[1]: False}
    if not m:
        return flags
    body = m.group(1).strip()
    if not body:
        return flags
    if "." in body.split("=")[0]:
        for key, val in re.findall(r"\.(\w+)\s*=\s*[\w:]*end_gap::(\w+)", body):
            flags[key] = val == "free"
    else:
        vals = re.findall(r"end_gap::(\w+)", body)
        for key, val in zip(names, vals):
            flags[key] = val == "free"
    return flags


def _method(text: str):
    local = "method_local" in text
    go, ge = map(int, re.search(r"gap_model_affine\(\s*(-?\d+)\s*,\s*(-?\d+)\s*\)", text).groups())
    cfg = {"mode": "local" if local else "global", "gap_open": go, "gap_extension": ge}
    if local:
        cfg.update(first_column_free=True, first_row_free=True, last_column_free=True, last_row_free=True)
    else:
        lead = _end_gaps(text, "leading")
        trail = _end_gaps(text, "trailing")
        cfg.update(first_column_free=lead["first_column"], first_row_free=lead["first_row"],
                   last_column_free=trail["last_column"], last_row_free=trail["last_row"])
    return cfg


def _line_of(src: str, pos: int) -> int:
    return src.count("\n", 0, pos) + 1


def extract_file(root: str, rel: str):
    src = open(os.path.join(root, rel)).read()
    base = re.search(r"inline constexpr auto base_config\s*=(.*?);", src, re.S).group(1)
    base_cfg = _method(base)
    out = []
    for m in re.finditer(r"DEFINE_TEST_VALUES\((\w+),(.*?)\n\)", src, re.S):
        name, body = m.group(1), m.group(2)
        conf = re.search(r"\.configurator\s*=(.*?),\s*\n\s*\.sequence1", body, re.S).group(1)
        s1 = re.search(r'\.sequence1\s*=\s*"([^"]*)"sv', body).group(1)
        s2 = re.search(r'\.sequence2\s*=\s*"([^"]*)"sv', body).group(1)
        exp_txt = re.search(r"\.expected_score\s*=\s*(.*?)\s*$", body, re.S).group(1).strip()
        expected = INT32_MIN if "lowest()" in exp_txt else int(exp_txt)
        cfg = dict(base_cfg) if "base_config" in conf else _method(conf)
        if "score_model_matrix" in conf:
            cfg["score"] = {"kind": "matrix",
                            "matrix": re.search(r"(blosum62_\w+)", conf).group(1)}
        else:
            mm = re.search(r"score_model_unitary\((.*)\)\s*$", conf.strip(), re.S).group(1)
            nums = re.findall(r"(-?\d+)\s*$|,\s*(-?\d+)", mm)
            ints = [int(x) for pair in nums for x in pair if x]
            cfg["score"] = {"kind": "unitary", "match": ints[-2], "mismatch": ints[-1]}
        out.append({"name": f"{os.path.basename(rel)[:-4]}::{name}",
                    "source": f"{rel}:{_line_of(src, m.start())}",
                    "config": cfg, "seq1": s1, "seq2": s2, "expected": expected})
    return out


def configuration_tests():
    """The two configuration smoke tests have no DEFINE_TEST_VALUES; their single vector
    is stated here with its location."""
    seq = "ACGTGACTGACACTACGACT"
    return [
        {"name": "configuration_test::prototype",
         "source": "test/unit/configuration/configuration_test.cpp:18-37",
         "config": {"mode": "global", "gap_open": -10, "gap_extension": -1,
                    "first_column_free": False, "first_row_free": True,
                    "last_column_free": False, "last_row_free": True,
                    "score": {"kind": "unitary", "match": 4, "mismatch": -5}},
         "seq1": seq, "seq2": seq, "expected": 80},
        {"name": "configure_aligner_saturated_test::prototype",
         "source": "test/unit/configuration/configure_aligner_saturated_test.cpp:18-42",
         "config": {"mode": "global", "gap_open": -10, "gap_extension": -1,
                    "first_column_free": False, "first_row_free": False,
                    "last_column_free": False, "last_row_free": False,
                    "score": {"kind": "unitary", "match": 4, "mismatch": -5}},
         "seq1": seq, "seq2": seq, "expected": 80},
    ]


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    vectors = []
    for rel in SCALAR_FILES:
        vectors.extend(extract_file(root, rel))
    vectors.extend(configuration_tests())
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")
    with open(dst, "w") as f:
        json.dump(vectors, f, indent=1)
    print(f"wrote {len(vectors)} vectors to {dst}")


if __name__ == "__main__":
    main()
