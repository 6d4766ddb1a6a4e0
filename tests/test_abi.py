"""CPU-side checks of the device library: it loads, exports every symbol include/pa_device.h
declares, and the configuration mirror validates rule chains like the reference."""
import ctypes
import os
import re

import pytest

from pairwise_aligner_amd import cfg, device, fastx, tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="pa_device.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(pa_\w+)\s*\(", text)))


def test_header_declares_the_binding_list():
    assert declared_symbols() == sorted(device.EXPORTS)
    assert declared_symbols("pa_fastx.h") == sorted(fastx.EXPORTS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(device.LIB_PATH)
    for name in declared_symbols() + declared_symbols("pa_fastx.h"):
        assert hasattr(lib, name), name
    assert lib.pa_abi_version() == 7


def test_configure_aligner_requires_score_and_gap_models():
    with pytest.raises(TypeError):
        cfg.configure_aligner(cfg.method_global(cfg.gap_model_affine(-10, -1)))
    with pytest.raises(TypeError):
        cfg.configure_aligner(cfg.score_model_unitary(4, -5))


def test_configure_aligner_translation():
    a = cfg.configure_aligner(cfg.method_global(cfg.gap_model_affine(cfg.score_model_unitary(4, -5), -10, -1),
                                                cfg.leading_end_gap(first_row=cfg.end_gap.free),
                                                cfg.trailing_end_gap(last_row=cfg.end_gap.free)))
    c = a.config
    assert (c.method, c.first_column_free, c.first_row_free, c.last_column_free, c.last_row_free) == \
        (device.METHOD_GLOBAL, False, True, False, True)
    assert (c.gap_open, c.gap_extension, c.match, c.mismatch) == (-10, -1, 4, -5)
    loc = cfg.configure_aligner(cfg.score_model_matrix(cfg.method_local(cfg.gap_model_affine(-11, -1)),
                                                       tables.blosum62_extended())).config
    assert loc.method == device.METHOD_LOCAL and loc.first_column_free and loc.last_row_free
    assert loc.symbols == b"ABCDEFGHIJKLMNOPQRSTUVWXYZ*" and len(loc.matrix) == 27 * 27


def test_chains_from_initial_and_from_the_method():
    """cfg.initial (initial.hpp:74) and methods without a predecessor (method_global.hpp:137-140,
    method_local.hpp:129-132) configure like the chains with the rule in the middle."""
    a = cfg.configure_aligner(cfg.score_model_unitary(cfg.gap_model_affine(cfg.method_local(), -10, -1), 4, -5)).config
    b = cfg.configure_aligner(cfg.method_local(cfg.gap_model_affine(cfg.score_model_unitary(cfg.initial, 4, -5),
                                                                    -10, -1))).config
    assert a == b and a.method == device.METHOD_LOCAL
    g = cfg.configure_aligner(cfg.score_model_unitary(cfg.gap_model_affine(
        cfg.method_global(cfg.leading_end_gap(first_column=cfg.end_gap.free), cfg.trailing_end_gap()), -2, -1),
        1, -1)).config
    assert (g.method, g.first_column_free, g.first_row_free, g.gap_open) == (device.METHOD_GLOBAL, True, False, -2)


def test_no_gpu_means_a_loud_error():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    with pytest.raises(device.PaError):
        device.Context(0)
