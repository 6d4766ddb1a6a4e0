"""FASTA / FASTQ records as the packed arrays the alignment entry points take (ctypes binding of
include/pa_fastx.h, implemented in libpa_device.so on the host; no GPU is needed).  The arrays
are views of the library's buffers (no copy), which live as long as any of them does.

    recs = fastx.read("reads.fastq")
    scores, flags = ctx.score_packed(recs.residues, recs.offsets, ...)

The reference reads no files; its one data file is the FASTQ fixture test/data/in.fastq
(test/data/datasources.cmake:6-8).
"""
from __future__ import annotations

import ctypes
import functools
import os
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import device

FASTA, FASTQ = 1, 2
UPPERCASE = 1

EXPORTS = ["pa_fastx_parse", "pa_fastx_read", "pa_fastx_get", "pa_fastx_destroy"]


class CView(ctypes.Structure):
    _fields_ = [
        ("format", ctypes.c_int32),
        ("n_records", ctypes.c_uint64),
        ("residues", ctypes.POINTER(ctypes.c_uint8)),
        ("offsets", ctypes.POINTER(ctypes.c_uint64)),
        ("names", ctypes.POINTER(ctypes.c_char)),
        ("name_offsets", ctypes.POINTER(ctypes.c_uint64)),
        ("qualities", ctypes.POINTER(ctypes.c_uint8)),
    ]


@functools.lru_cache(maxsize=1)
def _lib():
    if not os.path.exists(device.LIB_PATH):
        raise device.PaError(4, f"{device.LIB_PATH} is not built (python -m pairwise_aligner_amd.build)")
    lib = ctypes.CDLL(device.LIB_PATH)
    vp, c = ctypes.c_void_p, ctypes
    lib.pa_fastx_parse.argtypes = [c.c_char_p, c.c_uint64, c.c_uint32, c.POINTER(vp), c.c_char_p, c.c_size_t]
    lib.pa_fastx_parse.restype = c.c_int
    lib.pa_fastx_read.argtypes = [c.c_char_p, c.c_uint32, c.POINTER(vp), c.c_char_p, c.c_size_t]
    lib.pa_fastx_read.restype = c.c_int
    lib.pa_fastx_get.argtypes = [vp, c.POINTER(CView)]
    lib.pa_fastx_get.restype = c.c_int
    lib.pa_fastx_destroy.argtypes = [vp]
    lib.pa_fastx_destroy.restype = c.c_int
    return lib


@dataclass
class Records:
    """Parsed records: record r is residues[offsets[r]:offsets[r + 1]], its name
    name_bytes[name_offsets[r]:name_offsets[r + 1]]."""
    format: int
    residues: np.ndarray            # uint8
    offsets: np.ndarray             # uint64, n + 1 entries
    name_bytes: bytes
    name_offsets: np.ndarray        # uint64, n + 1 entries
    qualities: Optional[np.ndarray]  # uint8, FASTQ only (same offsets as residues)

    def __len__(self) -> int:
        return len(self.offsets) - 1

    @property
    def names(self) -> List[str]:
        o = self.name_offsets.tolist()
        return [self.name_bytes[o[r]:o[r + 1]].decode("latin-1") for r in range(len(self))]

    def sequence(self, r: int) -> bytes:
        return self.residues[self.offsets[r]:self.offsets[r + 1]].tobytes()

    def sequences(self) -> List[bytes]:
        return [self.sequence(r) for r in range(len(self))]


class _Handle:
    """Owns a pa_fastx handle; destroyed when the last array viewing its buffers is gone."""

    def __init__(self, handle: ctypes.c_void_p):
        self.handle = handle

    def __del__(self):
        if self.handle:
            try:
                _lib().pa_fastx_destroy(self.handle)
            except Exception:   # interpreter shutdown: the library may already be unloaded
                pass
            self.handle = None


def _view(owner: _Handle, ptr, ctype, count: int, dtype) -> np.ndarray:
    """numpy array over library memory, keeping `owner` alive (no copy)."""
    if count == 0:
        return np.zeros(0, dtype)
    buf = (ctype * count).from_address(ctypes.addressof(ptr.contents))
    buf._owner = owner
    return np.frombuffer(buf, dtype=dtype)


def _collect(handle: ctypes.c_void_p) -> Records:
    owner = _Handle(handle)
    v = CView()
    st = _lib().pa_fastx_get(handle, ctypes.byref(v))
    if st != device.PA_OK:
        raise device.PaError(st, "pa_fastx_get")
    n = v.n_records
    offsets = _view(owner, v.offsets, ctypes.c_uint64, n + 1, np.uint64)
    nbytes = int(offsets[-1])
    residues = _view(owner, v.residues, ctypes.c_uint8, nbytes, np.uint8)
    name_off = _view(owner, v.name_offsets, ctypes.c_uint64, n + 1, np.uint64)
    name_bytes = ctypes.string_at(v.names, int(name_off[-1])) if name_off[-1] else b""
    qualities = _view(owner, v.qualities, ctypes.c_uint8, nbytes, np.uint8) if v.format == FASTQ else None
    return Records(v.format, residues, offsets, name_bytes, name_off, qualities)


def parse(text: bytes, uppercase: bool = False) -> Records:
    """Parses FASTA or FASTQ text (the format is taken from the first non-blank byte)."""
    handle = ctypes.c_void_p()
    err = ctypes.create_string_buffer(256)
    st = _lib().pa_fastx_parse(text, len(text), UPPERCASE if uppercase else 0, ctypes.byref(handle), err, 256)
    if st != device.PA_OK:
        raise device.PaError(st, err.value.decode())
    return _collect(handle)


def read(path: str, uppercase: bool = False) -> Records:
    """Reads and parses a FASTA or FASTQ file."""
    handle = ctypes.c_void_p()
    err = ctypes.create_string_buffer(256)
    st = _lib().pa_fastx_read(os.fsencode(path), UPPERCASE if uppercase else 0, ctypes.byref(handle), err, 256)
    if st != device.PA_OK:
        raise device.PaError(st, err.value.decode())
    return _collect(handle)
