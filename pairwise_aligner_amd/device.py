"""ctypes binding of the C ABI in include/pa_device.h (libpa_device.so, built for gfx950).

There is no CPU fallback: if the library is missing or no GPU is visible, every entry point
raises.  Sequences are passed as concatenated bytes plus uint64 offsets.
"""
from __future__ import annotations

import ctypes
import functools
import os
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpa_device.so")

PA_OK = 0
STATUS_NAMES = {0: "PA_OK", 1: "PA_ERR_INVALID_ARGUMENT", 2: "PA_ERR_INVALID_SYMBOL", 3: "PA_ERR_UNSUPPORTED",
                4: "PA_ERR_DEVICE", 5: "PA_ERR_OUT_OF_MEMORY", 6: "PA_ERR_RANGE"}
METHOD_GLOBAL, METHOD_LOCAL = 0, 1
SCORE_UNITARY, SCORE_MATRIX = 0, 1
LANES_AUTO, LANES_LEARNED, LANES_INT16, LANES_INT32 = 0, 1, 16, 32
PAIR_INT16, PAIR_INT32, PAIR_RECOMPUTED, PAIR_EMPTY = 0x01, 0x02, 0x04, 0x08

# Every symbol include/pa_device.h declares.
EXPORTS = ["pa_abi_version", "pa_ctx_create", "pa_ctx_destroy", "pa_ctx_last_error", "pa_score_batch",
           "pa_batch_create", "pa_batch_run", "pa_batch_results", "pa_batch_stats_get", "pa_batch_time",
           "pa_batch_destroy", "pa_model_create", "pa_model_destroy", "pa_model_param_count",
           "pa_model_set_params", "pa_model_get_params", "pa_features", "pa_model_forward", "pa_model_loss_grad",
           "pa_model_train_step", "pa_dp_labels", "pa_tune", "pa_score_one_to_many", "pa_batch_create_one_to_many",
           "pa_score_batch_dp", "pa_score_batch_dp_cells", "pa_score_batch_noalloc", "pa_model_set_margin"]
N_FEATURES, N_OUTPUTS, N_HIDDEN = 8, 2, 64


class PaError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class CTuning(ctypes.Structure):
    _fields_ = [
        ("cell_cost", ctypes.c_float),
        ("step_cost", ctypes.c_float),
        ("rows_mask", ctypes.c_uint32),
        ("wavefront_min_rows", ctypes.c_uint32),
        ("streams", ctypes.c_uint32),
        ("eager", ctypes.c_uint32),
        ("cost_order", ctypes.c_uint32),
        ("persistent", ctypes.c_uint32),
        ("wavefront_tail", ctypes.c_float),
    ]


@dataclass(frozen=True)
class Tuning:
    """Planner and launch parameters (pa_tuning); 0 fields take the library defaults."""
    cell_cost: float = 0.0
    step_cost: float = 0.0
    rows_mask: int = 0
    wavefront_min_rows: int = 0
    streams: int = 0
    eager: int = 0
    cost_order: int = 0   # learned policy: order waves by the predicted cost (1) or the geometric model (0)
    # striped launch form: 0 one block per 16 tasks, the int16 local launches that differ only in R
    # fused into one launch with one work queue; 1 resident blocks taking tasks from a work queue
    # per launch; 2 as 1 with the int16 local launches fused; 3 no fusion; 4 as 0 with the global
    # method's int16 unitary-score launches also fused (two launches by rows per lane)
    persistent: int = 0
    # long pairs go to the wavefront kernel only where their striped sweep would exceed this
    # multiple of the batch's striped work per resident lane group (< 0: every long pair)
    wavefront_tail: float = 0.0

    def to_c(self) -> CTuning:
        return CTuning(self.cell_cost, self.step_cost, self.rows_mask, self.wavefront_min_rows, self.streams,
                       self.eager, self.cost_order, self.persistent, self.wavefront_tail)

    @staticmethod
    def from_c(c: CTuning) -> "Tuning":
        return Tuning(round(c.cell_cost, 4), round(c.step_cost, 4), c.rows_mask, c.wavefront_min_rows, c.streams,
                      c.eager, c.cost_order, c.persistent, round(c.wavefront_tail, 4))


class CConfig(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("first_column_free", ctypes.c_int32),
        ("first_row_free", ctypes.c_int32),
        ("last_column_free", ctypes.c_int32),
        ("last_row_free", ctypes.c_int32),
        ("gap_open", ctypes.c_int32),
        ("gap_extension", ctypes.c_int32),
        ("score_kind", ctypes.c_int32),
        ("match", ctypes.c_int32),
        ("mismatch", ctypes.c_int32),
        ("matrix_dim", ctypes.c_int32),
        ("matrix", ctypes.POINTER(ctypes.c_int32)),
        ("matrix_symbols", ctypes.POINTER(ctypes.c_uint8)),
        ("lane_policy", ctypes.c_int32),
        ("model", ctypes.c_void_p),
        ("tuning", ctypes.POINTER(CTuning)),
    ]


class CStats(ctypes.Structure):
    _fields_ = [("n_pairs", ctypes.c_uint64), ("cells", ctypes.c_uint64), ("padded_cells", ctypes.c_uint64),
                ("n_int16", ctypes.c_uint64), ("n_int32", ctypes.c_uint64), ("n_recomputed", ctypes.c_uint64),
                ("n_launches", ctypes.c_uint64), ("plan_ms", ctypes.c_double), ("n_profiled", ctypes.c_uint64)]


@dataclass(frozen=True)
class DeviceConfig:
    """Flat form of an aligner configuration, one-to-one with pa_config."""
    method: int = METHOD_GLOBAL
    first_column_free: bool = False
    first_row_free: bool = False
    last_column_free: bool = False
    last_row_free: bool = False
    gap_open: int = -10
    gap_extension: int = -1
    score_kind: int = SCORE_UNITARY
    match: int = 4
    mismatch: int = -5
    symbols: bytes = b""
    matrix: Optional[tuple] = None
    lane_policy: int = LANES_AUTO
    model: Optional["Model"] = None     # PA_LANES_LEARNED
    tuning: Optional[Tuning] = None     # planner/launch parameters (None: defaults)


@functools.lru_cache(maxsize=1)
def lib() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"device library missing: {LIB_PATH}; run __graft_entry__.build() (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    L.pa_abi_version.restype = i32
    L.pa_ctx_create.argtypes = [ctypes.POINTER(vp), i32]
    L.pa_ctx_destroy.argtypes = [vp]
    L.pa_ctx_last_error.argtypes = [vp]
    L.pa_ctx_last_error.restype = ctypes.c_char_p
    L.pa_score_batch.argtypes = [vp, ctypes.POINTER(CConfig), vp, vp, vp, vp, u64, vp, vp, vp]
    L.pa_batch_create.argtypes = [vp, ctypes.POINTER(CConfig), vp, vp, vp, vp, u64, vp, ctypes.POINTER(vp)]
    L.pa_score_one_to_many.argtypes = [vp, ctypes.POINTER(CConfig), vp, u64, vp, vp, u64, vp, vp, vp]
    L.pa_batch_create_one_to_many.argtypes = [vp, ctypes.POINTER(CConfig), vp, u64, vp, vp, u64, vp,
                                              ctypes.POINTER(vp)]
    L.pa_score_batch_noalloc.argtypes = [vp, ctypes.POINTER(CConfig), vp, vp, vp, vp, u64, vp, vp, vp]
    L.pa_score_batch_dp.argtypes = [vp, ctypes.POINTER(CConfig), vp, vp, vp, vp, u64, vp, vp, vp, vp, vp]
    L.pa_score_batch_dp_cells.argtypes = [vp, ctypes.POINTER(CConfig), vp, vp, vp, vp, u64, vp, vp, vp, vp, vp, vp,
                                          vp]
    L.pa_batch_run.argtypes = [vp, vp]
    L.pa_batch_results.argtypes = [vp, vp, vp, vp]
    L.pa_batch_stats_get.argtypes = [vp, ctypes.POINTER(CStats)]
    L.pa_batch_time.argtypes = [vp, vp, i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.pa_batch_destroy.argtypes = [vp, vp]
    L.pa_model_create.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(vp)]
    L.pa_model_destroy.argtypes = [vp]
    L.pa_model_param_count.restype = i32
    L.pa_model_set_params.argtypes = [vp, vp, vp]
    L.pa_model_set_margin.argtypes = [vp, ctypes.c_float]
    L.pa_model_get_params.argtypes = [vp, vp, vp]
    L.pa_features.argtypes = [vp, ctypes.POINTER(CConfig), vp, vp, vp, vp, u64, vp]
    L.pa_model_forward.argtypes = [vp, vp, vp, u64, vp]
    L.pa_model_loss_grad.argtypes = [vp, vp, vp, vp, vp, u64, vp, vp]
    L.pa_model_train_step.argtypes = [vp, vp, vp, vp, vp, u64, ctypes.c_float, vp]
    L.pa_dp_labels.argtypes = [vp, ctypes.POINTER(CConfig), vp, vp, vp, vp, u64, vp, vp]
    L.pa_tune.argtypes = [vp, ctypes.POINTER(CConfig), vp, vp, vp, vp, u64, vp, u64, i32, ctypes.POINTER(CTuning)]
    for name in ["pa_ctx_create", "pa_ctx_destroy", "pa_score_batch", "pa_batch_create", "pa_batch_run",
                 "pa_batch_results", "pa_batch_stats_get", "pa_batch_time", "pa_batch_destroy", "pa_model_create",
                 "pa_model_destroy", "pa_model_set_params", "pa_model_get_params", "pa_features",
                 "pa_model_forward", "pa_model_loss_grad", "pa_model_train_step", "pa_dp_labels", "pa_tune",
                 "pa_score_one_to_many", "pa_batch_create_one_to_many", "pa_score_batch_dp",
                 "pa_score_batch_dp_cells", "pa_score_batch_noalloc", "pa_model_set_margin"]:
        getattr(L, name).restype = i32
    return L


def _residues(b, name: str) -> np.ndarray:
    if isinstance(b, (bytes, bytearray, memoryview)):
        b = np.frombuffer(bytes(b), dtype=np.uint8)
    b = np.asarray(b)
    if b.dtype != np.uint8:
        raise TypeError(f"{name}: residues must be uint8 (got {b.dtype})")
    b = np.ascontiguousarray(b)
    return b if b.size else np.zeros(1, dtype=np.uint8)


def _offsets(o, buf: np.ndarray, name: str) -> np.ndarray:
    o = np.asarray(o)
    if o.ndim != 1 or o.size < 1 or not np.issubdtype(o.dtype, np.integer):
        raise TypeError(f"{name}: offsets must be a non-empty 1-d integer array")
    if o.dtype != np.uint64:
        if np.any(o < 0):
            raise ValueError(f"{name}: negative offset")
        o = o.astype(np.uint64)
    o = np.ascontiguousarray(o)
    if o.size > 1 and np.any(o[1:] < o[:-1]):
        raise ValueError(f"{name}: offsets decrease")
    if int(o[-1]) > buf.size:
        raise ValueError(f"{name}: last offset {int(o[-1])} beyond the {buf.size} residues")
    return o


def checked_batch(b1, o1, b2, o2, pair_method=None, one_query=False):
    """Inputs of a batch as the C ABI reads them: contiguous uint8 residues, uint64 offsets that
    stay inside their buffers, equally many pairs on both sides (one_query: o1 holds the query's
    two offsets), a uint8 method per pair.  Raises instead of letting the library read out of
    bounds."""
    b1, b2 = _residues(b1, "seq1"), _residues(b2, "seq2")
    o1, o2 = _offsets(o1, b1, "off1"), _offsets(o2, b2, "off2")
    n = len(o2) - 1
    if not one_query and len(o1) != len(o2):
        raise ValueError(f"{len(o1) - 1} first sequences against {n} second sequences")
    pm = None
    if pair_method is not None:
        pm = np.ascontiguousarray(pair_method, dtype=np.uint8)
        if pm.shape != (n,):
            raise ValueError(f"pair_method holds {pm.size} entries for {n} pairs")
    return b1, o1, b2, o2, pm


def pack_sequences(seqs: Sequence[bytes]):
    lens = np.fromiter((len(s) for s in seqs), dtype=np.uint64, count=len(seqs))
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    joined = b"".join(seqs)
    buf = np.frombuffer(joined, dtype=np.uint8).copy() if joined else np.zeros(1, dtype=np.uint8)
    return buf, off


class _CConfigHolder:
    """Keeps the ctypes arrays a CConfig points to alive."""

    def __init__(self, cfg: DeviceConfig):
        self.c = CConfig(cfg.method, int(cfg.first_column_free), int(cfg.first_row_free),
                         int(cfg.last_column_free), int(cfg.last_row_free), cfg.gap_open, cfg.gap_extension,
                         cfg.score_kind, cfg.match, cfg.mismatch, 0, None, None, cfg.lane_policy,
                         cfg.model.handle if cfg.model is not None else None)
        if cfg.tuning is not None:
            self.tun = cfg.tuning.to_c()
            self.c.tuning = ctypes.pointer(self.tun)
        if cfg.score_kind == SCORE_MATRIX:
            self.mat = (ctypes.c_int32 * len(cfg.matrix))(*cfg.matrix)
            self.sym = (ctypes.c_uint8 * len(cfg.symbols))(*cfg.symbols)
            self.c.matrix_dim = len(cfg.symbols)
            self.c.matrix = self.mat
            self.c.matrix_symbols = self.sym


class Context:
    """One HIP device + stream (pa_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = lib()
        h = ctypes.c_void_p()
        st = self._lib.pa_ctx_create(ctypes.byref(h), device)
        if st != PA_OK:
            raise PaError(st, f"cannot create a context on device {device}")
        self.handle = h

    def error(self) -> str:
        return self._lib.pa_ctx_last_error(self.handle).decode()

    def check(self, st: int):
        if st != PA_OK:
            raise PaError(st, self.error())

    def close(self):
        if self.handle:
            self._lib.pa_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def score(self, seqs1: Sequence[bytes], seqs2: Sequence[bytes], cfg: DeviceConfig, pair_method=None):
        b1, o1 = pack_sequences(seqs1)
        b2, o2 = pack_sequences(seqs2)
        return self.score_packed(b1, o1, b2, o2, cfg, pair_method)

    def score_packed(self, b1, o1, b2, o2, cfg: DeviceConfig, pair_method=None):
        b1, o1, b2, o2, pm = checked_batch(b1, o1, b2, o2, pair_method)
        n = len(o1) - 1
        holder = _CConfigHolder(cfg)
        scores = np.zeros(max(n, 1), dtype=np.int32)
        flags = np.zeros(max(n, 1), dtype=np.uint8)
        st = self._lib.pa_score_batch(self.handle, ctypes.byref(holder.c), b1.ctypes.data, o1.ctypes.data,
                                      b2.ctypes.data, o2.ctypes.data, n, None if pm is None else pm.ctypes.data,
                                      scores.ctypes.data, flags.ctypes.data)
        self.check(st)
        return scores[:n], flags[:n]

    def score_packed_noalloc(self, b1, o1, b2, o2, cfg: DeviceConfig, pair_method=None):
        """As score_packed, through pa_score_batch_noalloc: no device memory is allocated; every
        device buffer is a slice of the library's 64 MiB per-device scratch."""
        b1, o1, b2, o2, pm = checked_batch(b1, o1, b2, o2, pair_method)
        n = len(o1) - 1
        holder = _CConfigHolder(cfg)
        scores = np.zeros(max(n, 1), dtype=np.int32)
        flags = np.zeros(max(n, 1), dtype=np.uint8)
        st = self._lib.pa_score_batch_noalloc(self.handle, ctypes.byref(holder.c), b1.ctypes.data, o1.ctypes.data,
                                              b2.ctypes.data, o2.ctypes.data, n,
                                              None if pm is None else pm.ctypes.data, scores.ctypes.data,
                                              flags.ctypes.data)
        self.check(st)
        return scores[:n], flags[:n]

    def score_packed_dp(self, b1, o1, b2, o2, cfg: DeviceConfig, pair_method=None, gaps: bool = False):
        """Scores, flags and the final DP column and row of every pair (pa_score_batch_dp):
        dp_column[(o1[p]-o1[0]) + p + i] = H(i, len2), dp_row[(o2[p]-o2[0]) + p + j] = H(len1, j).
        With gaps (pa_score_batch_dp_cells), also the gap-state halves of the same cells in the same
        layout: returns (scores, flags, dp_column, dp_row, dp_column_gap, dp_row_gap)."""
        b1, o1, b2, o2, pm = checked_batch(b1, o1, b2, o2, pair_method)
        n = len(o1) - 1
        holder = _CConfigHolder(cfg)
        scores = np.zeros(max(n, 1), dtype=np.int32)
        flags = np.zeros(max(n, 1), dtype=np.uint8)
        n1 = int(o1[-1] - o1[0]) + n
        n2 = int(o2[-1] - o2[0]) + n
        dp_column = np.zeros(max(n1, 1), dtype=np.int32)
        dp_row = np.zeros(max(n2, 1), dtype=np.int32)
        if not gaps:
            st = self._lib.pa_score_batch_dp(self.handle, ctypes.byref(holder.c), b1.ctypes.data, o1.ctypes.data,
                                             b2.ctypes.data, o2.ctypes.data, n, None if pm is None else pm.ctypes.data,
                                             scores.ctypes.data, flags.ctypes.data, dp_column.ctypes.data,
                                             dp_row.ctypes.data)
            self.check(st)
            return scores[:n], flags[:n], dp_column[:n1], dp_row[:n2]
        column_gap = np.zeros_like(dp_column)
        row_gap = np.zeros_like(dp_row)
        st = self._lib.pa_score_batch_dp_cells(self.handle, ctypes.byref(holder.c), b1.ctypes.data, o1.ctypes.data,
                                               b2.ctypes.data, o2.ctypes.data, n,
                                               None if pm is None else pm.ctypes.data, scores.ctypes.data,
                                               flags.ctypes.data, dp_column.ctypes.data, dp_row.ctypes.data,
                                               column_gap.ctypes.data, row_gap.ctypes.data)
        self.check(st)
        return scores[:n], flags[:n], dp_column[:n1], dp_row[:n2], column_gap[:n1], row_gap[:n2]

    def score_one_to_many(self, query, b2, o2, cfg: DeviceConfig, pair_method=None):
        """Scores of (query, target p) for every packed target (pa_score_one_to_many)."""
        q = np.frombuffer(bytes(query), dtype=np.uint8).copy() if len(query) else np.zeros(1, dtype=np.uint8)
        _, _, b2, o2, pm = checked_batch(q, np.array([0, len(query)], dtype=np.uint64), b2, o2, pair_method,
                                         one_query=True)
        n = len(o2) - 1
        holder = _CConfigHolder(cfg)
        scores = np.zeros(max(n, 1), dtype=np.int32)
        flags = np.zeros(max(n, 1), dtype=np.uint8)
        st = self._lib.pa_score_one_to_many(self.handle, ctypes.byref(holder.c), q.ctypes.data, len(query),
                                            b2.ctypes.data, o2.ctypes.data, n,
                                            None if pm is None else pm.ctypes.data, scores.ctypes.data,
                                            flags.ctypes.data)
        self.check(st)
        return scores[:n], flags[:n]


class Batch:
    """A device-resident batch (pa_batch): planned and uploaded once, run many times."""

    def __init__(self, ctx: Context, b1, o1, b2, o2, cfg: DeviceConfig, pair_method=None):
        self.ctx = ctx
        self.handle = None
        b1, o1, b2, o2, pm = checked_batch(b1, o1, b2, o2, pair_method)
        self.n = len(o1) - 1
        self._keep = (b1, o1, b2, o2)
        holder = _CConfigHolder(cfg)
        h = ctypes.c_void_p()
        st = ctx._lib.pa_batch_create(ctx.handle, ctypes.byref(holder.c), b1.ctypes.data, o1.ctypes.data,
                                      b2.ctypes.data, o2.ctypes.data, self.n,
                                      None if pm is None else pm.ctypes.data, ctypes.byref(h))
        ctx.check(st)
        self.handle = h

    @classmethod
    def one_to_many(cls, ctx: Context, query, b2, o2, cfg: DeviceConfig, pair_method=None) -> "Batch":
        """A device-resident batch of one query against every packed target
        (pa_batch_create_one_to_many)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.n = len(o2) - 1
        q = np.frombuffer(bytes(query), dtype=np.uint8).copy() if len(query) else np.zeros(1, dtype=np.uint8)
        self.handle = None
        _, _, b2, o2, pm = checked_batch(q, np.array([0, len(query)], dtype=np.uint64), b2, o2, pair_method,
                                         one_query=True)
        self._keep = (q, b2, o2)
        holder = _CConfigHolder(cfg)
        h = ctypes.c_void_p()
        st = ctx._lib.pa_batch_create_one_to_many(ctx.handle, ctypes.byref(holder.c), q.ctypes.data, len(query),
                                                  b2.ctypes.data, o2.ctypes.data, self.n,
                                                  None if pm is None else pm.ctypes.data, ctypes.byref(h))
        ctx.check(st)
        self.handle = h
        return self

    def run(self):
        self.ctx.check(self.ctx._lib.pa_batch_run(self.ctx.handle, self.handle))

    def results(self):
        scores = np.zeros(max(self.n, 1), dtype=np.int32)
        flags = np.zeros(max(self.n, 1), dtype=np.uint8)
        self.ctx.check(self.ctx._lib.pa_batch_results(self.ctx.handle, self.handle, scores.ctypes.data,
                                                      flags.ctypes.data))
        return scores[:self.n], flags[:self.n]

    def stats(self) -> dict:
        s = CStats()
        self.ctx.check(self.ctx._lib.pa_batch_stats_get(self.handle, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in CStats._fields_}

    def time(self, iters: int):
        tot, main = ctypes.c_double(), ctypes.c_double()
        self.ctx.check(self.ctx._lib.pa_batch_time(self.ctx.handle, self.handle, iters, ctypes.byref(tot),
                                                   ctypes.byref(main)))
        return tot.value, main.value

    def close(self):
        if self.handle:
            self.ctx._lib.pa_batch_destroy(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Model:
    """The learned range/cost predictor (pa_model): a 3-layer FP32 MLP living on the GPU."""

    def __init__(self, ctx: Context, seed: int = 0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx.check(ctx._lib.pa_model_create(ctx.handle, seed, ctypes.byref(h)))
        self.handle = h
        self.n_params = ctx._lib.pa_model_param_count()
        self.margin = 0.10   # the library's default lane margin
        self.sigma = None    # held-out relative error of the range output (set by a trainer)

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        self.ctx.check(self.ctx._lib.pa_model_get_params(self.ctx.handle, self.handle, p.ctypes.data))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        self.ctx.check(self.ctx._lib.pa_model_set_params(self.ctx.handle, self.handle, p.ctypes.data))

    def set_margin(self, margin: float):
        """Relative safety margin of the lane choice (pa_model_set_margin)."""
        self.margin = float(margin)
        self.ctx.check(self.ctx._lib.pa_model_set_margin(self.handle, ctypes.c_float(margin)))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), N_OUTPUTS), dtype=np.float32)
        self.ctx.check(self.ctx._lib.pa_model_forward(self.ctx.handle, self.handle, x.ctypes.data, len(x),
                                                      y.ctypes.data))
        return y

    def loss_grad(self, x, target, weight=None):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        grad = np.zeros(self.n_params, dtype=np.float32)
        self.ctx.check(self.ctx._lib.pa_model_loss_grad(self.ctx.handle, self.handle, x.ctypes.data, t.ctypes.data,
                                                        None if w is None else w.ctypes.data, len(x),
                                                        loss.ctypes.data, grad.ctypes.data))
        return float(loss[0]), grad

    def train_step(self, x, target, weight=None, lr: float = 1e-3) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self.ctx.check(self.ctx._lib.pa_model_train_step(self.ctx.handle, self.handle, x.ctypes.data, t.ctypes.data,
                                                         None if w is None else w.ctypes.data, len(x),
                                                         ctypes.c_float(lr), loss.ctypes.data))
        return float(loss[0])

    def close(self):
        if self.handle:
            self.ctx._lib.pa_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def features(ctx: Context, cfg: DeviceConfig, b1, o1, b2, o2) -> np.ndarray:
    b1, o1, b2, o2, _ = checked_batch(b1, o1, b2, o2)
    holder = _CConfigHolder(cfg)
    n = len(o1) - 1
    out = np.zeros((max(n, 1), N_FEATURES), dtype=np.float32)
    ctx.check(ctx._lib.pa_features(ctx.handle, ctypes.byref(holder.c), b1.ctypes.data, o1.ctypes.data,
                                   b2.ctypes.data, o2.ctypes.data, n, out.ctypes.data))
    return out[:n]


def dp_labels(ctx: Context, cfg: DeviceConfig, b1, o1, b2, o2):
    """(max of H over each pair's DP matrix, modelled+measured cost in ns) from an int32 DP pass."""
    b1, o1, b2, o2, _ = checked_batch(b1, o1, b2, o2)
    holder = _CConfigHolder(cfg)
    n = len(o1) - 1
    max_h = np.zeros(max(n, 1), dtype=np.int32)
    cost = np.zeros(max(n, 1), dtype=np.float32)
    ctx.check(ctx._lib.pa_dp_labels(ctx.handle, ctypes.byref(holder.c), b1.ctypes.data, o1.ctypes.data,
                                    b2.ctypes.data, o2.ctypes.data, n, max_h.ctypes.data, cost.ctypes.data))
    return max_h[:n], cost[:n]


def tune(ctx: Context, cfg: DeviceConfig, b1, o1, b2, o2, pair_method=None, sample_pairs: int = 0,
         iters: int = 0) -> Tuning:
    """Runs the library's autotuner (pa_tune) on a sample of the batch; returns the fastest tuning."""
    b1, o1, b2, o2, pm = checked_batch(b1, o1, b2, o2, pair_method)
    holder = _CConfigHolder(cfg)
    n = len(o1) - 1
    out = CTuning()
    ctx.check(ctx._lib.pa_tune(ctx.handle, ctypes.byref(holder.c), b1.ctypes.data, o1.ctypes.data, b2.ctypes.data,
                               o2.ctypes.data, n, None if pm is None else pm.ctypes.data, sample_pairs, iters,
                               ctypes.byref(out)))
    return Tuning.from_c(out)


@functools.lru_cache(maxsize=None)
def default_context(device: int = 0) -> Context:
    return Context(device)
