"""Training driver of the learned lane dispatch.

The DP itself produces the labels (pa_dp_labels: an int32 pass that keeps the maximum of H over
every pair's matrix, and the launch timings shared out over the pairs by modelled cost); the
features and the network's forward, backward and Adam steps run on the GPU (pa_features,
pa_model_*).  The trained network predicts, per pair,
    y0 = max of H / (max substitution score * min(m, n))    -> lane width (int16 when it fits)
    y1 = log2(cost in ns + 1) / 32                           -> cost estimate
"""
from __future__ import annotations

import time
from typing import List, Sequence, Tuple

import numpy as np

from . import device, workloads


# Time of a pair in int16 lanes over its time in int32 lanes: the measured pass ratio of the two
# on the mixed DNA/protein stream (623 against 1,040 ms for the same 500,000 pairs).
INT16_COST_RATIO = 0.6


def targets(dc: device.DeviceConfig, o1, o2, max_h, cost, smax: int) -> np.ndarray:
    mn = np.minimum(np.diff(o1), np.diff(o2)).astype(np.float64)
    y0 = np.where(mn > 0, max_h.astype(np.float64) / (max(smax, 1) * np.maximum(mn, 1)), 0.0)
    y1 = np.log2(cost.astype(np.float64) + 1.0) / 32.0
    return np.stack([y0, y1], axis=1).astype(np.float32)


def max_substitution(dc: device.DeviceConfig) -> int:
    if dc.score_kind == device.SCORE_UNITARY:
        return max(dc.match, dc.mismatch)
    return max(dc.matrix)


def training_pairs(kind: str, seed: int, count: int, lo: int = 100, hi: int = 5000):
    """Pairs like the mixed workload (identity uniform 50-100%), from an independent seed."""
    rng = np.random.default_rng(seed)
    alpha, p = (workloads.DNA, None) if kind == "dna" else (workloads.AMINO, workloads.AMINO_FREQ)
    b1, o1 = workloads.random_packed(rng, rng.integers(lo, hi + 1, size=count), alpha, p)
    div = 1.0 - rng.uniform(0.5, 1.0, size=count)
    b2, o2 = workloads.mutate_packed(rng, b1, o1, div, 0.1, 0.7, alpha, p)
    return b1, o1, b2, o2


def build_dataset(ctx: device.Context, sets: Sequence[Tuple[device.DeviceConfig, tuple]]):
    xs, ys = [], []
    for dc, (b1, o1, b2, o2) in sets:
        max_h, cost = device.dp_labels(ctx, dc, b1, o1, b2, o2)
        x = device.features(ctx, dc, b1, o1, b2, o2)
        xs.append(x)
        ys.append(targets(dc, o1, o2, max_h, cost, max_substitution(dc)))
    return np.concatenate(xs), np.concatenate(ys)


def train(ctx: device.Context, x: np.ndarray, y: np.ndarray, steps: int = 1500, batch: int = 2048,
          lr: float = 2e-3, seed: int = 0, log=None, holdout: float = 0.2) -> Tuple[device.Model, List[float]]:
    """Adam on minibatches; the range output carries 4x the weight of the cost output.

    A `holdout` share of the samples is kept out of training and sets the model's lane margin
    from the range output's relative under-prediction, (true - predicted) / predicted, over the
    held-out pairs whose maximum of H (true or predicted) is 16,384 or more -- where a pair could
    reach the int16 limit -- or, with fewer than 20 such pairs, 8,192 or more.  The margin is the
    quantile at which an int16 guess stops paying: a pair costs about INT16_COST_RATIO of its int32
    time in int16 lanes and, when it saturates, its int32 recomputation on top, so int16 pays while
    the chance of saturating stays below 1 - INT16_COST_RATIO; the margin is the INT16_COST_RATIO
    quantile of the under-prediction (at least 0).  With fewer than 20 pairs either way the
    library default stays.  model.sigma keeps the standard deviation of the relative error over
    the same pairs."""
    model = device.Model(ctx, seed=seed)
    rng = np.random.default_rng(seed + 1)
    order = rng.permutation(len(x))
    n_held = int(len(x) * holdout) if len(x) >= 100 else 0
    held, fit = order[:n_held], order[n_held:]
    w = np.tile(np.array([[4.0, 1.0]], dtype=np.float32), (batch, 1))
    losses = []
    t0 = time.perf_counter()
    for step in range(steps):
        idx = fit[rng.integers(0, len(fit), size=batch)]
        rate = lr * (0.1 ** (step / steps))
        losses.append(model.train_step(x[idx], y[idx], w, lr=rate))
        if log and (step % 250 == 0 or step == steps - 1):
            log(f"[learned] step {step}: loss {losses[-1]:.5f} ({time.perf_counter() - t0:.1f} s)")
    if n_held:
        pred = np.maximum(model.forward(x[held])[:, 0].astype(np.float64), 1e-6)
        true = y[held, 0].astype(np.float64)
        # y0 * smax * min(m, n), the lengths and smax recovered from features 0, 1 and 6
        f = x[held].astype(np.float64)
        scale = f[:, 6] * 16 * np.minimum(np.exp2(f[:, 0] * 16) - 1, np.exp2(f[:, 1] * 16) - 1)
        for floor in (16384, 8192):
            ok = np.maximum(true, pred) * scale >= floor
            if np.count_nonzero(ok) >= 20:
                rel = (true[ok] - pred[ok]) / pred[ok]
                model.sigma = float(np.std(rel))
                model.set_margin(max(0.0, float(np.percentile(rel, 100 * INT16_COST_RATIO))))
                if log:
                    log(f"[learned] held-out pairs scoring {floor} or more: {np.count_nonzero(ok)}, relative error "
                        f"sigma {model.sigma:.4f}, lane margin {model.margin:.4f}")
                break
    return model, losses
