"""Learned lane dispatch: the MLP's GPU forward/backward kernels against the numpy oracle (itself
pinned by finite differences), the GPU features, and what the trained network gains over the
bound-based policy on the same pairs with zero score differences."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from oracle import oracle as O
from pairwise_aligner_amd import device, learned, tables, workloads


def _random_problem(seed, n=96):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, size=M.F * M.H + M.H + M.H * M.H + M.H + M.H * M.O + M.O)
    x = rng.uniform(0, 1, size=(n, M.F))
    t = rng.uniform(0, 1, size=(n, M.O))
    w = rng.uniform(0.5, 2.0, size=(n, M.O))
    return p, x, t, w


def test_mlp_oracle_gradient_matches_finite_differences():
    p, x, t, w = _random_problem(0)
    _, g = M.loss_grad(p, x, t, w)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(len(p), 60, replace=False), [len(p) - 1, len(p) - 2, 0]])
    fd = M.finite_difference(p, x, t, idx, w)
    assert np.allclose(g[idx], fd, rtol=1e-5, atol=1e-8)


def test_mlp_oracle_planted_mistake_is_caught():
    """Dropping the ReLU derivative of layer 1 must be visible against finite differences."""
    p, x, t, w = _random_problem(2)
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    y, (xx, z1, h1, z2, h2) = M.forward(p, x)
    dy = 2.0 * w * (y - t) / len(x)
    dh2 = (dy @ W3.T) * (z2 > 0)
    wrong_gb1 = (dh2 @ W2.T).sum(0)                       # missing * (z1 > 0)
    off = M.F * M.H
    fd = M.finite_difference(p, x, t, list(range(off, off + M.H)), w)
    assert not np.allclose(wrong_gb1, fd, rtol=1e-3, atol=1e-6)


@pytest.mark.gpu
def test_gpu_forward_and_gradient_match_oracle():
    ctx = device.default_context(0)
    model = device.Model(ctx, seed=3)
    p64 = model.get_params().astype(np.float64)
    p, x, t, w = _random_problem(4, n=1000)
    x = x.astype(np.float32)
    y_gpu = model.forward(x)
    y_ref, _ = M.forward(p64, x)
    assert np.allclose(y_gpu, y_ref, rtol=1e-4, atol=1e-5)
    loss_gpu, g_gpu = model.loss_grad(x, t, w)
    loss_ref, g_ref = M.loss_grad(p64, x.astype(np.float64), t, w)
    assert abs(loss_gpu - loss_ref) <= 1e-4 * abs(loss_ref) + 1e-6
    scale = np.max(np.abs(g_ref))
    assert np.max(np.abs(g_gpu - g_ref)) <= 1e-4 * scale, np.max(np.abs(g_gpu - g_ref)) / scale


@pytest.mark.gpu
def test_gpu_features():
    ctx = device.default_context(0)
    rng = np.random.default_rng(5)
    s1, _ = workloads.dna_pairs(rng, 4, 300, 600)
    s2 = [s1[0], s1[1][::-1], s1[2][:150], b"A" * 200]
    b1, o1 = workloads.pack(s1)
    b2, o2 = workloads.pack(s2)
    f = device.features(ctx, device.DeviceConfig(match=2, mismatch=-3), b1, o1, b2, o2)
    assert f.shape == (4, device.N_FEATURES)
    assert abs(f[0, 4] - 1.0) < 1e-6 and abs(f[0, 5] - 1.0) < 1e-6 and f[0, 3] < 1e-6   # identical
    assert f[2, 4] > 0.99                          # a prefix is contained
    assert f[3, 4] < 0.2                           # poly-A against random DNA
    assert np.allclose(f[:, 0], np.log2(np.diff(o1) + 1) / 16, rtol=1e-6)
    assert np.all(f[:, 7] == 1.0) and np.allclose(f[:, 6], 2 / 16)


def _mix32(x):
    x = np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7feb352d) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846ca68b) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def _minimizers(seq, code, k, bits, w=8):
    """Positions' hashes of the sliding-window (w, k)-minimizers of seq: the smallest (hash,
    position) of every window of w consecutive k-mers (one window over all k-mers when there are
    fewer than w), each distinct position once."""
    c = np.array([code[b] for b in seq], dtype=np.uint64) & ((1 << bits) - 1)
    K = len(seq) - k + 1
    if K <= 0:
        return []
    v = np.zeros(K, dtype=np.uint64)
    for j in range(k):
        v = (v << np.uint64(bits)) | c[j:j + K]
    h = _mix32(v & ((1 << (k * bits)) - 1))
    keys = (h << np.uint64(32)) | np.arange(K, dtype=np.uint64)
    nw = K - w + 1 if K > w else 1
    mins = [int(keys[j:j + w].min()) for j in range(nw)]
    out, last = [], None
    for key in mins:
        if key != last:
            out.append(key >> 32)
        last = key
    return out


@pytest.mark.gpu
def test_gpu_minimizer_containment_matches_numpy():
    """Features 4 and 5 (minimizer containment and the identity it implies) against a numpy
    restatement of the same definition, for DNA (2-bit codes, k = 8) and protein (5-bit, k = 3)."""
    ctx = device.default_context(0)
    rng = np.random.default_rng(11)
    for kind in ("dna", "protein"):
        b1, o1, b2, o2 = learned.training_pairs(kind, 12, 40, 5, 700)
        s1, s2 = workloads.unpack(b1, o1), workloads.unpack(b2, o2)
        if kind == "dna":
            dc = device.DeviceConfig(match=2, mismatch=-3)
            seen = sorted(set(b1.tobytes()) | set(b2.tobytes()))
            code, k, bits = {c: i for i, c in enumerate(seen)}, 8, 2
        else:
            t = tables.blosum62_standard()
            dc = device.DeviceConfig(score_kind=device.SCORE_MATRIX, symbols="".join(x for x, _ in t).encode(),
                                     matrix=tuple(v for _, r in t for v in r))
            code, k, bits = {c: c & 31 for c in range(256)}, 3, 5
        f = device.features(ctx, dc, b1, o1, b2, o2)
        for p in range(len(s1)):
            m1 = {h & 32767 for h in _minimizers(s1[p], code, k, bits)}
            m2 = _minimizers(s2[p], code, k, bits)
            cont = (sum((h & 32767) in m1 for h in m2) / len(m2)) if (m2 and len(s1[p]) >= k) else 0.0
            assert abs(f[p, 4] - cont) < 1e-6, (kind, p, f[p, 4], cont)
            assert abs(f[p, 5] - (cont ** (1 / k) if cont > 0 else 0.0)) < 1e-5


@pytest.mark.gpu
def test_learned_dispatch_gains_over_bound_policy():
    """Train on DP-generated labels, then score held-out protein pairs whose length bound does
    not prove int16 safety: the learned policy runs more of them in int16 than the bound policy,
    and every policy -- the learned one also with its tasks ordered by the predicted cost --
    returns the same scores (all equal to the oracle on a sample)."""
    ctx = device.default_context(0)
    t = tables.blosum62_standard()
    prot = device.DeviceConfig(method=device.METHOD_LOCAL, score_kind=device.SCORE_MATRIX,
                               symbols="".join(s for s, _ in t).encode(), matrix=tuple(v for _, r in t for v in r),
                               gap_open=-11, gap_extension=-1)
    dna = device.DeviceConfig(method=device.METHOD_LOCAL, match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
    sets = [(prot, learned.training_pairs("protein", 100, 3000, 100, 4000)),
            (dna, learned.training_pairs("dna", 101, 1500, 100, 4000))]
    x, y = learned.build_dataset(ctx, sets)
    model, losses = learned.train(ctx, x, y, steps=800, batch=1024)
    assert np.mean(losses[-50:]) < 0.25 * np.mean(losses[:10])

    b1, o1, b2, o2 = learned.training_pairs("protein", 202, 600, 3000, 5000)   # held out
    results = {}
    for name, policy, mdl, tuning in (("int32", device.LANES_INT32, None, None), ("auto", device.LANES_AUTO, None, None),
                                      ("learned", device.LANES_LEARNED, model, None),
                                      ("learned_cost_order", device.LANES_LEARNED, model, device.Tuning(cost_order=1))):
        dc = device.DeviceConfig(**{**prot.__dict__, "lane_policy": policy, "model": mdl, "tuning": tuning})
        bt = device.Batch(ctx, b1, o1, b2, o2, dc)
        bt.run()
        results[name] = bt.results() + (bt.stats(),)
    s32, _, st32 = results["int32"]
    sa, _, sta = results["auto"]
    sl, fl, stl = results["learned"]
    sc, _, stc = results["learned_cost_order"]
    assert np.array_equal(s32, sa) and np.array_equal(s32, sl) and np.array_equal(s32, sc)
    assert stc["n_int16"] == stl["n_int16"]   # the cost output orders tasks; it does not pick widths
    assert stl["n_int16"] > sta["n_int16"], (stl, sta)
    want = O.score_packed(b1, o1[:41], b2, o2[:41], O.OracleConfig(mode=O.LOCAL, score_kind=O.MATRIX,
                          symbols=prot.symbols, matrix=prot.matrix, gap_open=-11, gap_extension=-1), threads=8)
    assert np.array_equal(s32[:40].astype(np.int64), want)


def _timed_policies(ctx, w, base, model, iters=3):
    out = {}
    for name, policy, mdl in (("int32", device.LANES_INT32, None), ("bound", device.LANES_AUTO, None),
                              ("int16", device.LANES_INT16, None), ("learned", device.LANES_LEARNED, model)):
        dc = device.DeviceConfig(**{**base.__dict__, "lane_policy": policy, "model": mdl})
        bt = device.Batch(ctx, w["b1"], w["o1"], w["b2"], w["o2"], dc)
        bt.run()                      # plans (and caches) the int32 recomputation of saturated pairs
        ms, _ = bt.time(iters)
        scores, flags = bt.results()
        out[name] = dict(ms=ms / iters, scores=scores, flags=flags, stats=bt.stats())
        bt.close()
    return out


@pytest.mark.gpu
def test_learned_dispatch_beats_speculation_where_int16_saturates():
    """Protein pairs whose local scores straddle the int16 limit (4.8-7.8 kbp at 80-100%
    identity: about one in six scores 32,767 or more).  The length bound runs them all in int32;
    int16 speculation runs them all in int16 and recomputes the saturated ones in int32; the
    learned dispatcher (its lane margin calibrated on held-out pairs) sends the pairs it predicts
    high to int32 at once.  Learned is faster than both speculation and the bound (speculation
    against the bound depends on how much the recomputation of the saturated pairs costs against
    the int16 gain: printed, not asserted); learned recomputes far fewer pairs; every policy
    returns the fixed-int32 scores (and the oracle's on a sample)."""
    ctx = device.default_context(0)
    t = tables.blosum62_standard()
    prot = device.DeviceConfig(method=device.METHOD_LOCAL, score_kind=device.SCORE_MATRIX,
                               symbols="".join(s for s, _ in t).encode(), matrix=tuple(v for _, r in t for v in r),
                               gap_open=-11, gap_extension=-1)
    sets = [(prot, learned.training_pairs("protein", 300, 2500, 100, 5000)),
            (prot, learned.training_pairs("protein", 301, 1500, 4500, 8000))]
    x, y = learned.build_dataset(ctx, sets)
    model, _ = learned.train(ctx, x, y, steps=800, batch=1024)
    print(f"lane margin {model.margin:.4f}, held-out sigma {model.sigma}")
    w = workloads.config_saturating(count=6000)
    r = _timed_policies(ctx, w, prot, model)
    n = len(w["o1"]) - 1
    for name in ("bound", "int16", "learned"):
        assert np.array_equal(r[name]["scores"], r["int32"]["scores"]), name
    want = O.score_packed(w["b1"], w["o1"][:13], w["b2"], w["o2"][:13], O.OracleConfig(
        mode=O.LOCAL, score_kind=O.MATRIX, symbols=prot.symbols, matrix=prot.matrix, gap_open=-11, gap_extension=-1),
        threads=8)
    assert np.array_equal(r["int32"]["scores"][:12].astype(np.int64), want)
    saturated = int(np.count_nonzero(r["int32"]["scores"] >= 32767))
    assert saturated > 0.08 * n
    spec, lrn, bound = r["int16"], r["learned"], r["bound"]
    print({k: (round(v["ms"], 3), v["stats"]["n_int16"], v["stats"]["n_recomputed"]) for k, v in r.items()})
    assert spec["stats"]["n_recomputed"] >= saturated
    assert lrn["stats"]["n_recomputed"] < 0.25 * spec["stats"]["n_recomputed"]
    assert lrn["stats"]["n_int16"] > 0.5 * n
    assert lrn["ms"] < 0.95 * spec["ms"], (lrn["ms"], spec["ms"])
    assert lrn["ms"] < 0.95 * bound["ms"], (lrn["ms"], bound["ms"])
    model.close()


@pytest.mark.gpu
def test_learned_one_to_many_reads_the_query_once():
    """One query against many targets under the learned dispatcher: the feature kernel reads the
    single uploaded query for every pair, so the lane widths it picks equal those of the same pairs
    with the query written out per pair, and the scores equal fixed int32 (and the oracle's on a
    sample)."""
    ctx = device.default_context(0)
    t = tables.blosum62_standard()
    prot = device.DeviceConfig(method=device.METHOD_LOCAL, score_kind=device.SCORE_MATRIX,
                               symbols="".join(s for s, _ in t).encode(), matrix=tuple(v for _, r in t for v in r),
                               gap_open=-11, gap_extension=-1)
    sets = [(prot, learned.training_pairs("protein", 400, 1500, 100, 5000))]
    x, y = learned.build_dataset(ctx, sets)
    model, _ = learned.train(ctx, x, y, steps=300, batch=1024)
    b1, o1, b2, o2 = learned.training_pairs("protein", 401, 400, 200, 4500)
    query = workloads.unpack(b1, o1)[0]
    targets = workloads.unpack(b2, o2)
    rb1, ro1 = workloads.pack([query] * len(targets))
    lrn = device.DeviceConfig(**{**prot.__dict__, "lane_policy": device.LANES_LEARNED, "model": model})
    i32 = device.DeviceConfig(**{**prot.__dict__, "lane_policy": device.LANES_INT32})
    s_one, f_one = ctx.score_one_to_many(query, b2, o2, lrn)
    s_rep, f_rep = ctx.score_packed(rb1, ro1, b2, o2, lrn)
    s_32, _ = ctx.score_one_to_many(query, b2, o2, i32)
    assert np.array_equal(f_one & (device.PAIR_INT16 | device.PAIR_INT32),
                          f_rep & (device.PAIR_INT16 | device.PAIR_INT32))
    assert np.array_equal(s_one, s_32) and np.array_equal(s_rep, s_32)
    assert np.count_nonzero(f_one & device.PAIR_INT16) > 0
    want = O.score_packed(rb1, ro1[:21], b2, o2[:21], O.OracleConfig(
        mode=O.LOCAL, score_kind=O.MATRIX, symbols=prot.symbols, matrix=prot.matrix, gap_open=-11, gap_extension=-1),
        threads=8)
    assert np.array_equal(s_32[:20].astype(np.int64), want)
    model.close()
