"""Benchmark: GCUPS of batched affine-gap score-only DP on one MI355X, with a CPU baseline timed
beside it on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mixed|dna150|protein|uneven|long|
                    long_int16|protein_1xn] [--mixed-pairs N] [--no-extra] [--e2e-reps R] [--pairs N]

A step is one pass of the DP over the whole batch (every pair of the workload), inputs already
resident in HBM.  GCUPS = sum over pairs of m*n / step time / 1e9 (interior cells, as the
reference's benchmark fixture counts them).

Default run: the headline is the mixed DNA/protein stream (500,000 pairs, lengths 100-5,000,
identity 50-100%, both halves local) under the learned lane dispatcher -- the metric's "learned
dispatch vs fixed int32" -- with the same pairs under fixed int32 lanes, int16 speculation and the
length-bound policy beside it.  Every other configuration follows in `configs`: DNA 150 nt global
(10,000 pairs), protein local (100,000 pairs), the uneven 1,000,000-pair batch (autotuned, graph
replay), the 2,000 long pairs (int32 as configured, and re-centred int16), and one protein query
against 200,000 targets; each with its roofline and a sampled CPU baseline.  The line printed
last is the result: one JSON object.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pairwise_aligner_amd import device, tables, workloads  # noqa: E402

METRIC = "GCUPS score-only affine DP on 1 GPU vs CPU ref; learned dispatch vs fixed int32"
# MI355X peaks: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz; packed int16 issues two ops per lane.
PEAK_INT16_TOPS = 256 * 4 * 16 * 2.4e9 * 2 / 1e12   # 78.6
PEAK_INT32_TOPS = PEAK_INT16_TOPS / 2                # 39.3
OPS_PER_CELL = {"global": 8, "local": 10}
COPY_PROBE_BYTES = 256 << 20
# CPU threads of the baseline: the host cores visible to this process, at most 16 (one GPU's
# share of a GPU host's cores).
MAX_CPU_THREADS = 16
# HBM bytes per pass measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (one run each),
# scaled by the counters' reading for a plain copy of known size; see profiles/.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def log(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def measured_traffic(name, digest):
    """HBM bytes per pass for this workload, if profiles/ holds a measurement of these inputs."""
    try:
        t = json.load(open(TRAFFIC_FILE)).get(name)
    except (OSError, ValueError):
        return None
    if not t or t.get("inputs_sha256") != digest:
        return None
    return t["bytes_per_pass"]


def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return dict(model=model, logical_cpus_visible=visible, os_cpu_count=os.cpu_count(),
                threads_used=max(1, min(visible, MAX_CPU_THREADS)))


def blosum_config(method, go=-11, ge=-1, policy=device.LANES_AUTO):
    t = tables.blosum62_standard()
    return device.DeviceConfig(method=method, score_kind=device.SCORE_MATRIX,
                               symbols="".join(s for s, _ in t).encode(), matrix=tuple(v for _, r in t for v in r),
                               gap_open=go, gap_extension=ge, lane_policy=policy)


def oracle_cfg(dc):
    from oracle import oracle as O
    return O.OracleConfig(mode=dc.method, first_column_free=dc.first_column_free, first_row_free=dc.first_row_free,
                          last_column_free=dc.last_column_free, last_row_free=dc.last_row_free,
                          gap_open=dc.gap_open, gap_extension=dc.gap_extension, score_kind=dc.score_kind,
                          match=dc.match, mismatch=dc.mismatch, symbols=dc.symbols, matrix=dc.matrix)


def cells_of(o1, o2):
    return int(np.sum(np.diff(o1).astype(np.int64) * np.diff(o2).astype(np.int64)))


def roofline(ops, cells, kernel_ms, peak, traffic=None):
    achieved = ops * cells / (kernel_ms / 1e3) / 1e12 if kernel_ms > 0 else 0.0
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak, 2), "unit": "TOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic, "ops_per_cell": ops, "kernel_ms": round(kernel_ms, 4)}


# --------------------------------------------------------------------------------------------
# CPU baseline
# --------------------------------------------------------------------------------------------

def cpu_baseline(halves, gpu_scores, budget_s=12.0, one_core=False, cell_cap=2e12):
    """Times the CPU baseline -- oracle/simd_port.cpp, the reference's inter-sequence SIMD scheme
    (one pair per lane, 32 int16 / 16 int32 lanes per AVX-512 register, int16 where the bulk's
    bound allows) restated for AVX-512BW / AVX2 -- on the host's cores, over a prefix of the same
    pairs sized to about `budget_s` seconds, and compares the GPU scores of that prefix with it.
    `halves`: [(b1, o1, b2, o2, device config, pair methods)], sampled alike (the mixed stream's
    two halves).  The prefix has at least 32 pairs per thread (full bulks for every thread) unless
    that exceeds `cell_cap` cells."""
    from oracle import oracle as O
    cpu = host_cpu()
    threads = cpu["threads_used"]
    isa = {512: "AVX-512BW", 256: "AVX2"}.get(O.simd_isa(), "none")

    def prefix(k_of_half, reps=1):
        tot_cells, tot_s, got = 0, 0.0, []
        for (b1, o1, b2, o2, dc, pm), k in zip(halves, k_of_half):
            t0 = time.perf_counter()
            for _ in range(reps):
                g = O.simd_score_packed(b1, o1[:k + 1], b2, o2[:k + 1], oracle_cfg(dc),
                                        pair_method=None if pm is None else pm[:k], threads=nthreads)
            got.append(g)
            tot_s += time.perf_counter() - t0
            tot_cells += reps * cells_of(o1[:k + 1], o2[:k + 1])
        return got, tot_cells, tot_s

    def first_k(o1, o2, cells):
        cum = np.cumsum(np.diff(o1).astype(np.int64) * np.diff(o2).astype(np.int64))
        return int(min(len(cum), max(1, np.searchsorted(cum, cells) + 1))), cum

    def sized(target_s):
        # calibrate on a prefix of each half (about 1e8 cells, and full bulks for every thread
        # unless that exceeds cell_cap), grow every half's prefix towards the time budget, and
        # repeat the whole batch when even all of it takes less
        ks = []
        for b1, o1, b2, o2, _, _ in halves:
            n = len(o1) - 1
            k, cum = first_k(o1, o2, 1e8)
            want = min(32 * nthreads, n)
            k = max(k, want) if cum[want - 1] <= cell_cap else max(1, int(np.searchsorted(cum, cell_cap)))
            ks.append(min(k, n))
        got, cells, dt = prefix(ks)
        for _ in range(3):
            full = all(k == len(h[1]) - 1 for h, k in zip(halves, ks))
            if full or dt >= target_s * 0.5:
                break
            scale = min(64.0, target_s / max(dt, 1e-6))
            ks = [max(k, first_k(o1, o2, cells_of(o1[:k + 1], o2[:k + 1]) * scale)[0])
                  for (b1, o1, b2, o2, _, _), k in zip(halves, ks)]
            got, cells, dt = prefix(ks)
        reps = 1
        if dt < target_s * 0.5 and all(k == len(h[1]) - 1 for h, k in zip(halves, ks)):
            reps = int(min(1000, np.ceil(target_s * 0.5 / max(dt, 1e-6))))
            got, cells, dt = prefix(ks, reps)
        return got, cells, dt, ks, reps

    nthreads = threads
    got, cells, dt, ks, reps = sized(budget_s)
    mism = 0
    for g, s, k in zip(got, gpu_scores, ks):
        mism += int(np.count_nonzero(s[:k].astype(np.int64) != g.astype(np.int64)))
    n_total = sum(len(h[1]) - 1 for h in halves)
    res = dict(value=round(cells / dt / 1e9, 3), unit="GCUPS", cores=threads, kind="port",
               sample=(f"first {'+'.join(str(k) for k in ks)} of {n_total} pairs"
                       f"{f', {reps} times' if reps > 1 else ''} ({cells} cells) in {dt:.2f} s; "
                       f"inter-sequence SIMD port of the reference (oracle/simd_port.cpp, {isa}, int16 lanes where "
                       f"the bound allows, int32 otherwise) on {threads} threads"),
               isa=isa, host_cpu=cpu["model"], host_logical_cpus_visible=cpu["logical_cpus_visible"],
               host_os_cpu_count=cpu["os_cpu_count"], mismatches_vs_gpu=mism)
    if one_core:
        nthreads = 1
        _, cells1, dt1, ks1, reps1 = sized(budget_s / 4)
        res["value_1core"] = round(cells1 / dt1 / 1e9, 3)
        res["sample_1core"] = (f"first {'+'.join(str(k) for k in ks1)} pairs{f', {reps1} times' if reps1 > 1 else ''} "
                               f"({cells1} cells) in {dt1:.2f} s")
    return res


# --------------------------------------------------------------------------------------------
# GPU timing
# --------------------------------------------------------------------------------------------

def gpu_run(ctx, b1, o1, b2, o2, dc, pm, steps, warmup, query=None):
    """Creates the device-resident batch, runs `warmup` untimed passes and times `steps` passes
    (HIP events on the context stream: whole passes, and the main DP launches alone)."""
    if query is not None:
        batch = device.Batch.one_to_many(ctx, query, b2, o2, dc, pm)
    else:
        batch = device.Batch(ctx, b1, o1, b2, o2, dc, pm)
    for _ in range(warmup):
        batch.run()
    total_ms, main_ms = batch.time(steps)
    scores, flags = batch.results()
    st = batch.stats()
    batch.close()
    return dict(ms_per_step=total_ms / steps, kernel_ms=main_ms, scores=scores, flags=flags, stats=st)


def plan_summary(st):
    return {k: st[k] for k in ("n_int16", "n_int32", "n_recomputed", "padded_cells", "n_launches", "n_profiled")}


def dtype_of(stats_list):
    n16 = sum(s["n_int16"] for s in stats_list)
    n32 = sum(s["n_int32"] for s in stats_list) + sum(s["n_recomputed"] for s in stats_list)
    return "int16" if n32 == 0 else ("int32" if n16 == 0 else "int16+int32")


def end_to_end(ctx, w, dc, cells, want_scores, reps):
    """End-to-end time of the one-shot entry point (pa_score_batch / pa_score_one_to_many): host
    planning, upload of the sequences, the pass, download of scores and flags; the median of
    `reps` calls after one warm-up call, and the host planning time of one call."""
    def once():
        t0 = time.perf_counter()
        if "query" in w:
            got, _ = ctx.score_one_to_many(w["query"], w["b2"], w["o2"], dc, w.get("pair_method"))
        else:
            got, _ = ctx.score_packed(w["b1"], w["o1"], w["b2"], w["o2"], dc, w.get("pair_method"))
        return time.perf_counter() - t0, got
    _, got = once()
    times = [once()[0] for _ in range(reps)]
    med = float(np.median(times))
    b = device.Batch(ctx, w["b1"], w["o1"], w["b2"], w["o2"], dc, w.get("pair_method"))
    plan_ms = b.stats()["plan_ms"]
    b.close()
    log(f"[bench] end to end: median {med * 1e3:.1f} ms over {reps} calls, planning {plan_ms:.1f} ms")
    return dict(median_ms=round(med * 1e3, 3), gcups=round(cells / med / 1e9, 2), reps=reps,
                plan_and_upload_ms=round(plan_ms, 2), scores_identical=bool(np.array_equal(got, want_scores)))


def from_files(ctx, w, dc, cells, want_scores, reps):
    """Sequences from FASTA files to scores: the two sides are written as FASTA files (60-column
    lines), then read back with pairwise_aligner_amd.fastx (host, parallel) and scored with the
    one-shot entry point; the median of `reps` runs, and the parse rate of the files alone."""
    import shutil
    import tempfile
    from pairwise_aligner_amd import fastx
    d = tempfile.mkdtemp(prefix="pa_fastx_")
    try:
        paths, nbytes = [], 0
        for side in ("1", "2"):
            b, o = w["b" + side], w["o" + side]
            raw = b.tobytes()
            parts = []
            for r in range(len(o) - 1):
                seq = raw[o[r]:o[r + 1]]
                parts.append(b">s%d\n" % r)
                parts.extend(seq[k:k + 60] + b"\n" for k in range(0, len(seq), 60))
            path = os.path.join(d, f"seq{side}.fa")
            with open(path, "wb") as f:
                f.write(b"".join(parts))
            paths.append(path)
            nbytes += os.path.getsize(path)
        parse_s, total_s = [], []
        for k in range(reps + 1):
            t0 = time.perf_counter()
            r1, r2 = fastx.read(paths[0]), fastx.read(paths[1])
            t1 = time.perf_counter()
            got, _ = ctx.score_packed(r1.residues, r1.offsets, r2.residues, r2.offsets, dc, w.get("pair_method"))
            t2 = time.perf_counter()
            if k == 0:
                same_input = bool(np.array_equal(r1.residues, w["b1"]) and np.array_equal(r2.offsets, w["o2"]))
                same_scores = bool(np.array_equal(got, want_scores))
                continue
            parse_s.append(t1 - t0)
            total_s.append(t2 - t0)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    parse_med, total_med = float(np.median(parse_s)), float(np.median(total_s))
    log(f"[bench] from FASTA files: {nbytes / 1e6:.0f} MB parsed in {parse_med * 1e3:.1f} ms, "
        f"files to scores {total_med * 1e3:.1f} ms")
    return dict(fasta_bytes=nbytes, parse_ms=round(parse_med * 1e3, 2),
                parse_gb_per_s=round(nbytes / parse_med / 1e9, 2), files_to_scores_ms=round(total_med * 1e3, 2),
                gcups=round(cells / total_med / 1e9, 2), reps=reps, input_identical=same_input,
                scores_identical=same_scores)


# --------------------------------------------------------------------------------------------
# The configurations
# --------------------------------------------------------------------------------------------

WORKLOAD_NAMES = {
    "dna150": "DNA global, 10,000 pairs x 150 nt (10% substitutions, 1% indels), +4/-5, gap (-10,-1), int16",
    "protein": "protein local BLOSUM62, 100,000 pairs, lengths 50-500, gap (-11,-1), int16 with int32 recompute",
    "uneven": "1,000,000 DNA pairs, lognormal lengths (median 250, sigma 0.8) 50-10,000, half global half local, "
              "+2/-3, gap (-5,-2), autotuned, graph replay",
    "long": "2,000 DNA pairs 10-100 kbp, 15% divergence, global, +1/-4, gap (-6,-1), int32 wavefront kernel",
    "long_int16": "the same 2,000 long pairs in re-centred int16 lanes (wavefront kernel)",
    "protein_1xn": "one protein query (300 aa) against 200,000 targets 50-500, local BLOSUM62, gap (-11,-1), "
                   "query profile",
    "mixed": "mixed DNA/protein stream, {n} pairs, lengths 100-5,000, identity 50-100%, local (DNA +2/-3 gap "
             "(-5,-2); protein BLOSUM62 gap (-11,-1)), learned lane dispatch",
}


def make_workload(name, pairs=None):
    count = {} if pairs is None else {"count": pairs}
    if name == "dna150":
        w = workloads.config_dna150(**count)
        return w, device.DeviceConfig(method=device.METHOD_GLOBAL, match=4, mismatch=-5, gap_open=-10,
                                      gap_extension=-1), "global"
    if name == "protein":
        return workloads.config_protein(**count), blosum_config(device.METHOD_LOCAL), "local"
    if name == "protein_1xn":
        return workloads.config_protein_1xn(**count), blosum_config(device.METHOD_LOCAL), "local"
    if name == "uneven":
        w = workloads.config_uneven(**count)
        return w, device.DeviceConfig(method=device.METHOD_GLOBAL, match=2, mismatch=-3, gap_open=-5,
                                      gap_extension=-2), "mixed"
    if name in ("long", "long_int16"):
        w = workloads.config_long(**count)
        return w, device.DeviceConfig(method=device.METHOD_GLOBAL, match=1, mismatch=-4, gap_open=-6, gap_extension=-1,
                                      lane_policy=device.LANES_INT32 if name == "long" else device.LANES_AUTO), "global"
    raise SystemExit(f"unknown workload {name}")


def bench_config(name, steps, warmup, pairs=None, with_cpu=True, e2e_reps=0, files=False, w=None, dc=None,
                 method=None, cpu=None):
    """One configuration: GPU GCUPS over `steps` passes, roofline of the pass's DP launches, and
    the CPU baseline on a sample of the same pairs (or `cpu`, a result computed already)."""
    t0 = time.perf_counter()
    if w is None:
        w, dc, method = make_workload(name, pairs)
    n = len(w["o2"]) - 1
    cells = cells_of(w["o1"], w["o2"])
    digest = workloads.digest(w["b1"], w["o1"], w["b2"], w["o2"])
    log(f"[bench] {name}: {n} pairs, {cells:.3e} cells, generated in {time.perf_counter() - t0:.1f} s")
    ctx = device.default_context(0)
    tuned = None
    if name == "uneven":
        # the library's autotuner times candidate planner/launch parameters on a 5% sample
        t1 = time.perf_counter()
        tuning = device.tune(ctx, dc, w["b1"], w["o1"], w["b2"], w["o2"], w.get("pair_method"),
                             sample_pairs=max(2000, n // 20), iters=4)
        dc = dataclasses.replace(dc, tuning=tuning)
        tuned = dict(tuning=dataclasses.asdict(tuning), seconds=round(time.perf_counter() - t1, 2))
        log(f"[bench] {name}: autotuned {tuned}")
    one_to_one = None
    if "query" in w:
        ref = gpu_run(ctx, w["b1"], w["o1"], w["b2"], w["o2"], dc, None, steps, warmup)
        one_to_one = dict(ms_per_step=round(ref["ms_per_step"], 4))
    r = gpu_run(ctx, w["b1"], w["o1"], w["b2"], w["o2"], dc, w.get("pair_method"), steps, warmup, query=w.get("query"))
    gcups = cells / (r["ms_per_step"] / 1e3) / 1e9
    ops = 9 if method == "mixed" else OPS_PER_CELL[method]
    peak = PEAK_INT32_TOPS if r["stats"]["n_int16"] == 0 else PEAK_INT16_TOPS
    rl = roofline(ops, cells, r["kernel_ms"], peak, measured_traffic(name, digest))
    log(f"[bench] {name}: {gcups:.1f} GCUPS ({r['ms_per_step']:.3f} ms/step, DP launches {r['kernel_ms']:.3f} ms), "
        f"{rl['achieved']:.2f} of {rl['peak']:.1f} TOP/s")
    out = dict(workload=WORKLOAD_NAMES.get(name, name) if pairs is None else f"{name} with {pairs} pairs",
               gcups=round(gcups, 2), ms_per_step=round(r["ms_per_step"], 4), steps=steps, pairs=n, cells=cells,
               dtype=dtype_of([r["stats"]]), roofline=rl, plan=plan_summary(r["stats"]), inputs_sha256=digest)
    if tuned:
        out["autotuned"] = tuned
        # every launch form (pa_tuning.persistent) with the tuner's other parameters, on the same
        # pairs: ms per pass
        names = ["default_local_fused", "work_queue_per_launch", "work_queue_local_fused", "one_block_per_16_tasks",
                 "local_and_global_fused"]
        ab, same = {}, True
        for form, form_name in enumerate(names):
            if form == dc.tuning.persistent:
                ab[form_name] = round(r["ms_per_step"], 3)
                continue
            r2 = gpu_run(ctx, w["b1"], w["o1"], w["b2"], w["o2"],
                         dataclasses.replace(dc, tuning=dataclasses.replace(dc.tuning, persistent=form)),
                         w.get("pair_method"), steps, warmup)
            ab[form_name] = round(r2["ms_per_step"], 3)
            same = same and bool(np.array_equal(r2["scores"], r["scores"]))
        ab["tuned"] = names[dc.tuning.persistent]
        ab["scores_identical"] = same
        out["launch_form_ab"] = ab
        log(f"[bench] {name}: launch forms {out['launch_form_ab']}")
    if one_to_one is not None:
        one_to_one["speedup_of_profile"] = round(one_to_one["ms_per_step"] / r["ms_per_step"], 3)
        one_to_one["scores_identical"] = bool(np.array_equal(ref["scores"], r["scores"]))
        out["one_to_one_same_pairs"] = one_to_one
    if cpu is not None:
        out["cpu_baseline"] = cpu
    elif with_cpu:
        out["cpu_baseline"] = cpu_baseline([(w["b1"], w["o1"], w["b2"], w["o2"], dc, w.get("pair_method"))],
                                           [r["scores"]], one_core=(name == "protein"))
        log(f"[bench] {name}: CPU baseline {out['cpu_baseline']['value']} GCUPS "
            f"({out['cpu_baseline']['mismatches_vs_gpu']} mismatches)")
    if e2e_reps:
        out["end_to_end"] = end_to_end(ctx, w, dc, cells, r["scores"], e2e_reps)
        if files:
            out["end_to_end"]["from_fasta_files"] = from_files(ctx, w, dc, cells, r["scores"], min(e2e_reps, 5))
    out["_scores"] = r["scores"]
    return out


def bench_long(steps, warmup):
    """The 2,000 long pairs: int32 lanes as configured, and the same pairs in re-centred int16
    lanes; one CPU baseline (int32, as the configuration asks) for both."""
    w, dc32, method = make_workload("long")
    res32 = bench_config("long", steps, warmup, w=w, dc=dc32, method=method)
    dc16 = dataclasses.replace(dc32, lane_policy=device.LANES_AUTO)
    res16 = bench_config("long_int16", steps, warmup, w=w, dc=dc16, method=method, cpu=res32["cpu_baseline"])
    res16["scores_identical_to_int32"] = bool(np.array_equal(res16["_scores"], res32["_scores"]))
    return res32, res16


def bench_mixed(count, steps, warmup, side_steps=5, only=None):
    """Learned lane dispatch on the mixed stream: the same pairs under the learned policy (timed
    over `steps` passes: the headline), fixed int32 lanes, int16 speculation (every pair int16,
    saturated pairs recomputed in int32) and the length-bound policy (timed over
    min(steps, side_steps) passes each).  The network is trained first, on labels the DP produces
    for separately seeded training pairs.  `only` (a set of leg names) runs just those legs, with
    no CPU baseline and no saturating slice: a short run to profile one leg."""
    from pairwise_aligner_amd import learned
    t0 = time.perf_counter()
    w = workloads.config_mixed(count=count)
    log(f"[bench] mixed: {count} pairs generated in {time.perf_counter() - t0:.1f} s")
    ctx = device.default_context(0)
    dna = device.DeviceConfig(method=device.METHOD_LOCAL, match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
    prot = blosum_config(device.METHOD_LOCAL)
    t1 = time.perf_counter()
    sets = [(prot, learned.training_pairs("protein", 1001, 4000)), (dna, learned.training_pairs("dna", 1002, 2000)),
            (prot, learned.training_pairs("protein", 1003, 1000, 4500, 8000))]
    x, y = learned.build_dataset(ctx, sets)
    model, losses = learned.train(ctx, x, y, steps=1500, batch=2048, log=log)
    train_s = time.perf_counter() - t1
    halves = [("dna", dna, w["dna"]), ("protein", prot, w["protein"])]
    cells = sum(cells_of(h["o1"], h["o2"]) for _, _, h in halves)
    digest = workloads.digest(*[h[k] for _, _, h in halves for k in ("b1", "o1", "b2", "o2")])
    out = {"workload": WORKLOAD_NAMES["mixed"].format(n=f"{count:,}"), "pairs": count, "cells": cells,
           "inputs_sha256": digest, "train_seconds": round(train_s, 2),
           "train_loss_first_last": [round(float(np.mean(losses[:20])), 5), round(float(np.mean(losses[-20:])), 5)],
           "range_error_sigma_heldout_high_scores": None if model.sigma is None else round(model.sigma, 5),
           "lane_margin": round(model.margin, 5),
           "policies": {}}
    ref_scores = {}
    side = min(steps, side_steps)
    # the last leg orders each launch's waves by the network's predicted cost instead of the
    # geometric model (an A/B of the cost output; the planner's default is the geometric order)
    legs = (("learned", device.LANES_LEARNED, steps, warmup, None), ("int32", device.LANES_INT32, side, 1, None),
            ("int16_speculative", device.LANES_INT16, side, 1, None), ("bound", device.LANES_AUTO, side, 1, None),
            ("learned_cost_order", device.LANES_LEARNED, side, 1, device.Tuning(cost_order=1)))
    for name, policy, k, wu, tuning in legs:
        if only and name not in only:
            continue
        total_ms = kernel_ms = 0.0
        stats, mism = [], 0
        for kind, dc, h in halves:
            dcp = dataclasses.replace(dc, lane_policy=policy, model=model if policy == device.LANES_LEARNED else None,
                                      tuning=tuning)
            r = gpu_run(ctx, h["b1"], h["o1"], h["b2"], h["o2"], dcp, None, k, wu)
            total_ms += r["ms_per_step"]
            kernel_ms += r["kernel_ms"]
            stats.append(r["stats"])
            if kind in ref_scores:
                mism += int(np.count_nonzero(ref_scores[kind] != r["scores"]))
            else:
                ref_scores[kind] = r["scores"]
        leg = {"gcups": round(cells / (total_ms / 1e3) / 1e9, 2), "ms_per_step": round(total_ms, 3), "steps": k,
               "kernel_ms": round(kernel_ms, 3), "dtype": dtype_of(stats),
               "n_int16": sum(s["n_int16"] for s in stats), "n_int32": sum(s["n_int32"] for s in stats),
               "n_recomputed": sum(s["n_recomputed"] for s in stats),
               "score_mismatches_vs_learned": mism,
               "roofline": roofline(OPS_PER_CELL["local"], cells, kernel_ms,
                                    PEAK_INT32_TOPS if policy == device.LANES_INT32 else PEAK_INT16_TOPS,
                                    measured_traffic(f"mixed_{name}", digest))}
        out["policies"][name] = leg
        log(f"[bench] mixed/{name}: {leg}")
    pol = out["policies"]
    if only:
        model.close()
        return out
    for name in ("int16_speculative", "bound"):
        out[f"speedup_{name}_vs_int32"] = round(pol[name]["gcups"] / pol["int32"]["gcups"], 3)
    out["speedup_learned_vs_int32"] = round(pol["learned"]["gcups"] / pol["int32"]["gcups"], 3)
    out["speedup_learned_vs_int16_speculative"] = round(pol["learned"]["gcups"] / pol["int16_speculative"]["gcups"], 3)
    # the CPU baseline on a prefix of each half (the same pairs), and an oracle spot check
    out["cpu_baseline"] = cpu_baseline([(h["b1"], h["o1"], h["b2"], h["o2"], dc, None) for _, dc, h in halves],
                                       [ref_scores[k] for k, _, _ in halves], one_core=True)
    log(f"[bench] mixed: CPU baseline {out['cpu_baseline']['value']} GCUPS")
    from oracle import oracle as O
    bad = 0
    for kind, dc, h in halves:
        k = 40
        want = O.score_packed(h["b1"], h["o1"][:k + 1], h["b2"], h["o2"][:k + 1], oracle_cfg(dc),
                              threads=host_cpu()["threads_used"])
        bad += int(np.count_nonzero(ref_scores[kind][:k].astype(np.int64) != want))
    out["oracle_sample_mismatches"] = bad
    out["oracle_sample_pairs"] = 80
    out["saturating_slice"] = bench_saturating(model, prot, min(steps, side_steps))
    model.close()
    return out


def bench_saturating(model, prot, steps, count=20_000):
    """Where the lane choice decides: 20,000 protein pairs of 4.8-7.8 kbp at 80-100% identity,
    about a quarter of which score beyond the int16 lanes' range.  The same four policies with the
    same trained network; int16 speculation pays an int32 recomputation for every saturated pair
    (the batch's first run plans it, the timed runs replay it).  Each policy is timed with the
    default launch shape and with the shape the library's autotuner picks for it on this batch
    (pa_tune: long pairs split between the striped and the wavefront kernel, streams, launch form),
    and credited with the faster of the two: how a small batch of long pairs is best laid out
    depends on the widths the policy mixes."""
    t0 = time.perf_counter()
    w = workloads.config_saturating(count=count)
    cells = cells_of(w["o1"], w["o2"])
    log(f"[bench] saturating: {count} pairs, {cells:.3e} cells, generated in {time.perf_counter() - t0:.1f} s")
    ctx = device.default_context(0)
    out = {"workload": "20,000 protein pairs 4.8-7.8 kbp, identity 80-100%, local BLOSUM62 gap (-11,-1)",
           "pairs": count, "cells": cells, "policies": {}}
    ref = None
    for name, policy in (("learned", device.LANES_LEARNED), ("int32", device.LANES_INT32),
                         ("int16_speculative", device.LANES_INT16), ("bound", device.LANES_AUTO)):
        dc = dataclasses.replace(prot, lane_policy=policy, model=model if policy == device.LANES_LEARNED else None)
        r0 = gpu_run(ctx, w["b1"], w["o1"], w["b2"], w["o2"], dc, None, steps, 1)
        t1 = time.perf_counter()
        tuning = device.tune(ctx, dc, w["b1"], w["o1"], w["b2"], w["o2"], None, iters=4)
        tune_s = time.perf_counter() - t1
        r1 = gpu_run(ctx, w["b1"], w["o1"], w["b2"], w["o2"], dataclasses.replace(dc, tuning=tuning), None, steps, 1)
        # the policy is credited with the faster of its two measured shapes (the tuner times two
        # passes per candidate, which can mislead it where a recomputation follows the pass)
        r = r1 if r1["ms_per_step"] <= r0["ms_per_step"] else r0
        mism = 0 if ref is None else int(np.count_nonzero(ref != r["scores"]))
        mism += int(np.count_nonzero(r0["scores"] != r1["scores"]))
        ref = r["scores"] if ref is None else ref
        out["policies"][name] = {"gcups": round(cells / (r["ms_per_step"] / 1e3) / 1e9, 2),
                                 "ms_per_step": round(r["ms_per_step"], 3), "steps": steps,
                                 "shape": "autotuned" if r is r1 else "default",
                                 "default_shape_ms_per_step": round(r0["ms_per_step"], 3),
                                 "tuned_shape_ms_per_step": round(r1["ms_per_step"], 3),
                                 "tuning": dataclasses.asdict(tuning), "tune_seconds": round(tune_s, 1),
                                 "n_int16": r["stats"]["n_int16"], "n_int32": r["stats"]["n_int32"],
                                 "n_recomputed": r["stats"]["n_recomputed"], "score_mismatches_vs_learned": mism}
        log(f"[bench] saturating/{name}: {out['policies'][name]}")
    pol = out["policies"]
    out["scores_above_int16_top"] = int(np.count_nonzero(ref > 31744))
    for name in ("learned", "int16_speculative", "bound"):
        out[f"speedup_{name}_vs_int32"] = round(pol[name]["gcups"] / pol["int32"]["gcups"], 3)
    out["speedup_learned_vs_int16_speculative"] = round(pol["learned"]["gcups"] / pol["int16_speculative"]["gcups"], 3)
    return out


def copy_probe():
    import torch
    a = torch.arange(COPY_PROBE_BYTES // 4, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": COPY_PROBE_BYTES, "copies": 3, "check": int(b[-1].item())}))


def strip(d):
    return {k: v for k, v in d.items() if not k.startswith("_")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mixed")
    ap.add_argument("--no-extra", action="store_true", help="only the named workload")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy three times (scales the rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE counters for roofline.traffic)")
    ap.add_argument("--e2e-reps", type=int, default=10,
                    help="calls of the one-shot entry point timed end to end on the protein workload (0: skip)")
    ap.add_argument("--pairs", type=int, default=None, help="pair count of the named workload (not mixed)")
    ap.add_argument("--mixed-pairs", type=int, default=500_000, help="pairs of the mixed stream (config: 500,000)")
    ap.add_argument("--mixed-legs", default=None,
                    help="comma-separated policies of the mixed stream to run (learned, int32, int16_speculative, "
                         "bound, learned_cost_order); the headline leg must be among them; no CPU baseline, "
                         "no saturating slice (for profiling one leg)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (pairs are independent; multi-GPU sharding is not built); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    line = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "data": "synthetic (seeded generators in pairwise_aligner_amd/workloads.py)"}
    configs = {}
    if args.workload == "mixed":
        only = set(args.mixed_legs.split(",")) if args.mixed_legs else None
        if only is not None and "learned" not in only:
            sys.exit("--mixed-legs must include learned (the headline leg)")
        m = bench_mixed(args.mixed_pairs, args.steps, args.warmup, only=only)
        lead = m["policies"]["learned"]
        line.update({"value": lead["gcups"], "unit": "GCUPS", "ms_per_step": lead["ms_per_step"],
                     "dtype": lead["dtype"],
                     "config": {"workload": m["workload"], "pairs": m["pairs"], "cells": m["cells"],
                                "inputs_sha256": m["inputs_sha256"]},
                     "roofline": lead["roofline"], "cpu_baseline": m.get("cpu_baseline"),
                     "learned_vs_int32": m.get("speedup_learned_vs_int32"),
                     "learned_dispatch": {k: v for k, v in m.items()
                                          if k not in ("workload", "pairs", "cells", "inputs_sha256", "cpu_baseline")}})
    else:
        r = bench_config(args.workload, args.steps, args.warmup, pairs=args.pairs,
                         e2e_reps=args.e2e_reps if args.workload == "protein" else 0,
                         files=args.workload == "protein")
        line.update({"value": r["gcups"], "unit": "GCUPS", "ms_per_step": r["ms_per_step"], "dtype": r["dtype"],
                     "config": {"workload": r["workload"], "pairs": r["pairs"], "cells": r["cells"],
                                "inputs_sha256": r["inputs_sha256"]},
                     "roofline": r["roofline"], "cpu_baseline": r.get("cpu_baseline")})
        line["details"] = strip({k: v for k, v in r.items() if k not in ("roofline", "cpu_baseline")})
    if not args.no_extra:
        side = min(args.steps, 5)
        if args.workload != "dna150":
            configs["dna150_global_int16"] = strip(bench_config("dna150", max(args.steps, 50), args.warmup))
        if args.workload != "protein":
            configs["protein_local_100k"] = strip(bench_config("protein", args.steps, args.warmup,
                                                               e2e_reps=args.e2e_reps, files=args.e2e_reps > 0))
        if args.workload != "uneven":
            configs["uneven_1m_autotuned_graph"] = strip(bench_config("uneven", side, 1, e2e_reps=min(args.e2e_reps, 3)))
        if args.workload not in ("long", "long_int16"):
            r32, r16 = bench_long(min(args.steps, 3), 1)
            configs["long_2000_int32"] = strip(r32)
            configs["long_2000_recentred_int16"] = strip(r16)
        if args.workload != "protein_1xn":
            configs["protein_1xn_query_profile"] = strip(bench_config("protein_1xn", max(args.steps, 10), args.warmup,
                                                                      with_cpu=False))
        line["configs"] = configs
    print(json.dumps(line))


if __name__ == "__main__":
    main()
