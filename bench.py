"""Benchmark: GCUPS of batched affine-gap score-only DP on one MI355X, with the CPU oracle timed
beside it on the host cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload protein|dna150|uneven|long|mixed]

A step is one pass of the DP over the whole batch (all pairs of the workload), with inputs
already resident in HBM.  `value` is GCUPS = sum over pairs of m*n / step time / 1e9 (interior
cells, as the reference's benchmark fixture counts them).  The line printed last is the result:
one JSON object.
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

# MI355X peaks: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz; packed int16 issues two ops per lane.
PEAK_INT16_TOPS = 256 * 4 * 16 * 2.4e9 * 2 / 1e12   # 78.6
PEAK_INT32_TOPS = PEAK_INT16_TOPS / 2                # 39.3
PEAK_HBM_GBS = 8000.0
OPS_PER_CELL = {"global": 8, "local": 10}
COPY_PROBE_BYTES = 256 << 20
# HBM bytes per pass measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (one run each),
# scaled by the counters' reading for a plain copy of known size; see profiles/.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def measured_traffic(name, digest):
    """HBM bytes per pass for this workload, if profiles/ holds a measurement of these inputs."""
    try:
        t = json.load(open(TRAFFIC_FILE)).get(name)
    except (OSError, ValueError):
        return None
    if not t or t.get("inputs_sha256") != digest:
        return None
    return t


def log(msg):
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def blosum_config(method, go, ge, policy):
    t = tables.blosum62_standard()
    return device.DeviceConfig(method=method, score_kind=device.SCORE_MATRIX,
                               symbols="".join(s for s, _ in t).encode(), matrix=tuple(v for _, r in t for v in r),
                               gap_open=go, gap_extension=ge, lane_policy=policy)


def make_workload(name, pairs=None):
    """The named workload; `pairs` overrides the pair count of the DNA 150 nt and protein workloads
    (the batch-size sweep of the small DNA configuration)."""
    count = {} if pairs is None else {"count": pairs}
    if name == "dna150":
        w = workloads.config_dna150(**count)
        dc = device.DeviceConfig(method=device.METHOD_GLOBAL, match=4, mismatch=-5, gap_open=-10, gap_extension=-1)
        return w, dc, "global"
    if name == "protein":
        w = workloads.config_protein(**count)
        dc = blosum_config(device.METHOD_LOCAL, -11, -1, device.LANES_AUTO)
        return w, dc, "local"
    if name == "protein_1xn":
        w = workloads.config_protein_1xn(**count)
        dc = blosum_config(device.METHOD_LOCAL, -11, -1, device.LANES_AUTO)
        return w, dc, "local"
    if name == "uneven":
        w = workloads.config_uneven()
        dc = device.DeviceConfig(method=device.METHOD_GLOBAL, match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
        return w, dc, "mixed"
    if name in ("long", "long_int16"):
        # the configuration asks for int32 scores; long_int16 runs the same pairs in re-centred
        # int16 lanes (the planner's default for long global pairs within the spread bound)
        w = workloads.config_long()
        dc = device.DeviceConfig(method=device.METHOD_GLOBAL, match=1, mismatch=-4, gap_open=-6, gap_extension=-1,
                                 lane_policy=device.LANES_INT32 if name == "long" else device.LANES_AUTO)
        return w, dc, "global"
    raise SystemExit(f"unknown workload {name}")


def oracle_cfg(dc):
    from oracle import oracle as O
    return O.OracleConfig(mode=dc.method, first_column_free=dc.first_column_free, first_row_free=dc.first_row_free,
                          last_column_free=dc.last_column_free, last_row_free=dc.last_row_free,
                          gap_open=dc.gap_open, gap_extension=dc.gap_extension, score_kind=dc.score_kind,
                          match=dc.match, mismatch=dc.mismatch, symbols=dc.symbols, matrix=dc.matrix)


def cpu_baseline(w, dc, gpu_scores, budget_s=12.0):
    """Times the CPU baseline -- oracle/simd_port.cpp, the reference's inter-sequence SIMD scheme
    (one pair per lane, int16 lanes where the bound allows) restated for AVX-512BW / AVX2 -- on
    the host cores, on a bounded prefix of the same pairs, and compares the GPU scores of that
    prefix with it.  A one-thread rate on a smaller prefix is reported beside it."""
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    b1, o1, b2, o2 = w["b1"], w["o1"], w["b2"], w["o2"]
    n = len(o1) - 1
    oc = oracle_cfg(dc)
    pm_all = w.get("pair_method")
    cells_cum = np.cumsum(np.diff(o1).astype(np.int64) * np.diff(o2).astype(np.int64))
    isa = {512: "AVX-512BW", 256: "AVX2"}.get(O.simd_isa(), "none")

    def timed(sample, nthreads):
        pm = None if pm_all is None else pm_all[:sample]
        t0 = time.perf_counter()
        got = O.simd_score_packed(b1, o1[:sample + 1], b2, o2[:sample + 1], oc, pair_method=pm, threads=nthreads)
        return got, time.perf_counter() - t0

    def sized(target_s, nthreads):
        # calibrate on a prefix of about 1e8 cells, then size the sample to target_s seconds
        cal = int(min(n, max(64 * nthreads, np.searchsorted(cells_cum, 1e8) + 1)))
        got, dt = timed(cal, nthreads)
        rate = cells_cum[cal - 1] / max(dt, 1e-6)
        sample = int(min(n, max(cal, np.searchsorted(cells_cum, rate * target_s) + 1)))
        if sample > cal:
            got, dt = timed(sample, nthreads)
        else:
            sample = cal
        return got, dt, sample

    got, dt, sample = sized(budget_s, threads)
    cells = int(cells_cum[sample - 1])
    _, dt1, sample1 = sized(budget_s / 4, 1)
    cells1 = int(cells_cum[sample1 - 1])
    mism = int(np.count_nonzero(gpu_scores[:sample].astype(np.int64) != got.astype(np.int64)))
    return dict(value=cells / dt / 1e9, unit="GCUPS", cores=threads, kind="port",
                sample=f"first {sample} of {n} pairs ({cells} cells) in {dt:.2f} s; inter-sequence SIMD port "
                       f"of the reference (oracle/simd_port.cpp, {isa}, int16 lanes where the bound allows) "
                       f"on {threads} threads",
                value_1core=cells1 / dt1 / 1e9, sample_1core=f"first {sample1} pairs ({cells1} cells) in {dt1:.2f} s",
                isa=isa, mismatches_vs_gpu=mism, sample_pairs=sample)


def run(name, steps, warmup, with_cpu=True, autotune=False, pairs=None, e2e_reps=0):
    t0 = time.perf_counter()
    w, dc, method = make_workload(name, pairs)
    n = len(w["o1"]) - 1
    cells = int(np.sum(np.diff(w["o1"]).astype(np.int64) * np.diff(w["o2"]).astype(np.int64)))
    log(f"[bench] {name}: {n} pairs, {cells:.3e} cells, generated in {time.perf_counter() - t0:.1f} s")
    ctx = device.default_context(0)
    tuned = None
    if autotune:
        # the library's autotuner times candidate planner/launch parameters on a 5% sample
        t1 = time.perf_counter()
        tuning = device.tune(ctx, dc, w["b1"], w["o1"], w["b2"], w["o2"], w.get("pair_method"),
                             sample_pairs=max(2000, n // 20), iters=2)
        dc = dataclasses.replace(dc, tuning=tuning)
        tuned = dict(tuning=dataclasses.asdict(tuning), seconds=round(time.perf_counter() - t1, 2))
        log(f"[bench] {name}: autotuned {tuned}")
    one_to_one = None
    if "query" in w:
        # one query against many targets: the query-profile path, beside the one-to-one path
        # on the same pairs (the query repeated)
        ref = device.Batch(ctx, w["b1"], w["o1"], w["b2"], w["o2"], dc, w.get("pair_method"))
        for _ in range(warmup):
            ref.run()
        ref_total, ref_main = ref.time(steps)
        ref_scores, _ = ref.results()
        ref.close()
        one_to_one = dict(ms_per_step=ref_total / steps, main_kernel_ms=ref_main)
        batch = device.Batch.one_to_many(ctx, w["query"], w["b2"], w["o2"], dc, w.get("pair_method"))
    else:
        batch = device.Batch(ctx, w["b1"], w["o1"], w["b2"], w["o2"], dc, w.get("pair_method"))
    st = batch.stats()
    log(f"[bench] plan: {st}")
    for _ in range(warmup):
        batch.run()
    try:
        import torch
        torch.cuda.synchronize()
    except Exception:
        pass
    t1 = time.perf_counter()
    total_ms, main_ms = batch.time(steps)
    wall = time.perf_counter() - t1
    scores, flags = batch.results()
    st = batch.stats()
    ms_step = total_ms / steps
    gcups = cells / (ms_step / 1e3) / 1e9
    ops = OPS_PER_CELL["local" if method == "local" else "global"] if method != "mixed" else 9
    peak = PEAK_INT16_TOPS if dc.lane_policy != device.LANES_INT32 else PEAK_INT32_TOPS
    achieved = ops * cells / (main_ms / 1e3) / 1e12 if main_ms > 0 else 0.0
    res = dict(name=name, pairs=n, cells=cells, ms_per_step=ms_step, main_kernel_ms=main_ms, wall_s=wall,
               gcups=gcups, achieved_tops=achieved, peak_tops=peak, ops_per_cell=ops, stats=st,
               inputs_sha256=workloads.digest(w["b1"], w["o1"], w["b2"], w["o2"]), autotuned=tuned)
    if one_to_one is not None:
        one_to_one["scores_identical"] = bool(np.array_equal(ref_scores, scores))
        one_to_one["speedup_of_profile"] = one_to_one["ms_per_step"] / ms_step
        res["one_to_one"] = one_to_one
    if with_cpu:
        res["cpu"] = cpu_baseline(w, dc, scores)
    log(f"[bench] {name}: {gcups:.1f} GCUPS ({ms_step:.3f} ms/step, main kernels {main_ms:.3f} ms), "
        f"{achieved:.2f} of {peak:.1f} TOP/s")
    batch.close()
    if e2e_reps:
        res["end_to_end"] = end_to_end(ctx, w, dc, cells, scores, e2e_reps)
        if "query" not in w:
            res["end_to_end"]["from_fasta_files"] = from_files(ctx, w, dc, cells, scores, min(e2e_reps, 5))
    return res


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
    if "query" in w:
        b = device.Batch.one_to_many(ctx, w["query"], w["b2"], w["o2"], dc, w.get("pair_method"))
    else:
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


def run_mixed(count, steps, warmup, with_oracle=True):
    """Learned lane dispatch on the mixed workload: the same pairs under fixed int32 lanes, the
    length-bound policy and the learned policy.  Both halves are aligned locally (DNA +2/-3,
    gap (-5,-2); protein BLOSUM62, gap (-11,-1)).  The network is trained first, on labels the
    DP produces for separately seeded training pairs."""
    from pairwise_aligner_amd import learned
    t0 = time.perf_counter()
    w = workloads.config_mixed(count=count)
    log(f"[bench] mixed: {count} pairs generated in {time.perf_counter() - t0:.1f} s")
    ctx = device.default_context(0)
    dna = device.DeviceConfig(method=device.METHOD_LOCAL, match=2, mismatch=-3, gap_open=-5, gap_extension=-2)
    prot = blosum_config(device.METHOD_LOCAL, -11, -1, device.LANES_AUTO)
    t1 = time.perf_counter()
    sets = [(prot, learned.training_pairs("protein", 1001, 4000)), (dna, learned.training_pairs("dna", 1002, 2000))]
    x, y = learned.build_dataset(ctx, sets)
    model, losses = learned.train(ctx, x, y, steps=1500, batch=2048, log=log)
    train_s = time.perf_counter() - t1
    halves = [("dna", dna, w["dna"]), ("protein", prot, w["protein"])]
    cells = {k: int(np.sum(np.diff(h["o1"]).astype(np.int64) * np.diff(h["o2"]).astype(np.int64)))
             for k, _, h in halves}
    out = {"pairs": count, "cells": sum(cells.values()), "train_seconds": round(train_s, 2),
           "train_loss_first_last": [round(float(np.mean(losses[:20])), 5), round(float(np.mean(losses[-20:])), 5)]}
    ref_scores = {}
    for name, policy in (("int32", device.LANES_INT32), ("bound", device.LANES_AUTO), ("learned", device.LANES_LEARNED)):
        total_ms = 0.0
        n16 = n32 = nrec = 0
        mism = 0
        for kind, dc, h in halves:
            dcp = device.DeviceConfig(**{**dc.__dict__, "lane_policy": policy,
                                         "model": model if policy == device.LANES_LEARNED else None})
            bt = device.Batch(ctx, h["b1"], h["o1"], h["b2"], h["o2"], dcp)
            for _ in range(warmup):
                bt.run()
            ms, _ = bt.time(steps)
            total_ms += ms / steps
            sc, _ = bt.results()
            st = bt.stats()
            n16 += st["n_int16"]; n32 += st["n_int32"]; nrec += st["n_recomputed"]
            if kind in ref_scores:
                mism += int(np.count_nonzero(ref_scores[kind] != sc))
            else:
                ref_scores[kind] = sc
            bt.close()
        out[name] = {"gcups": round(out["cells"] / (total_ms / 1e3) / 1e9, 2), "ms_per_step": round(total_ms, 3),
                     "n_int16": n16, "n_int32": n32, "n_recomputed": nrec, "score_mismatches_vs_int32": mism}
        log(f"[bench] mixed/{name}: {out[name]}")
    out["speedup_learned_vs_int32"] = round(out["learned"]["gcups"] / out["int32"]["gcups"], 3)
    out["speedup_bound_vs_int32"] = round(out["bound"]["gcups"] / out["int32"]["gcups"], 3)
    if with_oracle:
        from oracle import oracle as O
        bad = 0
        for kind, dc, h in halves:
            k = 40
            want = O.score_packed(h["b1"], h["o1"][:k + 1], h["b2"], h["o2"][:k + 1], oracle_cfg(dc),
                                  threads=min(os.cpu_count() or 1, 16))
            bad += int(np.count_nonzero(ref_scores[kind][:k].astype(np.int64) != want))
        out["oracle_sample_mismatches"] = bad
        out["oracle_sample_pairs"] = 80
    model.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="protein")
    ap.add_argument("--no-extra", action="store_true", help="skip the secondary workloads")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device copy three times (scales the rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE counters for roofline.traffic)")
    ap.add_argument("--e2e-reps", type=int, default=10,
                    help="calls of the one-shot entry point timed end to end (0: skip)")
    ap.add_argument("--pairs", type=int, default=None,
                    help="pair count of the dna150 / protein workload (default: 10,000 / 100,000)")
    ap.add_argument("--mixed-pairs", type=int, default=100_000,
                    help="pairs of the mixed workload used for the learned-dispatch comparison (config: 500,000)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.arange(COPY_PROBE_BYTES // 4, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": COPY_PROBE_BYTES, "copies": 3, "check": int(b[-1].item())}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (pairs are independent; multi-GPU sharding is not built); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    if args.workload == "mixed":
        res = run_mixed(args.mixed_pairs, max(1, args.steps), args.warmup)
        print(json.dumps({"metric": "GCUPS score-only affine DP on 1 GPU vs CPU ref; learned dispatch vs fixed int32",
                          "value": res["learned"]["gcups"], "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": res["learned"]["ms_per_step"],
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int16+int32",
                          "data": "synthetic", "config": {"workload": f"mixed DNA/protein, {args.mixed_pairs} pairs"},
                          "mixed_learned_dispatch": res}))
        return
    main_res = run(args.workload, args.steps, args.warmup, autotune=args.workload == "uneven", pairs=args.pairs,
                   e2e_reps=args.e2e_reps)
    extra = mixed = uneven = None
    if not args.no_extra and args.workload != "dna150":
        extra = run("dna150", max(args.steps, 50), args.warmup, with_cpu=False)
    if not args.no_extra and args.workload != "uneven":
        uneven = run("uneven", 5, 1, with_cpu=False, autotune=True, e2e_reps=min(args.e2e_reps, 5))
    otm = None
    if not args.no_extra and args.workload != "protein_1xn":
        otm = run("protein_1xn", max(args.steps, 10), args.warmup, with_cpu=False)
    if not args.no_extra:
        mixed = run_mixed(args.mixed_pairs, 1, 1)

    cpu = main_res.pop("cpu")
    traffic = measured_traffic(args.workload, main_res["inputs_sha256"])
    frac = main_res["achieved_tops"] / main_res["peak_tops"]
    line = {
        "metric": "GCUPS score-only affine DP on 1 GPU vs CPU ref; learned dispatch vs fixed int32",
        "value": round(main_res["gcups"], 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16" if main_res["stats"]["n_int32"] == 0 else "int16+int32",
        "data": "synthetic (seeded generators in pairwise_aligner_amd/workloads.py)",
        "config": {"workload": {"protein": "protein local BLOSUM62, 100,000 pairs, lengths 50-500, gap (-11,-1)",
                                "dna150": "DNA global, 10,000 pairs x 150 nt, +4/-5, gap (-10,-1)",
                                "uneven": "1,000,000 DNA pairs lognormal lengths, half global half local",
                                "long": "2,000 DNA pairs 10-100 kbp, int32",
                                "long_int16": "2,000 DNA pairs 10-100 kbp, re-centred int16 lanes",
                                "protein_1xn": "one protein query (300 aa) against 200,000 targets 50-500, "
                                               "local BLOSUM62, gap (-11,-1), query profile"}.get(args.workload,
                                                                                                   args.workload)
                   if args.pairs is None else f"{args.workload} workload with {args.pairs} pairs",
                   "pairs": main_res["pairs"], "cells": main_res["cells"],
                   "inputs_sha256": main_res["inputs_sha256"]},
        "roofline": {"bound": "valu", "achieved": round(main_res["achieved_tops"], 3),
                     "peak": round(main_res["peak_tops"], 2), "unit": "TOP/s", "frac": round(frac, 4),
                     "traffic": None if traffic is None else traffic["bytes_per_pass"],
                     "ops_per_cell": main_res["ops_per_cell"], "kernel_ms": round(main_res["main_kernel_ms"], 4)},
        "cpu_baseline": {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample", "value_1core",
                                             "sample_1core", "isa")},
        "mismatches_vs_oracle_sample": cpu["mismatches_vs_gpu"],
        "plan": {k: main_res["stats"][k] for k in ("n_int16", "n_int32", "n_recomputed", "padded_cells", "n_launches",
                                                   "n_profiled")},
    }
    if main_res.get("end_to_end"):
        line["end_to_end"] = main_res["end_to_end"]
    if main_res.get("one_to_one"):
        line["one_to_one_same_pairs"] = {k: (round(v, 4) if isinstance(v, float) else v)
                                         for k, v in main_res["one_to_one"].items()}
    if main_res.get("autotuned"):
        line["autotuned"] = main_res["autotuned"]
    if mixed is not None:
        line["mixed_learned_dispatch"] = mixed
    if uneven is not None:
        line["uneven_1m_autotuned_graph"] = {
            "gcups": round(uneven["gcups"], 2), "ms_per_step": round(uneven["ms_per_step"], 3),
            "pairs": uneven["pairs"], "cells": uneven["cells"], "autotuned": uneven["autotuned"],
            "roofline_frac": round(uneven["achieved_tops"] / uneven["peak_tops"], 4),
            "n_launches": uneven["stats"]["n_launches"], "plan_ms": round(uneven["stats"]["plan_ms"], 1),
            "end_to_end": uneven.get("end_to_end")}
    if otm is not None:
        line["protein_1xn_query_profile"] = {
            "gcups": round(otm["gcups"], 2), "ms_per_step": round(otm["ms_per_step"], 4),
            "roofline_frac": round(otm["achieved_tops"] / otm["peak_tops"], 4),
            "one_to_one_ms_per_step": round(otm["one_to_one"]["ms_per_step"], 4),
            "speedup_of_profile": round(otm["one_to_one"]["speedup_of_profile"], 3),
            "scores_identical": otm["one_to_one"]["scores_identical"], "n_profiled": otm["stats"]["n_profiled"]}
    if extra is not None:
        line["dna150_global_int16"] = {"gcups": round(extra["gcups"], 2), "ms_per_step": round(extra["ms_per_step"], 4),
                                       "kernel_ms": round(extra["main_kernel_ms"], 4),
                                       "roofline_frac": round(extra["achieved_tops"] / extra["peak_tops"], 4)}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
