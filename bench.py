#!/usr/bin/env python3
"""Benchmark of the pairwise distance stage on one MI355X.

A step is one full pass of the pairwise stage (forward pass, reverse pass,
Myers-Miller recursion, identity count) over every pair of the workload, with
the sequences already resident in HBM.  The default workload is 1000 synthetic
protein sequences of length ~1523 (499,500 pairs, Gonnet 250, gap open 10.0,
extension 0.1).  The reported value is the batched kernel's GCUPS (one
persistent launch over a length-binned device work queue, the stages replayed
as a hipGraph); the plain one-wave-per-pair kernel is timed beside it on the
same input, and both must produce identical counts.  The same line carries the
uneven-length batching workload (2000 sequences, lognormal lengths), the
progressive stage, the k-tuple distances and the serving loop as sub-objects.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cuda-clustalw_amd"))

METRIC = "pairwise GCUPS (1000x1523), batched vs plain; serving p50/p99 latency"
PEAK_INT32_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12   # CUs x SIMDs x lanes x clock (Tops/s)
PEAK_HBM_GBS = 8000.0
OPS_PER_CELL = 10                                 # int32 ops per Gotoh cell update
PROFILE_OPS_PER_CELL = 10 + 2 * 27                # with prfScore's 27 multiply-adds


def ref_cells(lens, pairs, info):
    """Cell updates the reference CPU code performs (forward + reverse up to its
    early exit + every Myers-Miller node), from the kernels' per-pair cells."""
    n = lens[pairs[:, 0]].astype(np.int64)
    m = lens[pairs[:, 1]].astype(np.int64)
    ms, se1, se2, sb1, sb2 = (info[:, k].astype(np.int64) for k in range(5))
    rev = np.where(ms > 0, (se1 - sb1) * se2 + (se2 - sb2 + 1), 0)
    return n * m, rev, info[:, 7].view(np.uint32).astype(np.int64)


def host_cpu():
    """The host CPU the baseline ran on: model name, logical CPUs in the
    machine and CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline(codes, offs, lens, pairs, seconds, threads=None):
    """The oracle (C restatement of the reference) on a bounded pair sample."""
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O  # checker/baseline only
    sc = O.Scoring()
    if threads is None:
        threads = min(16, os.cpu_count() or 1)
    # calibrate with a small sample, then size the sample to ~`seconds`
    stride = max(1, len(pairs) // 64)
    probe = pairs[::stride][:64]
    t = time.perf_counter()
    O.pairwise_pairs(codes, offs, lens, probe, sc, nthreads=threads)
    dt = max(time.perf_counter() - t, 1e-3)
    per_pair = dt / len(probe)
    want = int(max(64, min(len(pairs), seconds / per_pair)))
    stride = max(1, len(pairs) // want)
    sample = pairs[::stride]
    t = time.perf_counter()
    O.pairwise_pairs(codes, offs, lens, sample, sc, nthreads=threads)
    dt = time.perf_counter() - t
    cells = (lens[sample[:, 0]].astype(np.int64) * lens[sample[:, 1]]).sum()
    return {"value": round(float(cells) / dt / 1e9, 4), "unit": "GCUPS", "cores": threads,
            "kind": "port", **host_cpu(),
            "sample": "every %dth pair of the workload (%d pairs, %.1f s, %d OpenMP threads)"
                      % (stride, len(sample), dt, threads)}


def small_stage(cpu: bool):
    """The 100 x ~97 workload (4,950 pairs) that the reference's CPU code runs in
    seconds: batched and plain GCUPS and cells/s, and the oracle on all pairs
    with one thread and with all cores."""
    import cwhip
    from cwhip import seqgen
    fam = seqgen.workload("small97")
    codes, offs, lens = fam.packed()
    n = len(lens)
    iu = np.triu_indices(n, 1)
    pairs = np.stack([iu[0], iu[1]], axis=1).astype(np.int32)
    conv = int((lens[pairs[:, 0]].astype(np.int64) * lens[pairs[:, 1]]).sum())
    ctx = cwhip.Context(0)
    ctx.load_sequences(codes, offs, lens)
    ctx.stage(cwhip.pair_params(max_len=int(lens.max()) + 1), pairs)
    out = {"workload": "small97", "sequences": n, "pairs": int(len(pairs)), "conventional_cells": conv}
    for name, k in (("batched", cwhip.KERNEL_BATCHED), ("plain", cwhip.KERNEL_PLAIN)):
        ctx.run_staged(kernel=k)
        ms, kms = ctx.run_staged(kernel=k, reps=5)
        if name == "batched":
            info = ctx.fetch_info()
            fwd, rev, mm = ref_cells(lens, pairs, info)
            upd = int(fwd.sum() + rev.sum() + mm.sum())
            out["reference_cell_updates"] = upd
            out["roofline"] = {"bound": "valu", "achieved": round(upd * OPS_PER_CELL / (kms * 1e-3) / 1e12, 3),
                               "peak": PEAK_INT32_TOPS, "unit": "Tops/s (int32)",
                               "frac": round(upd * OPS_PER_CELL / (kms * 1e-3) / 1e12 / PEAK_INT32_TOPS, 4)}
            out[name] = {"gcups": round(conv / (ms * 1e-3) / 1e9, 3), "ms": round(ms, 3),
                         "cells_per_s": round(upd / (ms * 1e-3), 1)}
        else:
            out[name] = {"gcups": round(conv / (ms * 1e-3) / 1e9, 3), "ms": round(ms, 3)}
    ctx.close()
    if cpu:
        out["cpu_baseline"] = cpu_baseline(codes, offs, lens, pairs, 60.0)
        out["cpu_baseline_1core"] = cpu_baseline(codes, offs, lens, pairs, 60.0, threads=1)
    return out


def batching_stage(cpu: bool, cpu_seconds: float):
    """The batching workload: 2000 sequences with lognormal(ln 400, 0.8) lengths
    clipped to [50, 3000] (1,999,000 pairs of very uneven size).  The autotuner
    picks the batched kernel's shapes first (its cost is reported, outside the
    timed region); then plain and batched each run the whole pair list and must
    give identical counts."""
    import cwhip
    from cwhip import seqgen
    fam = seqgen.workload("uneven")
    codes, offs, lens = fam.packed()
    n = len(lens)
    iu = np.triu_indices(n, 1)
    pairs = np.stack([iu[0], iu[1]], axis=1).astype(np.int32)
    conv = int((lens[pairs[:, 0]].astype(np.int64) * lens[pairs[:, 1]]).sum())
    ctx = cwhip.Context(0)
    params = cwhip.pair_params(max_len=int(lens.max()) + 1)
    ctx.load_sequences(codes, offs, lens)
    t0 = time.perf_counter()
    table, tune_ms = ctx.autotune(params, pairs)
    tune_wall = time.perf_counter() - t0
    ctx.stage(params, pairs)
    ctx.run_staged(kernel=cwhip.KERNEL_BATCHED)                    # warmup (captures the graph)
    ms_b, kms_b = ctx.run_staged(kernel=cwhip.KERNEL_BATCHED, reps=2)
    cnt_b = ctx.fetch_counts()
    info = ctx.fetch_info()
    fwd, rev, mm = ref_cells(lens, pairs, info)
    upd = int(fwd.sum() + rev.sum() + mm.sum())
    ms_p, _ = ctx.run_staged(kernel=cwhip.KERNEL_PLAIN)
    same = bool(np.array_equal(ctx.fetch_counts(), cnt_b))
    ctx.close()
    achieved = upd * OPS_PER_CELL / (kms_b * 1e-3) / 1e12
    out = {"workload": "uneven", "sequences": n, "pairs": int(len(pairs)),
           "length_mean": round(float(lens.mean()), 1), "length_min": int(lens.min()),
           "length_max": int(lens.max()), "conventional_cells": conv, "reference_cell_updates": upd,
           "batched": {"gcups": round(conv / (ms_b * 1e-3) / 1e9, 3), "ms": round(ms_b, 3),
                       "cells_per_s": round(upd / (ms_b * 1e-3), 1)},
           "plain": {"gcups": round(conv / (ms_p * 1e-3) / 1e9, 3), "ms": round(ms_p, 3),
                     "identical_to_batched": same},
           "batched_over_plain": round(ms_p / ms_b, 3),
           "autotune": {"table": table.as_dict(), "device_ms": round(tune_ms, 1),
                        "wall_s": round(tune_wall, 2),
                        "classes": "longer sequence <=128, <=256, <=512, <=1024, <=2048, longer"},
           "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_INT32_TOPS,
                        "unit": "Tops/s (int32)", "frac": round(achieved / PEAK_INT32_TOPS, 4)}}
    if cpu:
        out["cpu_baseline"] = cpu_baseline(codes, offs, lens, pairs, cpu_seconds)
    return out


def progressive_stage(cpu: bool):
    """The progressive stage at the progressive workload's shape (1000 sequences
    of ~1002 residues): device pairwise -> NJ tree -> guide -> device
    multiSeqAlign -> CLUSTAL text, all on the library.  The multiSeqAlign
    stage is timed on its own (twice, the faster run kept) and, with `cpu`, the
    oracle's single-threaded multiSeqAlign runs on the same guide as the CPU
    baseline and must give the same rows."""
    import cwhip
    from cwhip import pipeline, seqgen
    fam = seqgen.workload("prog1002")
    names, seqs = fam.names, fam.seqs
    ctx = cwhip.Context(0)
    tm = {}
    dnd, aln, rows, order = pipeline.align(ctx, names, seqs, dna=False, timings=tm)
    w, sim, steps = cwhip.guide_from_tree(dnd, names)
    best = None
    for _ in range(2):
        t = time.perf_counter()
        rows2, order2, st = ctx.msa(seqs, sim, w, steps)
        el = time.perf_counter() - t
        if best is None or el < best[0]:
            best = (el, st)
    el, st = best
    out = {"workload": "prog1002", "sequences": len(seqs),
           "length_mean": round(float(np.mean([len(x) for x in seqs])), 1),
           "profile_alignments": st["calls"], "device_batches": st["batches"],
           "mm_cell_updates": int(st["cells"]), "msa_s": round(el, 4),
           "msa_device_ms": round(st["device_ms"], 2),
           "mm_cells_per_s": round(st["cells"] / el, 1),
           "pipeline_s": {k: round(v, 4) for k, v in tm.items() if k.endswith("_s")},
           "aln_bytes": len(aln)}
    # algorithmic int32 ops: 10 per Myers-Miller cell update, 54 (27 multiply-adds)
    # per prfScore entry, each entry computed once per profile pair
    ops = st["cells"] * OPS_PER_CELL + st["prf_entries"] * 54
    dev_s = st["device_ms"] * 1e-3
    ach = ops / dev_s / 1e12
    # the per-cell form: every Myers-Miller cell update charged 10 + 2 x 27 ops,
    # as if prfScore were recomputed at each update (what the reference does)
    ach64 = st["cells"] * PROFILE_OPS_PER_CELL / dev_s / 1e12
    traffic, tsrc = None, None
    tfile = os.path.join(HERE, "profiles", "pmc_traffic_prog1002.json")
    if os.path.exists(tfile):
        t = json.load(open(tfile))
        traffic, tsrc = int(t["traffic_bytes_per_msa"]), t["summary"]
    out["roofline"] = {"bound": "valu", "achieved": round(ach, 4), "peak": PEAK_INT32_TOPS,
                       "unit": "Tops/s (int32)", "frac": round(ach / PEAK_INT32_TOPS, 5),
                       "achieved_64_ops_per_cell": round(ach64, 4),
                       "frac_64_ops_per_cell": round(ach64 / PEAK_INT32_TOPS, 5),
                       "traffic": traffic, "traffic_source": tsrc,
                       "hbm_gbs": round(traffic / dev_s / 1e9, 1) if traffic else None,
                       "note": "device time of the profile kernels (cw_prf_cols, cw_prf_gapcols, "
                               "cw_prf_score, cw_prf_wide, cw_prf_level, cw_prf_mm, "
                               "cw_prf_local) summed over the "
                               "dependency-level batches; achieved = 10 ops per Myers-Miller "
                               "cell update + 54 per prfScore entry (each computed once); "
                               "achieved_64_ops_per_cell = 64 ops per cell update; traffic = "
                               "FETCH_SIZE + WRITE_SIZE of those kernels per cwhip_msa"}
    if cpu:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle as O  # checker/baseline only
        t = time.perf_counter()
        r_rows, r_order, info = O.msa([x.astype(np.int32) for x in seqs], sim, w, steps,
                                      O.msa_params(False))
        dt = time.perf_counter() - t
        same = bool(np.array_equal(order2, r_order) and
                    all(np.array_equal(rows2[i], r_rows[i].astype(np.uint8)) for i in range(len(seqs))))
        out["cpu_baseline"] = {"value": round(st["cells"] / dt, 1), "unit": "Myers-Miller cell updates/s",
                               "cores": 1, "kind": "port", "seconds": round(dt, 2), **host_cpu(),
                               "sample": "the whole multiSeqAlign stage (999 profile alignments), 1 thread",
                               "identical_rows": same}
        out["speedup_vs_cpu"] = round(dt / el, 2)
    ctx.close()
    return out


def ktuple_stage(cpu: bool):
    """k-tuple distances (FastPairwiseAlign, the -QUICKTREE path) over all
    499,500 pairs of the 1000 x ~1523 workload: the kernel time of one
    cwhip_fast_pairwise_range (HIP events on the context's stream), pairs/s; 300
    pairs spread over the matrix compared bit for bit with the oracle; the CPU
    baseline is the oracle on one thread over the first 180 sequences' 16,110
    pairs (about 9 s)."""
    import cwhip
    from cwhip import seqgen
    fam = seqgen.workload("large1523")
    codes, offs, lens = fam.packed()
    n = len(lens)
    ctx = cwhip.Context(0)
    ctx.load_sequences(codes, offs, lens)
    ctx.fast_pairwise_range(0, 8, 0, 8)                 # warm-up
    t = time.perf_counter()
    pairs, dist = ctx.fast_pairwise_range(0, n, 0, n)
    wall = time.perf_counter() - t
    ms = ctx.fast_last_ms()
    ctx.close()
    out = {"workload": "large1523", "sequences": n, "pairs": int(len(dist)),
           "params": "protein ktup 1, window 5, window gap 3, 5 top diagonals, percent",
           "kernel_ms": round(ms, 1), "wall_s": round(wall, 3),
           "pairs_per_s": round(len(dist) / (ms * 1e-3), 1)}
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O  # checker/baseline only
    rng = np.random.default_rng(11)
    pick = np.sort(rng.choice(n, 25, replace=False))    # 300 pairs among 25 sequences
    sub = [np.asarray(fam.seqs[i], np.int32) for i in pick]
    po, do = O.fast_pairwise(sub, max_aln=2 * int(lens.max()))
    index = {(int(a), int(b)): k for k, (a, b) in enumerate(pairs)}
    got = np.array([dist[index[(int(pick[a]), int(pick[b]))]] for a, b in po])
    out["identical_sample"] = bool(np.array_equal(got.view(np.uint64), do.view(np.uint64)))
    # the same run at ~500 residues (1000 sequences, 499,500 pairs), where more
    # waves fit a CU's LDS than at ~1523
    fam5 = seqgen.workload("mid498")
    c5, o5, l5 = fam5.packed()
    ctx = cwhip.Context(0)
    ctx.load_sequences(c5, o5, l5)
    ctx.fast_pairwise_range(0, 8, 0, 8)
    p5, d5 = ctx.fast_pairwise_range(0, len(l5), 0, len(l5))
    ms5 = ctx.fast_last_ms()
    ctx.close()
    out["mid498"] = {"sequences": len(l5), "pairs": int(len(d5)), "length_mean": round(float(l5.mean()), 1),
                     "kernel_ms": round(ms5, 1), "pairs_per_s": round(len(d5) / (ms5 * 1e-3), 1)}
    if cpu:
        head = [np.asarray(s, np.int32) for s in fam.seqs[:180]]
        t = time.perf_counter()
        _, dh = O.fast_pairwise(head, max_aln=2 * int(lens.max()))
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(len(dh) / dt, 1), "unit": "pairs/s", "cores": 1,
                               "kind": "port", **host_cpu(),
                               "sample": "the first 180 sequences' %d pairs, %.1f s, one thread" % (len(dh), dt)}
        out["speedup_vs_cpu"] = round(out["pairs_per_s"] / out["cpu_baseline"]["value"], 1)
    return out


def serving_stage(rate: float, count: int, seed: int = 20260105, slots: int = 4):
    """Poisson arrivals at `rate` requests/s of `count` serving requests (N ~ U[10,500],
    L ~ lognormal(ln 300, 0.7) clipped [30, 2000]; 80% distances, 20% distances +
    progressive alignment) into the library's request loop (`slots` device slots,
    admission by device footprint).  Latency is submission to completion; throughput counts
    pairwise cells and Myers-Miller cell updates over the whole window.  Three
    requests are then re-run alone on the direct path and compared."""
    import cwhip
    from cwhip import pipeline, seqgen
    reqs = [seqgen.serving_request(seed, i) for i in range(count)]
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=count)
    arrive = np.cumsum(gaps) - gaps[0]
    srv = cwhip.Server(0, slots=slots, hbm_budget=128 << 30)
    jobs = []
    t0 = time.perf_counter()
    for (kind, fam), ta in zip(reqs, arrive):
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        jobs.append(srv.submit(kind, fam.seqs, fam.names))
    for j in jobs:
        srv.wait(j, 600)
    times = np.array([srv.times(j) for j in jobs])
    lat = times[:, 2] - times[:, 0]
    window = times[:, 2].max() - times[:, 0].min()
    cells = times[:, 3].sum()
    st = srv.stats()
    kinds = [k for k, _ in reqs]
    # bit-exactness spot check against the direct path
    same = True
    ctx = cwhip.Context(0)
    for i in (0, count // 2, count - 1):
        kind, fam = reqs[i]
        codes, offs, lens = fam.packed()
        ctx.load_sequences(codes, offs, lens)
        params = cwhip.pair_params(max_len=int(lens.max()))
        _, _, dist = ctx.pairwise_range(params, 0, len(lens), 0, len(lens), kernel=cwhip.KERNEL_PLAIN)
        same &= bool(np.array_equal(dist.view(np.uint32), srv.distances(jobs[i]).view(np.uint32)))
        if kind == cwhip.JOB_ALIGNMENT:
            dnd, aln, _, _ = pipeline.align(ctx, fam.names, fam.seqs, False)
            same &= (dnd == srv.text(jobs[i], 0) and aln == srv.text(jobs[i], 1))
    ctx.close()
    srv.close()
    kinds_a = np.array(kinds)

    def pct(sel, q):
        return round(float(np.percentile(lat[sel], q)), 4) if sel.any() else None

    return {"rate_per_s": rate, "requests": count, "alignment_jobs": int(sum(kinds)),
            "distance_jobs": int(count - sum(kinds)),
            "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
            "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
            "latency_mean_s": round(float(lat.mean()), 4),
            "distance_jobs_p50_p99_s": [pct(kinds_a == 0, 50), pct(kinds_a == 0, 99)],
            "alignment_jobs_p50_p99_s": [pct(kinds_a == 1, 50), pct(kinds_a == 1, 99)],
            "throughput_gcups": round(cells / window / 1e9, 2), "window_s": round(float(window), 3),
            "max_concurrent_jobs": st["max_running"], "peak_device_bytes": st["peak_bytes"], "budget_bytes": st["budget"],
            "slots": st["slots"], "spot_check_identical_to_direct_path": bool(same)}


def copy_probe(nbytes):
    """One device-to-device copy of nbytes (the HIP runtime's blit kernel): the
    reference dispatch whose FETCH_SIZE / WRITE_SIZE counters scale the
    per-dispatch HBM traffic of the DP kernel."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


VALU_PROBE_OPS = ("add_max", "add", "max", "add3", "max3", "mov_dpp", "subrev", "dp_mix")


def valu_probe(trips, waves):
    """The int32 VALU calibration kernels (libcwcal.so), one launch per op
    kind and waves-per-CU figure: a known count of wave instructions, timed
    with HIP events. Their SQ_INSTS_VALU readings calibrate that counter, and
    the fastest of them is the measured issue ceiling the DP kernels are
    compared with."""
    import ctypes as C
    lib = C.CDLL(os.path.join(HERE, "cuda-clustalw_amd", "libcwcal.so"))
    lib.cwcal_valu.argtypes = [C.c_int, C.c_int, C.c_int] + [C.POINTER(C.c_double)] * 3
    for op, name in enumerate(VALU_PROBE_OPS):
        for w in waves:
            ms, instrs, ops = C.c_double(), C.c_double(), C.c_double()
            rc = lib.cwcal_valu(op, trips, w, C.byref(ms), C.byref(instrs), C.byref(ops))
            if rc != 0:
                raise RuntimeError("cwcal_valu failed (%d)" % rc)
            print(json.dumps({"valu_probe": {"op": name, "trips": trips, "waves_per_cu": w,
                                             "ms": round(ms.value, 4), "loop_wave_instrs": instrs.value,
                                             "lane_ops_per_s": ops.value / (ms.value * 1e-3)}}), flush=True)


def serving_all(args):
    """The serving stream at the stated rate, then the same mix at higher rates."""
    out = {"serving": serving_stage(args.serve_rate, args.serve_count)}
    load = [float(r) for r in args.serve_load.split(",") if r.strip()]
    if load:
        keys = ("rate_per_s", "latency_p50_s", "latency_p99_s", "throughput_gcups",
                "max_concurrent_jobs", "spot_check_identical_to_direct_path")
        out["serving_load"] = [{k: r[k] for k in keys}
                               for r in (serving_stage(rate, 120) for rate in load)]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="large1523")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-progressive", action="store_true")
    ap.add_argument("--no-ktuple", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--stage", choices=["all", "batching", "progressive", "serving", "ktuple"], default="all",
                    help="progressive / serving: run only that part and print its object")
    ap.add_argument("--serve-rate", type=float, default=8.0, help="serving: requests per second")
    ap.add_argument("--serve-count", type=int, default=80, help="serving: number of requests")
    ap.add_argument("--serve-load", default="32,64",
                    help="serving: further rates (comma-separated, 120 requests each) for the "
                         "latency/throughput curve toward saturation; empty for none")
    ap.add_argument("--max-pairs", type=int, default=0, help="debug: cap the pair count")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run one device-to-device copy of this many bytes (counter calibration)")
    ap.add_argument("--valu-probe", type=int, default=0,
                    help="only run the int32 VALU calibration kernel for this many loop trips")
    ap.add_argument("--valu-probe-waves", default="4,8,16,32",
                    help="waves per CU of the VALU probe launches (comma-separated multiples of 4)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.valu_probe:
        valu_probe(args.valu_probe, [int(w) for w in args.valu_probe_waves.split(",")])
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only (the pairwise stage would shard by pair lists, "
              "one context per GPU); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.stage == "progressive":
        print(json.dumps({"progressive": progressive_stage(not args.no_cpu)}), flush=True)
        return
    if args.stage == "batching":
        print(json.dumps({"batching": batching_stage(not args.no_cpu, 5.0)}), flush=True)
        return
    if args.stage == "ktuple":
        print(json.dumps({"ktuple": ktuple_stage(not args.no_cpu)}), flush=True)
        return
    if args.stage == "serving":
        print(json.dumps(serving_all(args)), flush=True)
        return

    import cwhip
    from cwhip import seqgen

    fam = seqgen.workload(args.workload)
    codes, offs, lens = fam.packed()
    nseq = len(lens)
    iu = np.triu_indices(nseq, 1)
    pairs = np.stack([iu[0], iu[1]], axis=1).astype(np.int32)   # reference visiting order
    if args.max_pairs:
        pairs = pairs[:args.max_pairs]
    conv_cells = int((lens[pairs[:, 0]].astype(np.int64) * lens[pairs[:, 1]]).sum())

    ctx = cwhip.Context(0)
    params = cwhip.pair_params(max_len=int(lens.max()) + 1)
    ctx.load_sequences(codes, offs, lens)
    ctx.stage(params, pairs)

    def timed(kernel, steps, warmup):
        for _ in range(warmup):
            ctx.run_staged(kernel=kernel)
        tot, kern = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            ms, kms = ctx.run_staged(kernel=kernel)
            tot.append(ms)
            kern.append(kms)
        wall = time.perf_counter() - t0
        return tot, kern, wall

    tot_b, kern_b, wall_b = timed(cwhip.KERNEL_BATCHED, args.steps, args.warmup)
    cnt_b = ctx.fetch_counts()
    info = ctx.fetch_info()
    fwd, rev, mm = ref_cells(lens, pairs, info)
    upd_cells = int(fwd.sum() + rev.sum() + mm.sum())
    ms_step = float(np.mean(tot_b))
    kms = float(np.mean(kern_b))
    gcups = conv_cells / (ms_step * 1e-3) / 1e9

    out = {
        "metric": METRIC, "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded tree-evolved protein family)",
        "config": {"workload": args.workload, "sequences": nseq, "pairs": int(len(pairs)),
                   "length_mean": round(float(lens.mean()), 1), "length_min": int(lens.min()),
                   "length_max": int(lens.max()), "scoring": "Gonnet250 x10, pw gap 10.0/0.1",
                   "conventional_cells": conv_cells, "reference_cell_updates": upd_cells},
        "cells_per_s": round(upd_cells / (ms_step * 1e-3), 1),
        "batched": {"gcups": round(gcups, 3), "ms": round(ms_step, 3), "kernel_ms": round(kms, 3),
                    "wall_s": round(wall_b, 3)},
    }
    achieved = upd_cells * OPS_PER_CELL / (kms * 1e-3) / 1e12
    traffic, tsrc, issue = None, None, None
    tfile = os.path.join(HERE, "profiles", "pmc_traffic_%s.json" % args.workload)
    if os.path.exists(tfile) and not args.max_pairs:
        t = json.load(open(tfile))
        traffic, tsrc = int(t["traffic_bytes"]), t["summary"]
        if "valu_instr_per_cycle_per_simd" in t:
            # SQ_INSTS_VALU of the same launch, against the calibration probe
            # that issues the forward loop's opcode mix at the same occupancy
            issue = {"wave_instr_per_cycle_per_simd": t["valu_instr_per_cycle_per_simd"],
                     "clock_ghz": t["clock_ghz"],
                     "lane_slots_per_cell_update": round(t["valu_wave_instrs"] * 64 / upd_cells, 2),
                     "probe_dp_mix_16_waves": t["probe_dp_mix_instr_per_cycle_per_simd"]["16"],
                     "frac_of_probe": round(t["valu_instr_per_cycle_per_simd"] /
                                            t["probe_dp_mix_instr_per_cycle_per_simd"]["16"], 3),
                     "source": t["calibration"]}
    out["roofline"] = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_INT32_TOPS,
                       "unit": "Tops/s (int32)", "frac": round(achieved / PEAK_INT32_TOPS, 4),
                       "traffic": traffic, "traffic_source": tsrc,
                       "hbm_gbs": round(traffic / (kms * 1e-3) / 1e9, 1) if traffic else None,
                       "valu_issue": issue,
                       "note": "integer DP: %d int32 ops per reference cell update; "
                               "the HBM roof (%.0f GB/s) is not the bound" % (OPS_PER_CELL, PEAK_HBM_GBS)}

    if not args.no_plain:
        tot_p, kern_p, wall_p = timed(cwhip.KERNEL_PLAIN, args.steps, min(args.warmup, 1))
        cnt_p = ctx.fetch_counts()
        ms_p = float(np.mean(tot_p))
        out["plain"] = {"gcups": round(conv_cells / (ms_p * 1e-3) / 1e9, 3), "ms": round(ms_p, 3),
                        "kernel_ms": round(float(np.mean(kern_p)), 3),
                        "identical_to_batched": bool(np.array_equal(cnt_p, cnt_b))}
        out["batched_over_plain"] = round(ms_p / ms_step, 3)
    # the later stages make their own contexts; a context left open keeps its
    # stream, and with it one of the process's hardware queues
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(codes, offs, lens, pairs, args.cpu_seconds)
        out["cpu_baseline_1core"] = cpu_baseline(codes, offs, lens, pairs, 5.0, threads=1)
    if not args.no_small:
        out["small"] = small_stage(not args.no_cpu)
    if not args.no_batching:
        out["batching"] = batching_stage(not args.no_cpu, 5.0)
    if not args.no_progressive:
        out["progressive"] = progressive_stage(not args.no_cpu)
    if not args.no_ktuple:
        out["ktuple"] = ktuple_stage(not args.no_cpu)
    if not args.no_serving:
        out.update(serving_all(args))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
