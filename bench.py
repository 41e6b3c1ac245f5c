#!/usr/bin/env python3
"""Benchmark of the all-pairs motif-extraction hot path on one MI355X.

Metric: GCUPS (giga cell updates per second; one cell = one (k-factor of x, text position
of y) update, cells(pair) = (m-k+1)*n over all ordered pairs) of the batched persistent
kernel on the 1200 x 1000-nt Hamming workload (k=12, e=2), beside the plain
one-launch-per-pair GPU form, the oracle timed on host cores, and (--serving) p50/p99 job
latency of the continuous-batching request loop.

A step is one full all-pairs extraction over the resident, packed input (accumulators
cleared, persistent kernel, completion).  One JSON line is printed on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks
HBM_PEAK_GBS = 8000.0
# The published FP32 vector rate, 157.3 TFLOP/s, read as 32-bit lane operations (an FMA counts
# two flops): 78.6 T lane-ops/s = 256 CUs x 128 lanes x 2.4 GHz.  The kernels' issued VALU
# lane-ops (SQ_INSTS_VALU x 64) are measured against it.
VALU_LANE_OPS_PEAK_TOPS = 157.3 / 2
# Algorithmic view: the reference recurrence's int32 ops per cell (Hamming 8, edit 22) against
# one 32-bit op per lane per clock (256 x 64 x 2.4 GHz = 39.3 T).  The kernels are bit-sliced
# (32 cells per lane op) and, for Hamming, compute each unordered pair once, so this figure
# exceeds 1; it is reported beside the roofline, never as its fraction.
VALU_INT32_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
OPS_PER_CELL = {0: 8.0, 1: 22.0}
# Issue-rate ceilings measured on MI355X with a dependency-free micro-benchmark of one
# instruction kind at 32 waves per CU (T lane-ops/s); see profiles/valu_issue_rate.md.
VALU_CEILINGS_TOPS = {"v_xor_b32": 61.9, "v_bitop3_b32": 59.3, "v_alignbit_b32": 37.9}
PROFILES = os.path.join(ROOT, "profiles")


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="dna1200",
                    choices=["dna50", "dna1200", "prot500", "struct1000", "long4m", "hsapiens1200"],
                    help="dna50: 50 x 500 nt Hamming k=10 e=1; dna1200: 1200 x 1000 nt Hamming k=12 "
                         "e=2 (the metric's workload); prot500: 500 proteins of U[200,2000], edit k=8 "
                         "e=1; struct1000: 1000 x 600 nt, structured 8+[10,30]+8; long4m: one 4 Mnt "
                         "sequence, Hamming k=10 e=2; hsapiens1200: the reference's 1200 H. sapiens "
                         "genes x 1000 nt, Hamming k=12 e=2")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--plain-x", type=int, default=4,
                    help="x sequences timed in the plain one-launch-per-pair form")
    ap.add_argument("--cpu-x", type=int, default=0,
                    help="x sequences in the CPU baseline sample (0: sized for ~12 s)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving stream (p50/p99 job latency)")
    ap.add_argument("--serving-jobs", type=int, default=256)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES bytes (calibrates the "
                         "FETCH_SIZE/WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_probe(nbytes: int):
    import torch
    x = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}))


def traffic_from_profile(kernel: str, workload_tag: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the counter reads for a
    plain device copy of known size).  None when no summary covers this workload."""
    path = os.path.join(PROFILES, f"{kernel}_{workload_tag}_traffic.txt")
    if not os.path.exists(path):
        return None
    kv = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            key, val = line.split(":", 1)
            kv[key.strip()] = val.strip()
    try:
        cb = float(kv["copy_bytes"])
        fetch = float(kv["fetch_kib"]) * 1024 * cb / (float(kv["copy_fetch_kib"]) * 1024)
        write = float(kv["write_kib"]) * 1024 * cb / (float(kv["copy_write_kib"]) * 1024)
    except (KeyError, ValueError):
        return None
    return {"bytes": round(fetch + write), "fetch_bytes": round(fetch), "write_bytes": round(write),
            "source": os.path.relpath(path, ROOT)}


def valu_from_profile(kernel: str, workload_tag: str):
    """VALU wave-instructions of one dispatch of `kernel` on this workload, from the committed
    rocprofv3 --pmc summary (the SQ_INSTS_VALU row), or None."""
    path = os.path.join(PROFILES, f"{kernel}_{workload_tag}_pmc.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 2 and cells[0] == "SQ_INSTS_VALU":
            try:
                return {"valu_insts": int(cells[1].replace(",", "")), "source": os.path.relpath(path, ROOT)}
            except ValueError:
                return None
    return None


def calibration(distance: int, threads: int):
    """The oracle's time over motexOMP's on the same input and host (oracle/calibrate.py),
    for this distance at the nearest thread count, or None."""
    path = os.path.join(ROOT, "oracle", "calibration_hsapiens120.json")
    if not os.path.exists(path):
        return None
    cal = json.load(open(path))
    runs = [r for r in cal["runs"] if r["distance"] == distance]
    if not runs:
        return None
    r = min(runs, key=lambda r: abs(r["threads"] - threads))
    return {"ratio_oracle_over_motexomp": r["ratio_oracle_over_motexomp"], "threads": r["threads"],
            "input": cal["input"], "host": cal["host"], "source": "oracle/calibration_hsapiens120.json"}


def cpu_baseline(w, gpu_u, gpu_v, off, nx):
    """Times the oracle (OpenMP over x, host cores) on x sequences [0, nx) against every
    y and checks the GPU rows of the same sample bit for bit."""
    from oracle import oracle as O
    O.build()
    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    ou, ov, ooff, _ = O.all_pairs(w.seqs, w.distance, w.k, w.e, boxes=w.boxes, threads=cores,
                                  x_range=(0, nx))
    dt = time.perf_counter() - t0
    end = int(off[nx])
    parity = bool(np.array_equal(ou[:end], gpu_u[:end]) and np.array_equal(ov[:end], gpu_v[:end]))
    L = np.array(w.lens, dtype=np.int64)
    cells = float(np.sum(L[:nx] - w.k + 1) * np.sum(L))
    out = {"value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": cores, "kind": "port",
           "sample": f"x sequences 0..{nx - 1} against all {len(w.seqs)} texts "
                     f"({cells:.3e} cells, {dt:.1f} s, OpenMP over x)",
           "parity_on_sample": parity}
    return with_calibration(out, w.distance if not w.boxes else None, cores)


def with_calibration(out, distance, cores):
    """Adds the motexOMP-equivalent rate: the port's rate times its time ratio to motexOMP on
    the reference's own timed input (structured runs have no recorded motexOMP time)."""
    cal = calibration(distance, cores) if distance is not None else None
    if cal:
        out["calibration"] = cal
        out["motexomp_equivalent_value"] = round(out["value"] * cal["ratio_oracle_over_motexomp"], 4)
    return out


def cpu_baseline_prefix(ctx, w, params, seconds=12.0):
    """One long sequence: the oracle (like the reference) gives a single pair one thread, so
    a prefix sized for ~seconds of CPU work is timed, and the GPU checked on the same prefix."""
    from oracle import oracle as O
    O.build()
    P = int(min(len(w.seqs[0]), (seconds * 1.0e9) ** 0.5))
    pre = [w.seqs[0][:P]]
    t0 = time.perf_counter()
    ou, ov, _, _ = O.all_pairs(pre, w.distance, w.k, w.e, threads=1)
    dt = time.perf_counter() - t0
    gu, gv, _ = ctx.extract_all_pairs(pre, params)
    cells = float((P - w.k + 1) * P)
    out = {"value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": 1, "kind": "port",
           "sample": f"self pair of the first {P} symbols ({cells:.3e} cells, {dt:.1f} s; one "
                     "thread, as the reference gives a single sequence one thread)",
           "parity_on_sample": bool(np.array_equal(ou, gu) and np.array_equal(ov, gv))}
    return with_calibration(out, w.distance, 1)


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU's share only; --gpus {a.gpus} is not supported: "
              "this build runs on one GPU (the x sequences would be split across GPUs, one "
              "process each, with no collective).", file=sys.stderr)
        sys.exit(2)
    import motex_amd
    from motex_amd import MotexParams, synth

    if a.workload == "hsapiens1200":
        from motex_amd import refdata
        w = refdata.hsapiens(max(2, int(round(1200 * a.scale))))
    else:
        w = synth.WORKLOADS[a.workload](a.scale)
    params = MotexParams(k=w.k, e=w.e, distance=w.distance, alphabet=w.alphabet,
                         boxes=list(w.boxes), long_mode=w.long_mode)
    ctx = motex_amd.Context(0)
    t_up = time.perf_counter()
    batch = ctx.batch(w.seqs, params)
    t_up = time.perf_counter() - t_up
    for _ in range(a.warmup):
        batch.run()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        st = batch.run()
        kernel_ms.append(st["main_kernel_ms"])
    t1 = time.perf_counter()
    st = batch.stats()
    cells = st["cells"]
    wall_step = (t1 - t0) / a.steps
    value = cells / wall_step / 1e9
    gpu_u, gpu_v, off = batch.download()

    # plain one-launch-per-pair form on a sample of x sequences (every y each)
    nplain = min(a.plain_x, len(w.seqs))
    L = np.array(w.lens, dtype=np.int64)
    plain = None
    if nplain > 0:
        batch.clear()
        tp = time.perf_counter()
        pst = batch.run_plain(0, nplain)
        tp = time.perf_counter() - tp
        plain_cells = float(np.sum(np.maximum(L[:nplain] - w.k + 1, 0)) * np.sum(L[L >= w.k]))
        plain = {"value": round(plain_cells / tp / 1e9, 4), "unit": "GCUPS",
                 "device_value": round(plain_cells / max(pst["kernel_ms"], 1e-6) * 1e3 / 1e9, 4),
                 "sample": f"x 0..{nplain - 1} x all texts, {nplain * len(w.seqs)} launches"}
        pu, pv, _ = batch.download()
        end = int(off[nplain])
        plain["parity_with_batched"] = bool(np.array_equal(pu[:end], gpu_u[:end]) and
                                            np.array_equal(pv[:end], gpu_v[:end]))

    kms = float(np.mean(kernel_ms))
    ops = OPS_PER_CELL[w.distance]
    algo = cells * ops / (kms / 1e3) / 1e12
    kernel = (("struct_range" if st["kernel_variant"] == 1 else "struct_kernel") if w.boxes else
              "hd_batched_kernel" if w.distance == 0 else
              {2: "ed_slice_kernel", 3: "ed_slice_pipe_kernel"}.get(st["kernel_variant"], "ed_kernel"))
    tr = traffic_from_profile(kernel, a.workload) if a.scale == 1.0 else None
    vp = valu_from_profile(kernel, a.workload) if a.scale == 1.0 else None
    # issued VALU lane-ops of one dispatch (rocprof SQ_INSTS_VALU x 64) over the live kernel time
    achieved = vp["valu_insts"] * 64 / (kms / 1e3) / 1e12 if vp else None
    roofline = {"bound": "valu", "achieved": round(achieved, 3) if vp else None,
                "peak": VALU_LANE_OPS_PEAK_TOPS, "unit": "T lane-ops/s",
                "frac": round(achieved / VALU_LANE_OPS_PEAK_TOPS, 4) if vp else None,
                "traffic": tr["bytes"] if tr else None,
                "traffic_detail": tr,
                "valu_insts_per_dispatch": vp,
                "kernel": kernel,
                "kernel_ms": round(kms, 4),
                "cells_per_s": round(cells / (kms / 1e3), 1),
                "swept_cells_per_s": round(st["swept_cells"] / (kms / 1e3), 1),
                "symmetric": bool(st["symmetric"]),
                "graph_replay": bool(st["graph_replay"]),
                "valu_ceilings_measured": VALU_CEILINGS_TOPS,
                "note": "achieved = issued VALU lane-ops of one dispatch (SQ_INSTS_VALU x 64, "
                        "from the committed rocprofv3 --pmc summary of this workload) / the "
                        "kernel time measured here with HIP events; peak = the published 157.3 "
                        "TFLOP/s FP32 vector rate read as 78.6 T lane-ops/s",
                "algorithmic": {"ops_per_cell": ops, "ops_per_s_T": round(algo, 3),
                                "int32_peak_T": VALU_INT32_PEAK_TOPS,
                                "ratio_to_int32_peak": round(algo / VALU_INT32_PEAK_TOPS, 4),
                                "note": "the reference recurrence's ops per cell over every "
                                        "ordered pair's cells; above 1 because one lane op "
                                        "advances 32 cells and Hamming pairs are computed once "
                                        "per unordered pair"}}

    out = {
        "metric": "GCUPS batched vs plain-per-pair on 1 GPU at k=12,e=2; p50/p99 job latency",
        "value": round(value, 3),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": ("the reference's data/hsapiens.1200_genes.fa (tests/golden/refdata)" if a.workload == "hsapiens1200"
                 else "synthetic (seeded SplitMix64 generator, motex_amd/synth.py)"),
        "config": {"workload": w.name, "num_seqs": len(w.seqs), "total_nt": int(L.sum()),
                   "cells": cells, "k": w.k, "e": w.e, "distance": w.distance,
                   "tile": {"blocks": st["tile_blocks"], "rows": st["tile_rows"],
                            "waves_per_cu": st["waves_per_cu"]},
                   "work_items": st["work_items"], "upload_s": round(t_up, 3),
                   **({"handed_pairs": st["handed_pairs"]} if w.boxes else {})},
        "roofline": roofline,
        "edit_semantics": ("MaxShift recurrence of the reference's in-tree code (functions.cc:437-447, "
                           "bitminmax 923-967); equality with libflasm's flasm_ed is unpinned"
                           if w.distance == 1 else None),
        "plain": plain,
        "speedup_vs_plain": round(value / plain["value"], 2) if plain and plain["value"] else None,
    }
    if not a.no_cpu and len(w.seqs) == 1:
        out["cpu_baseline"] = cpu_baseline_prefix(ctx, w, params)
    elif not a.no_cpu:
        nx = a.cpu_x
        if nx <= 0:
            cores = len(os.sched_getaffinity(0))
            per_x = float(np.mean(L) - w.k + 1) * float(L.sum())
            # the oracle's cells per second per core on the box's host (measured: Hamming 1.5e9,
            # structured Hamming 0.5e9, edit 0.13e9), so the sample takes about 12 s
            per_core = 0.13e9 if w.distance else (0.5e9 if w.boxes else 1.4e9)
            nx = int(max(1, min(len(w.seqs), 12.0 * per_core * cores / per_x)))
        out["cpu_baseline"] = cpu_baseline(w, gpu_u, gpu_v, off, nx)
    if not a.no_serving:
        from motex_amd import serving
        sv = serving.run_stream(ctx, njobs=a.serving_jobs)
        sv["workload"] = ("serving: jobs of 20-400 DNA/PROT sequences of 100-2000 symbols, k 8-14, "
                          "e 0-3, d 0/1, q 20-80, Poisson arrivals at 70% of measured capacity")
        out["serving"] = sv
        out["p50_job_latency_ms"] = sv["p50_latency_ms"]
        out["p99_job_latency_ms"] = sv["p99_latency_ms"]
    batch.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
