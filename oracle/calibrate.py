"""Calibration of the oracle's OpenMP all-pairs loop against the reference motexOMP.

Test/measurement infrastructure only.  The oracle is what the benchmark can time on the GPU
machine's host cores; the reference motexOMP needs the binary-only libflasm.so and cannot
run there.  This script times the oracle on the input the reference was timed on -- the first
120 sequences of the reference's H. sapiens dataset (120 x 1000 nt, 1.424e10 cells), Hamming
k=12 e=2 and edit k=12 e=2 -- at 1 and 8 threads on the same 8-core host (Intel Xeon @ 2.70
GHz), and writes the ratio oracle time / motexOMP time to oracle/calibration_hsapiens120.json.
The benchmark divides its oracle throughput by that ratio to state a motexOMP-equivalent
figure beside it.

motexOMP times on that input and host, from the header line ``Run on P proc(s) in X secs``
(extraction only, motex.cc:492 -> 831):
    -d 0 -k 12 -e 2 -q 25 -t 1   97.89 s
    -d 0 -k 12 -e 2 -q 25 -t 8   13.17 s
    -d 1 -k 12 -e 2 -q 25 -t 8   16.50 s

Usage: python -m oracle.calibrate   (about a minute on 8 cores)
"""
from __future__ import annotations

import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "calibration_hsapiens120.json")
REFERENCE_S = [
    {"distance": 0, "threads": 1, "motexomp_s": 97.89},
    {"distance": 0, "threads": 8, "motexomp_s": 13.17},
    {"distance": 1, "threads": 8, "motexomp_s": 16.50},
]


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main() -> int:
    from oracle import oracle as O
    from motex_amd import refdata
    O.build()
    w = refdata.hsapiens(120)
    cells = float(sum(len(s) - w.k + 1 for s in w.seqs) * sum(len(s) for s in w.seqs))
    runs = []
    for r in REFERENCE_S:
        t0 = time.perf_counter()
        O.all_pairs(w.seqs, r["distance"], w.k, w.e, threads=r["threads"])
        dt = time.perf_counter() - t0
        runs.append({**r, "oracle_s": round(dt, 3), "ratio_oracle_over_motexomp": round(dt / r["motexomp_s"], 4),
                     "oracle_gcups": round(cells / dt / 1e9, 4),
                     "motexomp_gcups": round(cells / r["motexomp_s"] / 1e9, 4)})
        print(json.dumps(runs[-1]), flush=True)
    out = {"input": "first 120 sequences of the reference's H. sapiens dataset (120 x 1000 nt), k=12 e=2",
           "cells": cells, "host": cpu_model(), "host_threads": os.cpu_count(), "runs": runs}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print("wrote", os.path.relpath(OUT, ROOT))
    return 0


if __name__ == "__main__":
    sys.exit(main())
