"""Serving client: submit extraction jobs to the continuous-batching loop of the C ABI
(motex_submit / motex_poll / motex_wait) and replay the serving workload -- a stream of
jobs with Poisson arrivals at a fraction of measured capacity -- reporting p50/p99 job
latency and aggregate GCUPS.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import Job, MotexParams, _check, _row_ptrs, _seq_args, lib, u32p, uv_layout
from . import synth


class JobHandle:
    """A submitted job; keeps its inputs and output arrays alive until released."""

    def __init__(self, ctx, seqs: list[bytes], params: MotexParams):
        self.ctx, self.seqs, self.params = ctx, seqs, params
        self.off = uv_layout(seqs, params.nb_structs)
        n = max(int(self.off[-1]), 1)
        self.u = np.zeros(n, np.uint32)
        self.v = np.zeros(n, np.uint32)
        self._sa, self._lens = _seq_args(seqs)
        self._p, self._keep = params.to_c()
        self._up = _row_ptrs(self.u, self.off)
        self._vp = _row_ptrs(self.v, self.off)
        self._job = Job(self._sa, self._lens.ctypes.data_as(u32p), len(seqs), self._p,
                        C.cast(self._up, C.POINTER(u32p)), C.cast(self._vp, C.POINTER(u32p)))
        jid = C.c_uint64()
        _check(lib().motex_submit(ctx.h, C.byref(self._job), C.byref(jid)))
        self.id = jid.value

    def done(self) -> bool:
        d = C.c_int()
        _check(lib().motex_poll(self.ctx.h, self.id, C.byref(d)))
        return bool(d.value)

    def wait(self):
        _check(lib().motex_wait(self.ctx.h, self.id))
        n = int(self.off[-1])
        return self.u[:n], self.v[:n], self.off

    def times(self):
        lat, dev, cells = C.c_double(), C.c_double(), C.c_double()
        _check(lib().motex_job_times(self.ctx.h, self.id, C.byref(lat), C.byref(dev), C.byref(cells)))
        return lat.value, dev.value, cells.value

    def release(self):
        _check(lib().motex_job_release(self.ctx.h, self.id))


def job_params(w: synth.Workload) -> MotexParams:
    return MotexParams(k=w.k, e=w.e, distance=w.distance, alphabet=w.alphabet, boxes=list(w.boxes))


def run_stream(ctx, njobs: int = 256, load: float = 0.7, warm: int = 16, seed: int = 5,
               scale: float = 1.0) -> dict:
    """Replays njobs serving jobs with Poisson arrivals at `load` x measured capacity."""
    jobs = [synth.serving_job(i, scale) for i in range(njobs)]
    # capacity from a serial warm-up on the head of the same mix
    warm = min(warm, njobs)
    t0 = time.perf_counter()
    for w in jobs[:warm]:
        h = JobHandle(ctx, w.seqs, job_params(w))
        h.wait()
        h.release()
    cap = warm / (time.perf_counter() - t0)          # jobs per second, one at a time
    rate = load * cap
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, njobs))
    handles = []
    start = time.perf_counter()
    for w, at in zip(jobs, arrivals):
        delay = at - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        handles.append(JobHandle(ctx, w.seqs, job_params(w)))
    for h in handles:
        h.wait()
    makespan = time.perf_counter() - start
    lat, cells, dev = [], 0.0, 0.0
    for h in handles:
        l, d, c = h.times()
        lat.append(l)
        cells += c
        dev += d
        h.release()
    lat = np.array(lat)
    return {"jobs": njobs, "load": load, "capacity_jobs_per_s": round(cap, 3),
            "arrival_rate_jobs_per_s": round(rate, 3),
            "p50_latency_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_latency_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_latency_ms": round(float(lat.mean()), 3),
            "aggregate_gcups": round(cells / makespan / 1e9, 3),
            "makespan_s": round(makespan, 3), "cells": cells,
            "device_ms_sum": round(dev, 3)}
