"""Serving: every job of a concurrent mix gives the oracle's u/v, and the same u/v as when
it runs alone."""
import numpy as np
import pytest

import motex_amd
from motex_amd import MotexParams, serving, synth

pytestmark = pytest.mark.gpu

# Jobs of the 256-job serving stream at full size (20-40 sequences of 100-2000 symbols) whose
# oracle run takes about a second: Hamming DNA e=1..3, Hamming protein, edit protein e=3,
# edit DNA e=0 and e=2, edit protein e=2.
FULL_SIZE_JOBS = [16, 141, 235, 232, 13, 187, 80, 154]


def mixed_jobs():
    jobs = [synth.serving_job(i, scale=0.08) for i in range(10)]
    jobs += [synth.serving_job(i) for i in FULL_SIZE_JOBS]
    w4 = synth.dna1000_structured(scale=0.01)
    jobs.append(synth.Workload("structured", "DNA", w4.seqs, 0, w4.k, w4.e, 40, boxes=list(w4.boxes)))
    w3 = synth.prot500_edit(scale=0.01)
    jobs.append(synth.Workload("edit", "PROT", w3.seqs, 1, w3.k, w3.e, 30))
    return jobs


def test_results_equal_the_oracle_alone_and_in_the_mix(ctx, oracle):
    jobs = mixed_jobs()
    want = []
    for w in jobs:
        print(f"[serving] oracle {w.name} n={len(w.seqs)} k={w.k} e={w.e} d={w.distance} "
              f"boxes={w.boxes}", flush=True)
        ou, ov, _, _ = oracle.all_pairs(w.seqs, w.distance, w.k, w.e, boxes=list(w.boxes))
        want.append((ou, ov))
    for w, (ou, ov) in zip(jobs, want):
        print(f"[serving] alone {w.name}", flush=True)
        u, v, _ = ctx.extract_all_pairs(w.seqs, serving.job_params(w))
        assert np.array_equal(u, ou) and np.array_equal(v, ov), f"alone: {w.name}"
    # submit everything at once: jobs are admitted as slots free up and overlap on the GPU
    print("[serving] mix", flush=True)
    handles = [serving.JobHandle(ctx, w.seqs, serving.job_params(w)) for w in jobs]
    for w, h, (ou, ov) in zip(jobs, handles, want):
        u, v, off = h.wait()
        assert np.array_equal(u, ou) and np.array_equal(v, ov), f"in the mix: {w.name}"
        lat, dev, cells = h.times()
        assert lat > 0 and dev > 0 and cells > 0
        h.release()


def test_poll_and_release(ctx):
    w = synth.serving_job(3, scale=0.05)
    h = serving.JobHandle(ctx, w.seqs, serving.job_params(w))
    u, v, _ = h.wait()
    assert h.done()
    h.release()
    with pytest.raises(motex_amd.MotexError):
        h.release()          # unknown id now


def test_invalid_job_is_refused(ctx):
    w = synth.serving_job(1, scale=0.05)
    with pytest.raises(motex_amd.MotexError):
        serving.JobHandle(ctx, w.seqs, MotexParams(k=5, e=5))


def test_admission_refuses_jobs_over_the_budget():
    small = motex_amd.Context(0, hbm_budget=1 << 16)
    try:
        w = synth.dna1200_hamming(scale=0.05)
        with pytest.raises(motex_amd.MotexError) as ei:
            serving.JobHandle(small, w.seqs, serving.job_params(w))
        assert ei.value.code == -2
    finally:
        small.close()


def test_stream_replay_reports_latency(ctx):
    r = serving.run_stream(ctx, njobs=12, warm=4, scale=0.05)
    assert r["p99_latency_ms"] >= r["p50_latency_ms"] > 0
    assert r["aggregate_gcups"] > 0
