"""Runtime features around the batched kernels: HIP-graph replay of repeated runs and the
tile-shape autotuner.  Both change only how the work is launched, never the results, so
each case is checked against the oracle and against the eager run."""
import numpy as np
import pytest

import motex_amd
from motex_amd import MotexParams, synth

pytestmark = pytest.mark.gpu


def workload(kind):
    if kind == "hd":
        w = synth.dna50_hamming(scale=0.4)
    elif kind == "ed":
        w = synth.prot500_edit(scale=0.02)
    else:
        w = synth.dna1000_structured(scale=0.02)
    return w, MotexParams(w.k, w.e, w.distance, w.alphabet, boxes=w.boxes)


@pytest.mark.parametrize("kind", ["hd", "ed", "struct"])
def test_graph_replays_match_eager_and_oracle(ctx, oracle, kind, monkeypatch):
    w, p = workload(kind)
    print(f"[graph] {kind} seqs={len(w.seqs)}", flush=True)
    ou, ov, off, _ = oracle.all_pairs(w.seqs, w.distance, w.k, w.e, boxes=w.boxes)
    # eager reference run
    monkeypatch.setenv("MOTEX_GRAPH", "0")
    b = ctx.batch(w.seqs, p)
    st = b.run()
    assert st["graph_replay"] == 0
    u0, v0, _ = b.download()
    b.close()
    assert np.array_equal(u0, ou) and np.array_equal(v0, ov)
    monkeypatch.delenv("MOTEX_GRAPH")
    b = ctx.batch(w.seqs, p)
    for rep in range(5):
        if rep == 3:
            # eager work on the same stream and device arrays between replays
            b.clear()
            b.run_plain(0, 1)
            b.clear()
        st = b.run()
        assert st["graph_replay"] == 1
        u, v, _ = b.download()
        assert np.array_equal(u, u0), f"replay {rep}: u differs"
        assert np.array_equal(v, v0), f"replay {rep}: v differs"
    b.close()


def test_autotune_keeps_results_and_is_used(ctx, oracle):
    w, p = workload("hd")
    sample = w.seqs[:12]
    t = ctx.autotune(sample, p)
    print(f"[autotune] {t}", flush=True)
    assert t["candidates"] >= 2 and t["kernel_ms"] > 0
    assert t["blocks"] in (1, 2, 4) and t["rows"] >= 64 and t["waves_per_cu"] >= 8
    b = ctx.batch(w.seqs, p)
    st = b.run()
    assert (st["tile_blocks"], st["tile_rows"], st["waves_per_cu"]) == (t["blocks"], t["rows"], t["waves_per_cu"])
    u, v, off = b.download()
    ou, ov, _, _ = oracle.all_pairs(w.seqs, w.distance, w.k, w.e)
    assert np.array_equal(u, ou) and np.array_equal(v, ov)
    b.close()


def test_autotune_edit_distance(ctx, oracle):
    w, p = workload("ed")
    t = ctx.autotune(w.seqs[:6], p)
    assert t["candidates"] >= 2
    u, v, off = ctx.extract_all_pairs(w.seqs, p)
    ou, ov, _, _ = oracle.all_pairs(w.seqs, w.distance, w.k, w.e)
    assert np.array_equal(u, ou) and np.array_equal(v, ov)


def test_autotune_rejects_empty_sample(ctx):
    with pytest.raises(motex_amd.MotexError):
        ctx.autotune([], MotexParams(8, 1))
