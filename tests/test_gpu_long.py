"""Long-sequence mode (-L): one 4-Mnt sequence, Hamming k=10 e=2, GC 43%, 100 planted 10-mers.

The reference gives a single sequence one thread (motex.cc:509-511 parallelises over x only)
and its -L wavefront is compiled out (functions.cc:3010-3019); here the one self pair is cut
into row tiles x text parts over the persistent grid.  -L is a scheduling hint: results
never depend on it.

- a 50-knt prefix of the same generator (50 planted copies) through long_mode=True, bit-exact
  with the oracle;
- the full 4 Mnt on the GPU, checked through what the mathematics fixes at any size: with a
  single sequence every factor occurs in its own text, so u = 1 at every full window and 0
  before; v at sampled positions (the planted copies and random ones) equals a brute-force
  count of the text windows within Hamming distance e.
"""
import numpy as np
import pytest

from motex_amd import MotexParams, synth

pytestmark = pytest.mark.gpu


def test_long4m_prefix_long_mode(ctx, oracle):
    w = synth.dna4m_long(scale=0.0125)       # 50,000 nt, 50 planted copies
    assert w.long_mode and len(w.seqs) == 1 and len(w.seqs[0]) == 50_000
    p = MotexParams(k=w.k, e=w.e, long_mode=True)
    print(f"[long] prefix n={len(w.seqs[0])} k={w.k} e={w.e} -L", flush=True)
    u, v, _ = ctx.extract_all_pairs(w.seqs, p)
    ou, ov, _, _ = oracle.all_pairs(w.seqs, 0, w.k, w.e, threads=1)
    bad = np.nonzero((u != ou) | (v != ov))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}"
    # the hint changes nothing
    u2, v2, _ = ctx.extract_all_pairs(w.seqs, MotexParams(k=w.k, e=w.e))
    assert np.array_equal(u, u2) and np.array_equal(v, v2)


def brute_v(y: np.ndarray, p: int, k: int, e: int) -> int:
    """Windows of y within Hamming distance e of the factor ending at p."""
    win = np.lib.stride_tricks.sliding_window_view(y, k)
    f = y[p - k + 1:p + 1]
    d = np.zeros(len(win), np.int32)
    for c in range(k):
        d += win[:, c] != f[c]
    return int(np.count_nonzero(d <= e))


def test_long4m_full_size_properties(ctx):
    w = synth.dna4m_long()
    y = np.frombuffer(w.seqs[0], np.uint8)
    n, k, e = len(y), w.k, w.e
    print(f"[long] full n={n} k={k} e={e} -L", flush=True)
    b = ctx.batch(w.seqs, MotexParams(k=k, e=e, long_mode=True))
    st = b.run()
    u, v, _ = b.download()
    b.close()
    assert st["cells"] == float((n - k + 1) * n)
    assert np.all(u[:k - 1] == 0) and np.all(v[:k - 1] == 0)
    assert np.all(u[k - 1:] == 1)
    assert np.all(v[k - 1:] >= 1)
    # sampled positions: the planted copies' ends and random ones
    mo = np.frombuffer(w.planted[0], np.uint8)
    win = np.lib.stride_tricks.sliding_window_view(y, k)
    near = np.nonzero((win != mo).sum(axis=1) <= e)[0][:12] + k - 1
    rng = np.random.default_rng(4)
    sample = np.concatenate([near, rng.integers(k - 1, n, 12)])
    for p in sample:
        assert int(v[p]) == brute_v(y, int(p), k, e), f"position {p}"
