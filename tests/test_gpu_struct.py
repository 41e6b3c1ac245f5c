"""GPU parity of structured-motif extraction against the oracle (bit-exact u and v)."""
import numpy as np
import pytest

from motex_amd import MotexParams, synth

pytestmark = pytest.mark.gpu


def rand_seqs(rng, n, lo, hi, alpha):
    a = np.frombuffer(alpha, np.uint8)
    return [a[rng.integers(0, len(a), rng.integers(lo, hi + 1))].tobytes() for _ in range(n)]


def check(ctx, oracle, seqs, k, e, boxes, distance=0):
    print(f"[struct parity] n={len(seqs)} total={sum(map(len, seqs))} k={k} e={e} d={distance} "
          f"boxes={boxes}", flush=True)
    p = MotexParams(k=k, e=e, distance=distance, boxes=boxes)
    u, v, off = ctx.extract_all_pairs(seqs, p)
    ou, ov, ooff, _ = oracle.all_pairs(seqs, distance, k, e, boxes=boxes)
    bad = np.nonzero((u != ou) | (v != ov))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: gpu {u[bad[:5]]},{v[bad[:5]]} oracle {ou[bad[:5]]},{ov[bad[:5]]}"
    return u, v


def test_config4a_subset(ctx, oracle):
    w = synth.config4a(scale=0.02)  # 20 sequences of 600 nt, one planted structured motif
    u, v = check(ctx, oracle, w.seqs, w.k, w.e, list(w.boxes))
    assert v.sum() > 0


def test_reference_boxes_example(ctx, oracle):
    # the two-gap structure of data/boxes.txt: gaps [16,17] and [15,16], boxes 6/1 and 7/2
    rng = np.random.default_rng(3)
    seqs = rand_seqs(rng, 8, 40, 300, b"AC")
    check(ctx, oracle, seqs, 6, 1, [(16, 17, 6, 1), (15, 16, 7, 2)])


@pytest.mark.parametrize("seed,k,e,boxes,alpha", [
    (1, 4, 1, [(0, 6, 3, 0)], b"ACG"),                       # zero-length gaps allowed
    (2, 5, 1, [(2, 4, 5, 1), (1, 3, 4, 1), (0, 2, 3, 0)], b"AC"),  # three gaps
    (3, 6, 2, [(5, 40, 6, 2)], b"ACGT"),                     # wide gap range
    (4, 3, 1, [(1, 2, 2, 500)], b"ACGT"),                    # box error >= m: passes everywhere
])
def test_hamming_shapes(ctx, oracle, seed, k, e, boxes, alpha):
    rng = np.random.default_rng(seed)
    check(ctx, oracle, rand_seqs(rng, 7, 10, 200, alpha), k, e, boxes)


@pytest.mark.parametrize("seed,k,e,boxes", [
    (5, 5, 1, [(2, 6, 4, 1)]),
    (6, 6, 1, [(3, 5, 6, 1), (0, 2, 3, 0)]),
    (7, 16, 3, [(1, 3, 15, 2)]),                             # 64-bit words for both boxes
])
def test_edit_shapes(ctx, oracle, seed, k, e, boxes):
    rng = np.random.default_rng(seed)
    check(ctx, oracle, rand_seqs(rng, 6, 20, 160, b"ACG"), k, e, boxes, distance=1)


def test_plain_form_equals_batched(ctx):
    w = synth.config4a(scale=0.01)
    p = MotexParams(k=w.k, e=w.e, boxes=list(w.boxes))
    b = ctx.batch(w.seqs, p)
    b.run()
    ub, vb, _ = b.download()
    b.clear()
    b.run_plain()
    up, vp, _ = b.download()
    assert np.array_equal(ub, up) and np.array_equal(vb, vp)
    b.close()
