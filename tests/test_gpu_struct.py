"""GPU parity of structured-motif extraction against the oracle (bit-exact u and v)."""
import numpy as np
import pytest

from motex_amd import MotexError, MotexParams, synth

pytestmark = pytest.mark.gpu


def rand_seqs(rng, n, lo, hi, alpha):
    a = np.frombuffer(alpha, np.uint8)
    return [a[rng.integers(0, len(a), rng.integers(lo, hi + 1))].tobytes() for _ in range(n)]


@pytest.fixture(params=["range", "range-ordered", "symmetric", "ordered"])
def sym(request, monkeypatch):
    """Runs a test with each structured plan: range-batched (Hamming boxes: every row of x
    against packed text ranges, struct_range_kernels.hip; with one gap each unordered pair
    once, credited both ways) and range-batched over ordered pairs, and one pair per work item
    with the pair-symmetric plan (each unordered pair's box DPs serve both orders) and without."""
    monkeypatch.setenv("MOTEX_STRUCT_RANGE", "1" if request.param.startswith("range") else "0")
    monkeypatch.setenv("MOTEX_SYM", "0" if request.param in ("ordered", "range-ordered") else "1")
    return request.param


def planes(length, err):
    low = err.bit_length()
    return ((1 << low) - 1 - err + length).bit_length()


def range_eligible(seqs, k, e, boxes, distance):
    """Whether the range-batched kernel can take the batch: Hamming boxes, at most 4 gaps,
    counters of at most 7 planes, a whole text and the longest x in one 8192-diagonal window."""
    maxlen = max(len(s) for s in seqs)
    shapes = [(k, e)] + [(b[2], b[3]) for b in boxes]
    return (distance == 0 and len(boxes) <= 4 and 2 * maxlen + 32 <= 8192
            and all(planes(l, x) <= 7 for l, x in shapes))


def check(ctx, oracle, seqs, k, e, boxes, distance=0, sym=None):
    print(f"[struct parity] n={len(seqs)} total={sum(map(len, seqs))} k={k} e={e} d={distance} "
          f"boxes={boxes} plan={sym}", flush=True)
    p = MotexParams(k=k, e=e, distance=distance, boxes=boxes)
    b = ctx.batch(seqs, p)
    st = b.run()
    u, v, off = b.download()
    b.close()
    if sym == "symmetric":
        assert st["symmetric"] == int(distance == 0)
    elif sym == "ordered":
        assert st["symmetric"] == 0
    elif sym == "range" and range_eligible(seqs, k, e, boxes, distance):
        assert st["kernel_variant"] == 1 and st["symmetric"] == 1
    elif sym == "range-ordered" and range_eligible(seqs, k, e, boxes, distance):
        assert st["kernel_variant"] == 1 and st["symmetric"] == 0
    ou, ov, ooff, _ = oracle.all_pairs(seqs, distance, k, e, boxes=boxes)
    bad = np.nonzero((u != ou) | (v != ov))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: gpu {u[bad[:5]]},{v[bad[:5]]} oracle {ou[bad[:5]]},{ov[bad[:5]]}"
    return u, v, st


def test_dna1000_structured_subset(ctx, oracle, sym):
    w = synth.dna1000_structured(scale=0.02)  # 20 sequences of 600 nt, one planted structured motif
    u, v, _ = check(ctx, oracle, w.seqs, w.k, w.e, list(w.boxes), sym=sym)
    assert v.sum() > 0


def test_reference_boxes_example(ctx, oracle, sym):
    # the two-gap structure of data/boxes.txt: gaps [16,17] and [15,16], boxes 6/1 and 7/2
    rng = np.random.default_rng(3)
    seqs = rand_seqs(rng, 8, 40, 300, b"AC")
    check(ctx, oracle, seqs, 6, 1, [(16, 17, 6, 1), (15, 16, 7, 2)], sym=sym)


@pytest.mark.parametrize("seed,k,e,boxes,alpha", [
    (1, 4, 1, [(0, 6, 3, 0)], b"ACG"),                       # zero-length gaps allowed
    (2, 5, 1, [(2, 4, 5, 1), (1, 3, 4, 1), (0, 2, 3, 0)], b"AC"),  # three gaps
    (3, 6, 2, [(5, 40, 6, 2)], b"ACGT"),                     # wide gap range
    (4, 3, 1, [(1, 2, 2, 500)], b"ACGT"),                    # box error >= m: passes everywhere
    (8, 4, 1, [(1, 3, 4, 60)], b"ACGT"),                     # passes everywhere in some pairs only
    (9, 3, 0, [(0, 1, 3, 0)], b"A"),                         # one symbol: lists overflow, bitmap scan
    (10, 5, 1, [(4, 15, 6, 1)], b"ACG"),                     # x + y over 2048: lists in several runs
])
def test_hamming_shapes(ctx, oracle, sym, seed, k, e, boxes, alpha):
    rng = np.random.default_rng(seed)
    lo, hi = {8: (40, 120), 9: (60, 400), 10: (900, 2600)}.get(seed, (10, 200))
    check(ctx, oracle, rand_seqs(rng, 7, lo, hi, alpha), k, e, boxes, sym=sym)


@pytest.mark.parametrize("seed,k,e,boxes", [
    (5, 5, 1, [(2, 6, 4, 1)]),
    (6, 6, 1, [(3, 5, 6, 1), (0, 2, 3, 0)]),
    (7, 16, 3, [(1, 3, 15, 2)]),                             # 64-bit words for both boxes
])
def test_edit_shapes(ctx, oracle, seed, k, e, boxes):
    rng = np.random.default_rng(seed)
    check(ctx, oracle, rand_seqs(rng, 6, 20, 160, b"ACG"), k, e, boxes, distance=1)


def test_plain_form_equals_batched(ctx):
    w = synth.dna1000_structured(scale=0.01)
    p = MotexParams(k=w.k, e=w.e, boxes=list(w.boxes))
    b = ctx.batch(w.seqs, p)
    b.run()
    ub, vb, _ = b.download()
    b.clear()
    b.run_plain()
    up, vp, _ = b.download()
    assert np.array_equal(ub, up) and np.array_equal(vb, vp)
    b.close()


def test_edit_boxes_with_128_bit_words(ctx, oracle):
    rng = np.random.default_rng(41)
    a = np.frombuffer(b"AC", np.uint8)
    seqs = [a[rng.integers(0, 2, int(rng.integers(150, 260)))].tobytes() for _ in range(5)]
    p = MotexParams(k=36, e=8, distance=1, boxes=[(3, 5, 34, 7)])
    print(f"[struct] edit boxes 36/8 + [3,5] + 34/7 n={len(seqs)}", flush=True)
    u, v, _ = ctx.extract_all_pairs(seqs, p)
    ou, ov, _, _ = oracle.all_pairs(seqs, 1, 36, 8, boxes=[(3, 5, 34, 7)])
    assert np.array_equal(u, ou) and np.array_equal(v, ov)
    assert v.sum() > 0


def check_struct(ctx, oracle, seqs, d, k, e, boxes, tag):
    print(f"[struct] {tag} n={len(seqs)} lens={[len(s) for s in seqs]} d={d} k={k} e={e} boxes={boxes}", flush=True)
    p = MotexParams(k=k, e=e, distance=d, boxes=boxes)
    u, v, _ = ctx.extract_all_pairs(seqs, p)
    ou, ov, _, _ = oracle.all_pairs(seqs, d, k, e, boxes=boxes)
    bad = np.nonzero((u != ou) | (v != ov))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}"
    return u, v


@pytest.mark.parametrize("d", [0, 1])
def test_sequences_beyond_4096(ctx, oracle, sym, d):
    # 5-10 knt sequences (the reference's structured operator has no length limit)
    rng = np.random.default_rng(50 + d)
    seqs = rand_seqs(rng, 3, 5000, 10000 if d == 0 else 6000, b"ACGT")
    u, v = check_struct(ctx, oracle, seqs, d, 8, 1, [(10, 20, 8, 1)], "long")
    assert v.sum() > 0


def test_five_gaps(ctx, oracle, sym):
    rng = np.random.default_rng(55)
    seqs = rand_seqs(rng, 6, 60, 200, b"AC")
    u, v = check_struct(ctx, oracle, seqs, 0, 4, 1, [(2, 3, 4, 1), (1, 2, 3, 0), (2, 2, 4, 1), (0, 1, 3, 1),
                                                    (1, 3, 4, 2)], "five gaps")
    assert v.sum() > 0


def test_eight_gaps_edit(ctx, oracle):
    rng = np.random.default_rng(56)
    seqs = rand_seqs(rng, 4, 60, 120, b"AC")
    check_struct(ctx, oracle, seqs, 1, 3, 1, [(1, 2, 3, 1)] * 8, "eight gaps, edit")


def test_boxes_longer_than_64(ctx, oracle, sym):
    rng = np.random.default_rng(57)
    seqs = rand_seqs(rng, 4, 300, 600, b"AC")
    u, v = check_struct(ctx, oracle, seqs, 0, 100, 30, [(5, 9, 80, 24)], "long boxes")
    assert v.sum() > 0


def test_range_plan_is_the_default_for_hamming_boxes(ctx, monkeypatch):
    monkeypatch.delenv("MOTEX_STRUCT_RANGE", raising=False)
    w = synth.dna1000_structured(scale=0.01)
    b = ctx.batch(w.seqs, MotexParams(k=w.k, e=w.e, boxes=list(w.boxes)))
    st = b.run()
    b.close()
    assert st["kernel_variant"] == 1 and st["handed_pairs"] == 0


@pytest.mark.parametrize("qcap", [64, 2304])
def test_range_overflow_hands_pairs_to_the_pair_kernel(ctx, oracle, monkeypatch, qcap):
    # hit lists too small for a range's hits (about 2400 per 12-text range here):
    # those ranges' pairs go to the per-pair kernel through the overflow list, in the same run
    monkeypatch.setenv("MOTEX_STRUCT_RANGE", "1")
    monkeypatch.setenv("MOTEX_SR_HCAP", str(qcap))
    w = synth.dna1000_structured(scale=0.02)
    u, v, st = check(ctx, oracle, w.seqs, w.k, w.e, list(w.boxes), sym="range")
    print(f"[struct] qcap {qcap}: {st['handed_pairs']} pairs handed over", flush=True)
    n = len(w.seqs)
    if qcap == 64:  # every unit: each unordered pair once (pair-symmetric units, one gap)
        assert st["handed_pairs"] == n * (n + 1) // 2
    assert st["handed_pairs"] <= n * (n + 1) // 2
    assert v.sum() > 0


def test_range_overflow_list_full_is_an_error(ctx, monkeypatch):
    monkeypatch.setenv("MOTEX_STRUCT_RANGE", "1")
    monkeypatch.setenv("MOTEX_SR_HCAP", "64")
    monkeypatch.setenv("MOTEX_SR_OVF_CAP", "3")
    w = synth.dna1000_structured(scale=0.02)
    b = ctx.batch(w.seqs, MotexParams(k=w.k, e=w.e, boxes=list(w.boxes)))
    with pytest.raises(MotexError, match="overflowed"):
        b.run()
    b.close()


def test_range_plan_with_short_texts_between(ctx, oracle, monkeypatch):
    # texts shorter than the whole structure end a range; x shorter than it has no rows
    monkeypatch.setenv("MOTEX_STRUCT_RANGE", "1")
    rng = np.random.default_rng(61)
    seqs = rand_seqs(rng, 30, 5, 300, b"ACG")
    u, v, st = check(ctx, oracle, seqs, 5, 1, [(3, 9, 5, 1)], sym="range")
    assert st["kernel_variant"] == 1 and v.sum() > 0


@pytest.mark.parametrize("symmetric", ["1", "0"])
def test_range_overflow_ordered_and_symmetric(ctx, oracle, monkeypatch, symmetric):
    # some units overflow (hit slots of 2304 against about 2400 hits), both unit plans
    monkeypatch.setenv("MOTEX_STRUCT_RANGE", "1")
    monkeypatch.setenv("MOTEX_SYM", symmetric)
    monkeypatch.setenv("MOTEX_SR_HCAP", "2304")
    w = synth.dna1000_structured(scale=0.03)
    u, v, st = check(ctx, oracle, w.seqs, w.k, w.e, list(w.boxes), sym="range" if symmetric == "1" else "range-ordered")
    print(f"[struct] symmetric {symmetric}: {st['handed_pairs']} pairs handed over", flush=True)
    assert v.sum() > 0


def test_range_plan_dense_rows(ctx, oracle, monkeypatch):
    # 6-nt boxes with one error in DNA: about 38 hits per row of a range, beyond the
    # one-gap merge's register slots, so those rows take the chain walk
    monkeypatch.setenv("MOTEX_STRUCT_RANGE", "1")
    rng = np.random.default_rng(62)
    seqs = rand_seqs(rng, 40, 100, 250, b"ACGT")
    u, v, st = check(ctx, oracle, seqs, 6, 1, [(3, 8, 6, 1)], sym="range")
    assert st["kernel_variant"] == 1 and st["handed_pairs"] == 0 and v.sum() > 0


@pytest.mark.parametrize("symmetric", ["1", "0"])
def test_range_plan_in_several_launches(ctx, oracle, monkeypatch, symmetric):
    # slot memory for a handful of units (about 46 KB each here) per launch pair: the run is
    # several sweep/merge launch pairs sharing the accumulators, the self-pair flags and the
    # overflow list
    monkeypatch.setenv("MOTEX_STRUCT_RANGE", "1")
    monkeypatch.setenv("MOTEX_SYM", symmetric)
    monkeypatch.setenv("MOTEX_SR_SLOT_BYTES", "300000")
    rng = np.random.default_rng(63)
    seqs = rand_seqs(rng, 24, 150, 400, b"ACGT")
    u, v, st = check(ctx, oracle, seqs, 7, 1, [(2, 9, 7, 1)], sym="range" if symmetric == "1" else "range-ordered")
    assert v.sum() > 0


def test_range_plan_protein_alphabet(ctx, oracle, monkeypatch):
    # 20 letters: a larger NE window per unit (20 symbol rows in LDS)
    monkeypatch.setenv("MOTEX_STRUCT_RANGE", "1")
    w = synth.prot500_edit(scale=0.03)
    seqs = [s[:700] for s in w.seqs]
    u, v, st = check(ctx, oracle, seqs, 5, 1, [(3, 12, 5, 1)], sym="range")
    assert st["kernel_variant"] == 1 and v.sum() > 0


def test_range_plan_on_genes(ctx, oracle, monkeypatch):
    # the reference's H. sapiens genes (N symbols, repeats), one gap, on the default plan
    monkeypatch.delenv("MOTEX_STRUCT_RANGE", raising=False)
    from motex_amd import refdata
    seqs = refdata.hsapiens(60).seqs
    u, v, st = check(ctx, oracle, seqs, 8, 1, [(10, 30, 8, 1)])
    assert st["kernel_variant"] == 1 and v.sum() > 0


@pytest.mark.parametrize("ring", ["1", "0"], ids=["ring", "generic"])
@pytest.mark.parametrize("k,e,box", [
    (5, 0, (5, 0)),    # threshold plane 0
    (6, 1, (7, 2)),    # two shapes: the first takes the ring, the second the generic sweep
    (7, 2, (7, 2)),    # odd length: trips of 14 rows
    (8, 1, (8, 1)),
    (8, 4, (5, 3)),    # threshold plane 3 for the first shape
    (8, 7, (8, 7)),    # nearly every window passes: slots overflow to the per-pair kernel
    (9, 1, (9, 1)),    # odd length: trips of 18 rows
    (10, 3, (10, 3)),
    (11, 2, (11, 2)),  # beyond the ring lengths: the generic sweep either way
])
def test_range_ring_lengths(ctx, oracle, monkeypatch, ring, k, e, box):
    # box lengths 5-10 sweep with the leaving column in a register ring and rows with hits
    # queued in LDS (MOTEX_SR_RING=0: the generic sweep); some x are shorter than a box, some
    # shorter than a trip
    monkeypatch.setenv("MOTEX_STRUCT_RANGE", "1")
    monkeypatch.setenv("MOTEX_SR_RING", ring)
    rng = np.random.default_rng(k * 10 + e)
    seqs = rand_seqs(rng, 20, 3, 400, b"ACGT")
    check(ctx, oracle, seqs, k, e, [(2, 9, box[0], box[1])], sym="range")
