"""GPU parity of the edit-distance (MaxShift) path against the oracle (bit-exact u and v)."""
import numpy as np
import pytest

from motex_amd import MotexParams, synth

pytestmark = pytest.mark.gpu


def rand_seqs(rng, n, lo, hi, alpha):
    a = np.frombuffer(alpha, np.uint8)
    return [a[rng.integers(0, len(a), rng.integers(lo, hi + 1))].tobytes() for _ in range(n)]


@pytest.fixture(params=["1", "pipe", "0"], ids=["text_parallel", "text_parallel_pipelined", "one_text_per_lane"])
def kernel(request, monkeypatch):
    """Runs a test on the text-parallel edit kernel (32 texts per lane word; used for k in
    4..30) in its one-wave-per-item and pipelined (several waves per item) forms, and again
    with it switched off."""
    monkeypatch.setenv("MOTEX_ED_SLICE", "0" if request.param == "0" else "1")
    monkeypatch.setenv("MOTEX_ED_PIPE", "1" if request.param == "pipe" else "0")
    return request.param


def check(ctx, oracle, seqs, k, e, alphabet="DNA", kernel=None):
    print(f"[ed parity] n={len(seqs)} total={sum(map(len, seqs))} k={k} e={e} kernel={kernel}", flush=True)
    p = MotexParams(k=k, e=e, distance=1, alphabet=alphabet)
    b = ctx.batch(seqs, p)
    st = b.run()
    u, v, off = b.download()
    b.close()
    if kernel is not None:
        sliced = kernel != "0" and 4 <= k <= 30
        assert (st["kernel_variant"] in (2, 3)) == sliced
        if sliced:
            assert (st["kernel_variant"] == 3) == (kernel == "pipe")
    ou, ov, ooff, _ = oracle.all_pairs(seqs, 1, k, e)
    bad = np.nonzero((u != ou) | (v != ov))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: gpu {u[bad[:5]]},{v[bad[:5]]} oracle {ou[bad[:5]]},{ov[bad[:5]]}"
    return u, v


def test_prot500_subset(ctx, oracle, kernel):
    w = synth.prot500_edit(scale=0.03)  # 15 protein sequences, lengths 200..2000
    u, v = check(ctx, oracle, w.seqs, w.k, w.e, "PROT", kernel)
    assert v.sum() > 0


@pytest.mark.parametrize("pipe", ["0", "1"], ids=["one_wave", "pipelined"])
def test_text_groups_of_mixed_lengths(ctx, oracle, monkeypatch, pipe):
    # 70 texts: three groups of 32 texts sorted by length, ragged ends, short texts skipped
    # (on the text-parallel kernel, which the default choice keeps for larger inputs)
    monkeypatch.setenv("MOTEX_ED_SLICE", "1")
    monkeypatch.setenv("MOTEX_ED_PIPE", pipe)
    rng = np.random.default_rng(77)
    seqs = rand_seqs(rng, 70, 2, 900, b"ACGT")
    check(ctx, oracle, seqs, 7, 2, kernel="pipe" if pipe == "1" else "1")


def test_pipelined_strips_of_long_x(ctx, oracle, monkeypatch):
    # x up to 1,400 symbols: up to 22 strips, more than the 4 waves of an item, so waves run
    # several strips and the boundary buffers are reused
    monkeypatch.setenv("MOTEX_ED_SLICE", "1")
    monkeypatch.setenv("MOTEX_ED_PIPE", "1")
    rng = np.random.default_rng(78)
    seqs = rand_seqs(rng, 9, 60, 1400, b"ACGT")
    check(ctx, oracle, seqs, 9, 2, kernel="pipe")


@pytest.mark.parametrize("seed,k,e,alpha,n,lo,hi", [
    (1, 5, 1, b"ACGT", 10, 3, 400),        # short texts (< k) are skipped
    (2, 8, 2, b"AC", 8, 60, 700),          # dense hits, several strips
    (3, 1, 0, b"ACGT", 6, 1, 90),          # k = 1
    (4, 14, 3, b"ACGT", 6, 100, 600),      # widest 32-bit word case
    (5, 12, 0, b"AC", 6, 100, 300),        # e = 0
    (6, 20, 4, b"ACGT", 5, 100, 400),      # 64-bit words
    (7, 32, 6, b"AC", 4, 100, 300),        # k = 32 (one-text-per-lane kernel only)
    (11, 15, 3, b"ACGT", 40, 10, 500),     # text-parallel kernel past 32-bit words, two groups
    (12, 23, 22, b"A", 5, 20, 200),        # one symbol, e = k - 1: first k with AGPR state
    (13, 30, 7, b"AC", 36, 30, 300),       # largest k of the text-parallel kernel
    (8, 3, 2, b"ACGT", 8, 2, 70),          # e = k-1, texts shorter than one wave
    (9, 10, 9, b"A", 5, 10, 200),          # one symbol: widest words, every tie
    (10, 4, 1, b"ARNDCQEGHILKMFPSTWYV", 40, 4, 300),  # protein, more than one text group
    (14, 33, 6, b"AC", 5, 30, 300),        # 128-bit words from k = 33
    (15, 48, 12, b"ACGT", 5, 40, 400),
    (16, 64, 20, b"AC", 4, 60, 300),       # longest edit factor
    (17, 64, 63, b"A", 3, 60, 200),        # one symbol, e = k - 1: widest 128-bit words
])
def test_random_edges(ctx, oracle, kernel, seed, k, e, alpha, n, lo, hi):
    rng = np.random.default_rng(100 + seed)
    check(ctx, oracle, rand_seqs(rng, n, lo, hi, alpha), k, e, kernel=kernel)


def test_plain_form_equals_batched(ctx):
    rng = np.random.default_rng(5)
    seqs = rand_seqs(rng, 6, 100, 500, b"ACGT")
    p = MotexParams(k=6, e=1, distance=1)
    b = ctx.batch(seqs, p)
    b.run()
    ub, vb, _ = b.download()
    b.clear()
    b.run_plain()
    up, vp, _ = b.download()
    assert np.array_equal(ub, up) and np.array_equal(vb, vp)
    b.close()
