"""GPU parity of the Hamming-distance path against the oracle (bit-exact u and v)."""
import numpy as np
import pytest

import motex_amd
from motex_amd import MotexParams, synth

pytestmark = pytest.mark.gpu


def rand_seqs(rng, n, lo, hi, alpha):
    a = np.frombuffer(alpha, np.uint8)
    return [a[rng.integers(0, len(a), rng.integers(lo, hi + 1))].tobytes() for _ in range(n)]


@pytest.fixture(params=["0", "1"], ids=["ordered", "symmetric"])
def sym(request, monkeypatch):
    """Runs a test with the pair-symmetric Hamming plan forced off, then on (where the
    input allows it: every text fits one work item's range)."""
    monkeypatch.setenv("MOTEX_SYM", request.param)
    return request.param == "1"


def check(ctx, oracle, seqs, k, e, alphabet="DNA", sym=None):
    print(f"[hd parity] n={len(seqs)} total={sum(map(len, seqs))} k={k} e={e} sym={sym}", flush=True)
    p = MotexParams(k=k, e=e, distance=0, alphabet=alphabet)
    b = ctx.batch(seqs, p)
    st = b.run()
    u, v, off = b.download()
    b.close()
    if sym is not None and len(seqs) >= 2:
        assert st["symmetric"] == int(sym)
    ou, ov, ooff, _ = oracle.all_pairs(seqs, 0, k, e)
    assert np.array_equal(off[:-1], ooff)
    bad = np.nonzero((u != ou) | (v != ov))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: gpu {u[bad[:5]]},{v[bad[:5]]} oracle {ou[bad[:5]]},{ov[bad[:5]]}"
    return u, v


def test_dna50_full(ctx, oracle, sym):
    w = synth.dna50_hamming()
    u, v = check(ctx, oracle, w.seqs, w.k, w.e, sym=sym)
    assert v.sum() > 0


def test_dna1200_subset(ctx, oracle, sym):
    w = synth.dna1200_hamming(scale=0.05)  # 60 sequences of 1000 nt
    check(ctx, oracle, w.seqs, w.k, w.e, sym=sym)


@pytest.mark.parametrize("cap", ["8", "1"])
def test_symmetric_text_side_lists_full(ctx, oracle, monkeypatch, cap):
    # every range's list of first text-side occurrences holds only `cap` entries (rounded up
    # to 8): the rest take the direct atomics, with the same result
    monkeypatch.setenv("MOTEX_SYM", "1")
    monkeypatch.setenv("MOTEX_SYM_LIST_CAP", cap)
    rng = np.random.default_rng(11)
    check(ctx, oracle, rand_seqs(rng, 40, 200, 900, b"ACGT"), 8, 2, sym=True)


def test_symmetric_is_the_default_for_rare_occurrences(ctx):
    w = synth.dna1200_hamming(scale=0.05)
    b = ctx.batch(w.seqs, MotexParams(k=w.k, e=w.e))
    assert b.stats()["symmetric"] == 1
    b.close()
    # k=5, e=1 over ACGT: occurrences too dense for the text-side atomics to pay off
    b = ctx.batch(w.seqs, MotexParams(k=5, e=1))
    assert b.stats()["symmetric"] == 0
    b.close()


@pytest.mark.parametrize("seed,k,e,alpha,n,lo,hi", [
    (1, 5, 1, b"ACGT", 20, 3, 300),       # short sequences (< k) mixed in
    (2, 8, 2, b"ACGTN", 15, 50, 400),      # N is an ordinary symbol
    (3, 1, 0, b"ACGT", 8, 1, 100),         # k = 1
    (4, 12, 0, b"AC", 12, 100, 500),       # e = 0 (exact)
    (5, 12, 11, b"ACGT", 6, 12, 80),       # e = k-1
    (6, 31, 7, b"AC", 6, 40, 300),         # five counter planes
    (7, 64, 20, b"AC", 5, 60, 300),        # k = 64 (seven planes)
    (8, 10, 2, b"ARNDCQEGHILKMFPSTWYV", 10, 100, 900),  # protein alphabet
    (9, 6, 1, b"0123456789", 10, 50, 300),  # USR alphabet
    (10, 65, 20, b"AC", 5, 60, 300),       # beyond 64: generic single-block kernel
    (11, 200, 70, b"AC", 4, 150, 600),     # eleven-plane counters, long warm-up rows
    (12, 1024, 400, b"AC", 3, 1000, 1600),  # longest factor
    (13, 16, 2, b"ACGT", 12, 200, 900),     # k = 16: generic kernels (five counter planes)
    (14, 13, 3, b"ACGT", 12, 200, 900),     # odd k from 13: trips of 26 rows
    (15, 15, 1, b"ARNDCQEGHILKMFPSTWYV", 10, 100, 900),  # ring kernels on 20 symbols
    (16, 8, 1, b"ACGT", 30, 100, 700),      # occurrence rate 4e-4: pair-symmetric by default
])
def test_random_edges(ctx, oracle, sym, seed, k, e, alpha, n, lo, hi):
    rng = np.random.default_rng(seed)
    alphabet = "PROT" if len(alpha) == 20 else ("USR" if alpha[0:1] == b"0" else "DNA")
    # k = 1024 items have 1024 rows, so texts over 2048 - 1024 - 32 symbols are split across
    # items, which rules out the pair-symmetric plan even when it is asked for
    check(ctx, oracle, rand_seqs(rng, n, lo, hi, alpha), k, e, alphabet, sym=None if k >= 1024 else sym)


def test_long_sequence_split(ctx, oracle):
    # texts longer than one work item's range are split; u must still count the text once
    rng = np.random.default_rng(42)
    seqs = rand_seqs(rng, 2, 20000, 20000, b"AC") + [b"ACACACACACAC" * 100]
    check(ctx, oracle, seqs, 10, 2)


@pytest.mark.parametrize("k,e", [(12, 2), (10, 3), (7, 1), (13, 2)])
def test_long_and_short_mixed(ctx, oracle, k, e):
    # long sequences are cut into blocks of diagonals per row tile: interior blocks count
    # their hits in place, the blocks at a sequence's ends go through the hit queue; short
    # sequences in the same batch keep whole-text ranges.  A two-letter alphabet makes
    # occurrences dense.
    rng = np.random.default_rng(k * 10 + e)
    seqs = (rand_seqs(rng, 2, 17000, 26000, b"AC") + rand_seqs(rng, 5, 30, 900, b"ACGT") +
            rand_seqs(rng, 1, 9000, 9000, b"ACGT"))
    check(ctx, oracle, seqs, k, e)


def test_many_short_texts_keep_the_work_queue_small(ctx, monkeypatch):
    # pair-symmetric items hold at most three text ranges; with tens of thousands of short
    # texts that plan would need millions of items, so the ordered plan runs instead (even
    # when the symmetric one is asked for), and every factor still occurs in its own text
    monkeypatch.setenv("MOTEX_SYM", "1")
    rng = np.random.default_rng(7)
    seqs = rand_seqs(rng, 30000, 40, 40, b"ACGT")
    print(f"[hd] {len(seqs)} texts of 40 nt, k=12 e=2", flush=True)
    b = ctx.batch(seqs, MotexParams(k=12, e=2))
    st = b.run()
    u, v, off = b.download()
    b.close()
    assert st["symmetric"] == 0 and st["work_items"] < 2_000_000
    u, v = u.reshape(len(seqs), 40), v.reshape(len(seqs), 40)
    assert np.all(u[:, :11] == 0) and np.all(v[:, :11] == 0)  # no full window before k - 1
    full_u, full_v = u[:, 11:], v[:, 11:]
    assert np.all(full_u >= 1) and np.all(full_v >= full_u) and np.all(full_u <= len(seqs))


def test_empty_and_single(ctx, oracle):
    check(ctx, oracle, [b"ACGTACGTAC"], 4, 1)
    check(ctx, oracle, [b"ACG", b"AC"], 4, 1)   # every pair skipped


def test_plain_form_equals_batched(ctx, oracle):
    w = synth.dna50_hamming(scale=0.2)
    p = MotexParams(k=w.k, e=w.e)
    b = ctx.batch(w.seqs, p)
    b.run()
    ub, vb, off = b.download()
    b.clear()
    b.run_plain()
    up, vp, _ = b.download()
    assert np.array_equal(ub, up) and np.array_equal(vb, vp)
    b.close()


def test_extract_pair_accumulates(ctx, oracle):
    rng = np.random.default_rng(7)
    x, y = rand_seqs(rng, 2, 300, 600, b"ACG")
    p = MotexParams(k=7, e=2)
    u = np.full(len(x), 5, np.uint32)
    v = np.full(len(x), 9, np.uint32)
    ctx.extract_pair(x, y, p, u, v)
    ou, ov = oracle.pair(0, x, y, 7, 2)
    assert np.array_equal(u, ou + 5) and np.array_equal(v, ov + 9)


def test_invalid_params_raise(ctx):
    with pytest.raises(motex_amd.MotexError):
        ctx.extract_all_pairs([b"ACGT"], MotexParams(k=4, e=4))
    with pytest.raises(motex_amd.MotexError):
        ctx.extract_all_pairs([b"ACGT"], MotexParams(k=4, e=1, distance=2))


def test_u_counts_beyond_16_bits(ctx):
    """100,000 copies of AAAA, k=4, e=0: the one factor of every sequence occurs once in every
    sequence, so u = v = N exactly.  The ordered plan runs (occurrences are dense), and one
    work item then spans more than 65,535 texts: the per-row distinct-text count must not
    wrap (the reference keeps u as unsigned int, motex.cc:495-529)."""
    N = 100_000
    seqs = [b"AAAA"] * N
    print(f"[hd u-width] n={N} k=4 e=0", flush=True)
    b = ctx.batch(seqs, MotexParams(k=4, e=0))
    st = b.run()
    u, v, off = b.download()
    b.close()
    assert st["symmetric"] == 0
    u = u.reshape(N, 4)
    v = v.reshape(N, 4)
    assert np.all(u[:, :3] == 0) and np.all(v[:, :3] == 0)
    assert np.all(u[:, 3] == N), np.unique(u[:, 3])
    assert np.all(v[:, 3] == N), np.unique(v[:, 3])
