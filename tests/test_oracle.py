"""Pins the CPU oracle (oracle/motex_oracle.c) before it is trusted as the checker.

Every check here is something other than the oracle's own recurrence: brute-force window
comparison, facts recorded from the reference's libflasm as a black box, closed forms
(e = 0 is exact k-mer counting), symmetry, and exhaustive equivalence of bitminmax's case
analysis with its closed form.
"""
import os

import numpy as np
import pytest

from motex_amd import synth


def brute_hd_matrix(x: bytes, y: bytes, k: int):
    """Hamming distances of every full window pair: D[px-k+1, pt-k+1]."""
    xa = np.frombuffer(x, np.uint8)
    ya = np.frombuffer(y, np.uint8)
    if len(xa) < k or len(ya) < k:
        return np.zeros((0, 0), np.int32)
    xw = np.lib.stride_tricks.sliding_window_view(xa, k)
    yw = np.lib.stride_tricks.sliding_window_view(ya, k)
    return (xw[:, None, :] != yw[None, :, :]).sum(axis=2)


def brute_hd_uv(seqs, k, e):
    out_u, out_v = [], []
    for x in seqs:
        u = np.zeros(len(x), np.int64)
        v = np.zeros(len(x), np.int64)
        for y in seqs:
            if len(x) < k or len(y) < k:
                continue
            H = brute_hd_matrix(x, y, k) <= e
            cnt = H.sum(axis=1)
            v[k - 1:] += cnt
            u[k - 1:] += cnt > 0
        out_u.append(u)
        out_v.append(v)
    return np.concatenate(out_u), np.concatenate(out_v)


def rand_seqs(rng, n, lo, hi, alpha):
    a = np.frombuffer(alpha, np.uint8)
    return [a[rng.integers(0, len(a), rng.integers(lo, hi + 1))].tobytes() for _ in range(n)]


@pytest.mark.parametrize("seed,k,e,alpha", [
    (1, 5, 1, b"ACGT"), (2, 8, 2, b"ACGTN"), (3, 3, 0, b"AC"), (4, 12, 3, b"ACGT"),
    (5, 1, 0, b"ACGT"), (6, 20, 6, b"ARNDCQEGHILKMFPSTWYV"), (7, 64, 20, b"AC"),
    (8, 65, 20, b"AC"), (9, 130, 45, b"AC"),   # beyond one 64-bit word: sliding window count
])
def test_hd_matches_brute_force(oracle, seed, k, e, alpha):
    rng = np.random.default_rng(seed)
    seqs = rand_seqs(rng, 6, max(1, k - 2), k + 60, alpha)
    u, v, off, _ = oracle.all_pairs(seqs, 0, k, e)
    bu, bv = brute_hd_uv(seqs, k, e)
    assert np.array_equal(u.astype(np.int64), bu)
    assert np.array_equal(v.astype(np.int64), bv)


def test_hd_probe_facts(oracle):
    # Recorded from the reference's libflasm (black box): text-start overhang not counted.
    A, B = b"TTTTTACGTC", b"CGTCGGGGGG"
    u, v, off, _ = oracle.all_pairs([A, B], 0, 5, 1)
    assert (u[9], v[9]) == (1, 1)          # ACGTC
    assert v[4] == 2                       # TTTTT
    assert v[10 + 9] == 3                  # GGGGG (B's last factor)


def test_hd_transpose_symmetry(oracle):
    rng = np.random.default_rng(11)
    x, y = rand_seqs(rng, 2, 40, 70, b"ACG")
    hxy = {(int(t), int(p)) for t, p, _ in oracle.hits(0, x, y, 6, 2)}
    hyx = {(int(p), int(t)) for t, p, _ in oracle.hits(0, y, x, 6, 2)}
    assert hxy == hyx and len(hxy) > 0


def exact_kmer_uv(seqs, k):
    out_u, out_v = [], []
    for x in seqs:
        u = np.zeros(len(x), np.int64)
        v = np.zeros(len(x), np.int64)
        for p in range(k - 1, len(x)):
            w = x[p - k + 1:p + 1]
            occ = [sum(1 for t in range(len(y) - k + 1) if y[t:t + k] == w) for y in seqs]
            v[p] = sum(occ)
            u[p] = sum(1 for o in occ if o)
        out_u.append(u)
        out_v.append(v)
    return np.concatenate(out_u), np.concatenate(out_v)


def test_e0_hamming_is_exact_kmer_counting(oracle):
    rng = np.random.default_rng(21)
    seqs = rand_seqs(rng, 5, 10, 50, b"AC")
    u, v, off, _ = oracle.all_pairs(seqs, 0, 4, 0)
    eu, ev = exact_kmer_uv(seqs, 4)
    assert np.array_equal(u.astype(np.int64), eu)
    assert np.array_equal(v.astype(np.int64), ev)


def test_e0_edit_hits_are_exact_occurrences(oracle):
    # A popcount-0 word means the last k operations were diagonal matches, so every edit
    # hit at e=0 is an exact occurrence.  The greedy bitminmax choice can carry error bits
    # from before an occurrence into it, so the converse does not hold: the MaxShift
    # recurrence may miss exact occurrences (seed 22 below has one).
    rng = np.random.default_rng(22)
    seqs = rand_seqs(rng, 5, 10, 50, b"AC")
    total_exact = total_hit = 0
    for x in seqs:
        for y in seqs:
            ex = {(t + 3, p) for p in range(3, len(x)) for t in range(len(y) - 3)
                  if y[t:t + 4] == x[p - 3:p + 1]}
            h = {(int(t), int(p)) for t, p, _ in oracle.hits(1, x, y, 4, 0)}
            assert h <= ex
            total_exact += len(ex)
            total_hit += len(h)
    assert total_hit < total_exact and total_hit > 0.9 * total_exact


def test_ed_probe_facts(oracle):
    # Recorded from libflasm (black box): at e=1 an exact occurrence inside a text yields
    # three end positions (i-1, i, i+1), and a length-(k-1) text prefix at the text start
    # is an occurrence.
    x = b"GGGGGACGTCGGGGG"
    y = b"TTTTTTTACGTCTTTTTTT"
    h = oracle.hits(1, x, y, 5, 1)
    ends = sorted(int(t) for t, p, _ in h if p == 9)
    assert ends == [10, 11, 12]
    h2 = oracle.hits(1, b"TTTTTACGTC", b"CGTCGGGGGG", 5, 1)
    assert any(int(t) == 3 and int(p) == 9 for t, p, _ in h2)


def _bmm_closed(a, b, c):
    pcs = [bin(w).count("1") for w in (a, b, c)]
    m = min(pcs)
    return max(w for w, p in zip((a, b, c), pcs) if p == m)


def test_bitminmax_closed_form_exhaustive(oracle):
    L = oracle.lib()
    for a in range(32):
        for b in range(32):
            for c in range(32):
                assert L.orc_bitminmax(a, b, c) == _bmm_closed(a, b, c)


def test_recurse_order(oracle):
    S = oracle.recurse([1, 5], [2, 7])
    assert S.tolist() == [[1, 5], [1, 6], [1, 7], [2, 5], [2, 6], [2, 7]]


def merge_reference_definition(x, y, k, e, boxes, box_ok):
    """u, v of one pair from per-box flag matrices via numpy shifts (one gap)."""
    (gmin, gmax, bl, be), = boxes
    m, n = len(x), len(y)
    ok0, ok1 = box_ok
    ns = gmax - gmin + 1
    u = np.zeros(ns * m, np.int64)
    v = np.zeros(ns * m, np.int64)
    I, J = np.nonzero(ok0)
    for kx in range(ns):
        dj = gmin + kx + bl
        tot = np.zeros(m + 1, np.int64)
        for ly in range(ns):
            di = gmin + ly + bl
            ii, jj = I + di, J + dj
            inb = (ii <= n) & (jj <= m)
            good = np.zeros_like(inb)
            good[inb] = ok1[ii[inb], jj[inb]]
            np.add.at(tot, J[good], 1)
        v[kx * m: (kx + 1) * m] += tot[1:]
        u[kx * m: (kx + 1) * m] += tot[1:] > 0
    return u, v


def merge_reference_definition_gaps(x, y, boxes, box_ok, S):
    """u, v of one pair from per-box flag matrices, any number of gaps: for x-side
    combination kx and y-side combination ky, box 0 ending at (I, J) (1-based) makes a chain
    when box b ends at I + sum(gy_c + len_c), J + sum(gx_c + len_c) for every b."""
    m, n = len(x), len(y)
    I, J = np.nonzero(box_ok[0])
    u = np.zeros(len(S) * m, np.int64)
    v = np.zeros(len(S) * m, np.int64)
    for kx, gx in enumerate(S):
        tot = np.zeros(m + 1, np.int64)
        for gy in S:
            ii, jj = I.copy(), J.copy()
            good = np.ones(len(I), bool)
            for b, (_, _, bl, _) in enumerate(boxes):
                ii = ii + gy[b] + bl
                jj = jj + gx[b] + bl
                inb = (ii <= n) & (jj <= m)
                good &= inb
                idx = np.nonzero(good)[0]
                good[idx] = box_ok[b + 1][ii[idx], jj[idx]]
            np.add.at(tot, J[good], 1)
        v[kx * m: (kx + 1) * m] += tot[1:]
        u[kx * m: (kx + 1) * m] += tot[1:] > 0
    return u, v


@pytest.mark.parametrize("distance,boxes", [
    (0, [(2, 4, 3, 1), (0, 2, 4, 1)]),
    (0, [(1, 2, 3, 0), (2, 3, 3, 1), (0, 1, 4, 1)]),
    (1, [(1, 3, 3, 1), (0, 2, 4, 1)]),
])
def test_structured_merge_definition_several_gaps(oracle, distance, boxes):
    # the odometer walk over several gaps against the definition: every x-side and y-side
    # gap combination tried on the box flag matrices (Hamming flags from brute-force windows)
    rng = np.random.default_rng(41 + len(boxes) + distance)
    x, y = rand_seqs(rng, 2, 90, 140, b"AC")
    k, e = 4, 1
    mats = []
    for ell, err in [(k, e)] + [(b[2], b[3]) for b in boxes]:
        ok = np.zeros((len(y) + 1, len(x) + 1), bool)
        if distance == 0:
            ok[ell:, ell:] = (brute_hd_matrix(x, y, ell) <= err).T
        else:
            for t, p, _ in oracle.hits(distance, x, y, ell, err):
                ok[t + 1, p + 1] = True
        mats.append(ok)
    S = oracle.recurse([b[0] for b in boxes], [b[1] for b in boxes]).tolist()
    u, v = oracle.struct_pair(distance, x, y, k, e, boxes)
    bu, bv = merge_reference_definition_gaps(x, y, boxes, mats, S)
    assert np.array_equal(u.astype(np.int64), bu)
    assert np.array_equal(v.astype(np.int64), bv)
    assert v.sum() > 0
    # a dropped constraint on the last box changes the counts
    loose = mats[:-1] + [np.ones_like(mats[-1])]
    assert not np.array_equal(merge_reference_definition_gaps(x, y, boxes, loose, S)[1], bv)


@pytest.mark.parametrize("distance", [0, 1])
def test_structured_merge_definition(oracle, distance):
    rng = np.random.default_rng(31 + distance)
    x, y = rand_seqs(rng, 2, 80, 120, b"ACG")
    k, e, boxes = 4, 1, [(2, 5, 3, 1)]
    mats = []
    for ell, err in [(k, e), (boxes[0][2], boxes[0][3])]:
        ok = np.zeros((len(y) + 1, len(x) + 1), bool)
        for t, p, _ in oracle.hits(distance, x, y, ell, err):
            ok[t + 1, p + 1] = True
        mats.append(ok)
    if distance == 0:
        # the per-box flags themselves come from brute-force windows
        D = brute_hd_matrix(x, y, k) <= e
        assert np.array_equal(mats[0][k:, k:], D.T)
    u, v = oracle.struct_pair(distance, x, y, k, e, boxes)
    bu, bv = merge_reference_definition(x, y, k, e, boxes, mats)
    assert np.array_equal(u.astype(np.int64), bu)
    assert np.array_equal(v.astype(np.int64), bv)
    assert v.sum() > 0


def test_writer_double_filter_and_dedup(oracle, tmp_path):
    # (29/100)*100 is 28.999999999999996 in IEEE double, so u=29 of N=100 fails q=29 --
    # integer arithmetic (29*100 >= 29*100) would pass it.  functions.cc:1107.
    seqs = [b"ACGTACGT"] + [b"TTTTTTTT"] * 99
    k = 4
    off = np.arange(0, 801, 8, dtype=np.uint64)[:100]
    u = np.zeros(800, np.uint32)
    v = np.zeros(800, np.uint32)
    u[3], v[3] = 29, 40        # ACGT at j=3 of seq 0
    u[4], v[4] = 30, 40        # CGTA
    u[7], v[7] = 30, 40        # ACGT again at j=7: the first occurrence that passes
    u[11], v[11] = 50, 60      # TTTT in seq 1 at j=3 ...
    u[12], v[12] = 51, 61      # ... and at j=4: deduplicated, the first one is printed
    out = str(tmp_path / "o.motex")
    sm = str(tmp_path / "o.smile")
    nmot = oracle.write(out, seqs, u, v, off, alphabet="DNA", d=0, k=k, e=1, q=29,
                        smile_path=sm)
    body = open(out).read().splitlines()
    motifs = [l for l in body if l and not l.startswith("#")]
    assert motifs == ["CGTA 30 100 0.300000 40", "ACGT 30 100 0.300000 40",
                      "TTTT 50 100 0.500000 60"]
    assert nmot == 3
    sl = open(sm).read().splitlines()
    assert len(sl[0]) == 79 and sl[0].startswith("%%% 1 29/100 800 4 4 1 alphabet ACGTN$")
    assert sl[1] == " " * 79 and sl[2] == " " * 79 and sl[3] == "=" * 79
    assert sl[4] == "CGTA 1230 30\t40"


def test_dna50_generator_is_deterministic():
    a = synth.dna50_hamming()
    b = synth.dna50_hamming()
    assert a.seqs == b.seqs
    assert len(a.seqs) == 50 and all(len(s) == 500 for s in a.seqs)
    import hashlib
    h = hashlib.sha256(b"".join(a.seqs)).hexdigest()
    golden = os.path.join(os.path.dirname(__file__), "golden", "dna50_sha256.txt")
    assert open(golden).read().strip() == h


def test_wide_bitminmax_equals_the_64_bit_one(oracle):
    # the 128-bit restatement used for edit factors up to k = 64 against the 64-bit one, on
    # words that fit 64 bits, and on words whose high halves alone decide
    L = oracle.lib()
    rng = np.random.default_rng(3)
    out = np.zeros(2, np.uint64)
    for _ in range(3000):
        a, b, c = (int(w) for w in rng.integers(0, 1 << 20, 3, dtype=np.uint64))
        a &= int(rng.integers(0, 1 << 20))
        L.orc_bitminmax_w(a, 0, b, 0, c, 0, out.ctypes.data_as(oracle.u64p))
        assert int(out[0]) == L.orc_bitminmax(a, b, c) and int(out[1]) == 0
        L.orc_bitminmax_w(0, a, 0, b, 0, c, out.ctypes.data_as(oracle.u64p))
        assert int(out[1]) == L.orc_bitminmax(a, b, c) and int(out[0]) == 0


@pytest.mark.parametrize("k", [33, 48, 64])
def test_e0_edit_hits_are_exact_occurrences_for_long_factors(oracle, k):
    # factors past one 64-bit word: every e=0 hit is an exact occurrence, and long exact
    # repeats are found
    rng = np.random.default_rng(k)
    base = rand_seqs(rng, 1, 3 * k, 3 * k, b"AC")[0]
    seqs = [base, base[k // 2:] + base[: k // 2], rand_seqs(rng, 1, 2 * k, 4 * k, b"AC")[0]]
    total = 0
    for x in seqs:
        for y in seqs:
            ex = {(t + k - 1, p) for p in range(k - 1, len(x)) for t in range(len(y) - k + 1)
                  if y[t:t + k] == x[p - k + 1:p + 1]}
            h = {(int(t), int(p)) for t, p, _ in oracle.hits(1, x, y, k, 0)}
            assert h <= ex
            total += len(h)
    assert total >= 3 * (2 * k + 1)   # at least the three self pairs' exact windows
