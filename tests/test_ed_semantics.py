"""How far the oracle's edit distance (the reference's in-tree MaxShift recurrence,
functions.cc:437-447 with bitminmax 923-967) is from exact semi-global edit distance.

libflasm's flasm_ed (libflasm.h:71) ships only as a binary, which is never run here, so
which of the two it computes is not pinned.  This test measures the gap with a brute-force
Sellers dynamic program: for the factor f = x[px-k+1 .. px] and text position pt, the
least edit distance between f and any text factor ending at pt,

    D[0][t] = 0,  D[i][0] = i,
    D[i][t] = min(D[i-1][t-1] + (f[i-1] != y[t-1]), D[i-1][t] + 1, D[i][t-1] + 1),

an occurrence being D[k][pt+1] <= e.  The figures are pinned so that a change of either side
shows, and the same figures are stated where edit-distance results are reported.
"""
import numpy as np
import pytest

from motex_amd import synth


def sellers_hits(x: bytes, y: bytes, k: int, e: int) -> set:
    """{(pos_t, pos_x)} with exact semi-global edit distance <= e (all factors at once)."""
    xa = np.frombuffer(x, np.uint8)
    ya = np.frombuffer(y, np.uint8)
    if len(xa) < k or len(ya) == 0:
        return set()
    F = np.lib.stride_tricks.sliding_window_view(xa, k)   # factor f ends at px = row + k - 1
    n = len(ya)
    t = np.arange(n + 1)
    D = np.zeros((len(F), n + 1), np.int32)                # row 0: free start anywhere
    for i in range(1, k + 1):
        a = np.empty_like(D)
        a[:, 0] = i
        a[:, 1:] = np.minimum(D[:, :-1] + (F[:, i - 1][:, None] != ya[None, :]), D[:, 1:] + 1)
        # D[i][t] = min_s<=t (a[s] + t - s): the insertion chain along the row
        D = (np.minimum.accumulate(a - t, axis=1) + t).astype(np.int32)
    r, c = np.nonzero(D[:, 1:] <= e)
    return {(int(pt), int(row + k - 1)) for row, pt in zip(r, c)}


def maxshift_hits(oracle, x, y, k, e) -> set:
    return {(int(t), int(p)) for t, p, _ in oracle.hits(1, x, y, k, e)}


def compare(oracle, seqs, k, e):
    ms = ex = both = 0
    for x in seqs:
        for y in seqs:
            a = maxshift_hits(oracle, x, y, k, e)
            b = sellers_hits(x, y, k, e)
            ms += len(a)
            ex += len(b)
            both += len(a & b)
    return ms, ex, both


def test_sellers_brute_force_on_known_cases():
    # exact occurrence inside a text: the three end positions of e=1 (substitution at the
    # end, the match, a trailing insertion)
    h = sellers_hits(b"GGGGGACGTCGGGGG", b"TTTTTTTACGTCTTTTTTT", 5, 1)
    assert sorted(t for t, p in h if p == 9) == [10, 11, 12]
    # a length-(k-1) text prefix is one deletion away
    assert (3, 9) in sellers_hits(b"TTTTTACGTC", b"CGTCGGGGGG", 5, 1)
    # e = 0 is exact matching
    x, y = b"ACGTACGGTACG", b"TTACGTACGGAC"
    h = sellers_hits(x, y, 4, 0)
    assert h == {(t + 3, p) for p in range(3, len(x)) for t in range(len(y) - 3) if y[t:t + 4] == x[p - 3:p + 1]}


@pytest.mark.parametrize("shape,seqs_fn,k,e,expected", [
    # the 500-protein edit workload's shape: proteins of U[200,2000] (6 of them), k=8, e=1
    ("protein", lambda: synth.prot500_edit(scale=0.012).seqs, 8, 1, (20726, 20723, 20722)),
    # 12 random 60-nt DNA sequences, k=6, e=1
    ("dna", lambda: [np.frombuffer(b"ACGT", np.uint8)[np.random.default_rng(3).integers(0, 4, (12, 60))[i]].tobytes()
                     for i in range(12)], 6, 1, (6637, 6018, 5929)),
])
def test_maxshift_versus_exact_edit_distance(oracle, shape, seqs_fn, k, e, expected):
    ms, ex, both = compare(oracle, seqs_fn(), k, e)
    print(f"{shape}: MaxShift {ms} hits, exact {ex}, common {both}, "
          f"MaxShift only {ms - both}, exact only {ex - both}")
    assert (ms, ex, both) == expected
