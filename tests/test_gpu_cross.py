"""motex_extract_cross (the -I re-search, motex.cc:940-990): every x against every text y
only, bit-exact with the oracle's per-pair operator summed over the texts."""
import numpy as np
import pytest

from motex_amd import MotexParams

pytestmark = pytest.mark.gpu


def rand_seqs(rng, n, lo, hi, alpha):
    a = np.frombuffer(alpha, np.uint8)
    return [a[rng.integers(0, len(a), rng.integers(lo, hi + 1))].tobytes() for _ in range(n)]


@pytest.mark.parametrize("d,k,e,alpha", [(0, 8, 1, b"ACGT"), (0, 12, 2, b"ACGT"), (1, 8, 1, b"ACGT"),
                                         (1, 6, 1, b"ARNDCQEGHILKMFPSTWYV"), (0, 70, 20, b"AC")])
def test_cross_equals_the_oracle(ctx, oracle, d, k, e, alpha):
    rng = np.random.default_rng(k * 10 + d)
    ys = rand_seqs(rng, 40, 2, 600, alpha)              # some texts shorter than k
    xs = [y[5:5 + k] for y in ys[:10] if len(y) >= k + 5]
    xs += rand_seqs(rng, 6, k, 3 * k, alpha)             # longer motifs: all their factors
    xs += [alpha[:1] * max(1, k - 2)]                    # shorter than k: no factor
    print(f"[cross] d={d} k={k} e={e} nx={len(xs)} ny={len(ys)}", flush=True)
    p = MotexParams(k=k, e=e, distance=d, alphabet="PROT" if len(alpha) == 20 else "DNA")
    u, v, off = ctx.extract_cross(xs, ys, p)
    for i, x in enumerate(xs):
        ou = np.zeros(len(x), np.uint32)
        ov = np.zeros(len(x), np.uint32)
        for y in ys:
            oracle.pair(d, x, y, k, e, ou, ov)
        assert np.array_equal(u[int(off[i]):int(off[i + 1])], ou), i
        assert np.array_equal(v[int(off[i]):int(off[i + 1])], ov), i
    assert v.sum() > 0
