"""motex_extract_all_pairs_noalloc (no device allocation, static per-device scratch array)
gives exactly what motex_extract_all_pairs gives, call after call: the array keeps the last
call's contents, so a second, smaller call must not see the first call's counts."""
import numpy as np
import pytest

import motex_amd
from motex_amd import MotexParams, synth

pytestmark = pytest.mark.gpu


def both(ctx, w):
    p = MotexParams(k=w.k, e=w.e)
    print(f"[noalloc] n={len(w.seqs)} total={sum(map(len, w.seqs))} k={w.k} e={w.e}", flush=True)
    u, v, off = ctx.extract_all_pairs_noalloc(w.seqs, p)
    ru, rv, roff = ctx.extract_all_pairs(w.seqs, p)
    assert np.array_equal(off, roff)
    assert np.array_equal(u, ru) and np.array_equal(v, rv)
    return u, v


def test_noalloc_equals_the_allocating_entry_point_twice(ctx):
    big = synth.dna1200_hamming()          # 1200 x 1000 nt, k=12 e=2
    small = synth.dna50_hamming(scale=0.4)  # 20 x 500 nt, k=10 e=1
    u1, v1 = both(ctx, big)
    assert v1.sum() > 0
    u2, v2 = both(ctx, small)
    assert v2.sum() > 0
    # and once more at the large size after the small one
    both(ctx, synth.dna1200_hamming(scale=0.5))


def test_noalloc_accumulates_and_skips(ctx):
    seqs = [b"ACGTACGTAC", b"ACGTTCGTAC", b"ACG"]
    p = MotexParams(k=4, e=1)
    off = motex_amd.uv_layout(seqs)
    u = np.full(int(off[-1]), 3, np.uint32)
    v = np.full(int(off[-1]), 5, np.uint32)
    lib = motex_amd.lib()
    sa, lens = motex_amd._seq_args(seqs)
    pc, keep = p.to_c()
    import ctypes as C
    motex_amd._check(lib.motex_extract_all_pairs_noalloc(ctx.h, sa, lens.ctypes.data_as(motex_amd.u32p), len(seqs),
                                                         C.byref(pc), motex_amd._row_ptrs(u, off),
                                                         motex_amd._row_ptrs(v, off)))
    ru, rv, _ = ctx.extract_all_pairs(seqs, p)
    assert np.array_equal(u, ru + 3) and np.array_equal(v, rv + 5)


def test_noalloc_refuses_what_it_does_not_cover(ctx):
    with pytest.raises(motex_amd.MotexError) as ei:
        ctx.extract_all_pairs_noalloc([b"ACGTACGT"] * 3, MotexParams(k=4, e=1, distance=1))
    assert ei.value.code == -4
    # 8,000 x 1000 nt needs more than the 64 MiB array: refused before any device work
    rng = np.random.default_rng(0)
    a = np.frombuffer(b"ACGT", np.uint8)
    seqs = [a[rng.integers(0, 4, 1000)].tobytes() for _ in range(8000)]
    with pytest.raises(motex_amd.MotexError) as ei:
        ctx.extract_all_pairs_noalloc(seqs, MotexParams(k=12, e=2))
    assert ei.value.code == -2 and "scratch" in str(ei.value)
