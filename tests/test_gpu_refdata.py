"""GPU parity on the reference repository's own datasets (data/README:1-53), bit-exact with
the oracle: real genes with repeats and skewed composition, the B. subtilis regions of the
reference's examples with its structure file, and data/small.fa."""
import numpy as np
import pytest

from motex_amd import MotexParams, refdata

pytestmark = pytest.mark.gpu


def check(ctx, oracle, seqs, d, k, e, boxes=(), tag=""):
    print(f"[refdata] {tag} n={len(seqs)} total={sum(map(len, seqs))} d={d} k={k} e={e} "
          f"boxes={list(boxes)}", flush=True)
    p = MotexParams(k=k, e=e, distance=d, boxes=list(boxes))
    b = ctx.batch(seqs, p)
    st = b.run()
    u, v, off = b.download()
    b.close()
    ou, ov, ooff, _ = oracle.all_pairs(seqs, d, k, e, boxes=list(boxes))
    bad = np.nonzero((u != ou) | (v != ov))[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[:5]}: gpu {u[bad[:5]]},{v[bad[:5]]} oracle {ou[bad[:5]]},{ov[bad[:5]]}"
    return st, u, v


@pytest.mark.parametrize("sym", ["0", "1"], ids=["ordered", "symmetric"])
def test_hsapiens_hamming(ctx, oracle, monkeypatch, sym):
    monkeypatch.setenv("MOTEX_SYM", sym)
    w = refdata.hsapiens(100)
    st, u, v = check(ctx, oracle, w.seqs, 0, 12, 2, tag="hsapiens[:100]")
    assert st["symmetric"] == int(sym)


def test_hsapiens_edit(ctx, oracle):
    w = refdata.hsapiens(40)
    check(ctx, oracle, w.seqs, 1, 8, 1, tag="hsapiens[:40]")


def test_subtilis_hamming(ctx, oracle):
    # the reference's single-motif example: -d 0 -k 8 -e 1 (data/README step 6)
    check(ctx, oracle, refdata.subtilis(), 0, 8, 1, tag="subtilis")


def test_subtilis_structured_hamming(ctx, oracle):
    # the reference's structured example with data/boxes.txt (two gaps, boxes 6/1 and 7/2)
    check(ctx, oracle, refdata.subtilis()[:250], 0, 8, 1, boxes=refdata.boxes(), tag="subtilis[:250]")


def test_subtilis_structured_edit(ctx, oracle):
    check(ctx, oracle, refdata.subtilis()[:120], 1, 8, 1, boxes=refdata.boxes(), tag="subtilis[:120]")


@pytest.mark.parametrize("d", [0, 1])
def test_small(ctx, oracle, d):
    seqs = refdata.small()
    check(ctx, oracle, seqs, d, 5, 1, tag="small.fa")
    # the structure (5 + 16 + 6 + 15 + 7 = 49 symbols) skips the pairs with the 21- and
    # 47-symbol sequences
    check(ctx, oracle, seqs, d, 5, 1, boxes=refdata.boxes(), tag="small.fa")
