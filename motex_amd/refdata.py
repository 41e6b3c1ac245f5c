"""The reference repository's own datasets, kept as inputs under tests/golden/refdata/.

- ``hsapiens.1200_genes.fa.bz2``: 1200 H. sapiens gene sequences of 1000 nt
  (data/hsapiens.1200_genes.fa.bz2; the 1200 x 1000-nt Hamming workload's shape on real
  sequence, with its repeats and skewed composition);
- ``dnc_subtilis_330-30.seq.bz2``: 1062 B. subtilis upstream regions of 40-300 nt, the base
  input of the reference's simulators (data/README:1-53);
- ``small.fa``: three short DNA sequences (data/small.fa);
- ``boxes.txt``: the reference's two-gap structure file (data/boxes.txt: gaps [16,17] and
  [15,16], boxes 6/1 and 7/2).

The FASTA parse keeps what the reference's reader (motex.cc:225-300) does to these files:
header lines are skipped, sequence lines are joined and upper-cased.  The files hold only
A/C/G/T, so the alphabet check never fires.
"""
from __future__ import annotations

import bz2
import os

from .synth import Workload

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DIR = os.path.join(ROOT, "tests", "golden", "refdata")


def _open(name: str) -> bytes:
    path = os.path.join(DIR, name)
    if name.endswith(".bz2"):
        with bz2.open(path, "rb") as f:
            return f.read()
    with open(path, "rb") as f:
        return f.read()


def read_fasta(name: str) -> list[bytes]:
    seqs, cur = [], []
    for line in _open(name).split(b"\n"):
        if line.startswith(b">"):
            if cur:
                seqs.append(b"".join(cur))
            cur = []
        else:
            cur.append(line.strip().upper())
    if cur and b"".join(cur):
        seqs.append(b"".join(cur))
    return [s for s in seqs if s]


def boxes() -> list[tuple[int, int, int, int]]:
    """data/boxes.txt as (gap_min, gap_max, box_len, box_err) per gap (motex.cc:303-357)."""
    vals = [int(t) for t in _open("boxes.txt").split()]
    ng = vals[0]
    return [tuple(vals[1 + 4 * g: 5 + 4 * g]) for g in range(ng)]


def hsapiens(count: int | None = None) -> Workload:
    """The first `count` (all 1200 by default) H. sapiens genes, Hamming k=12 e=2 q=25."""
    seqs = read_fasta("hsapiens.1200_genes.fa.bz2")
    if count is not None:
        seqs = seqs[:count]
    return Workload(f"{len(seqs)} H. sapiens genes x 1000 nt (reference dataset), Hamming k=12 e=2 q=25",
                    "DNA", seqs, 0, 12, 2, 25)


def subtilis() -> list[bytes]:
    return read_fasta("dnc_subtilis_330-30.seq.bz2")


def small() -> list[bytes]:
    return read_fasta("small.fa")
