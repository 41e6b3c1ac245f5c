"""Seeded synthetic inputs for the five benchmark workloads.

Every byte is produced from SplitMix64 with integer-threshold sampling only (no floating
point in the sampled data), so a given (workload, seed) yields identical sequences on every
machine.  The base seed is ``0x4D6F54655800 + workload id`` ("MoTeX\\0" + id: 1, 2, 3,
0x4A, 0x4B in the order of WORKLOADS); serving jobs mix their job id into the seed.

The planted-motif scheme follows the intent of the reference's simulators
(data/motif_sim.c:165-236, data/str_motif_sim.c:215-346) but overwrites at a uniform
offset instead of appending, so the stated sequence lengths hold exactly.

Command line: ``python -m motex_amd.synth WORKLOAD OUT.fa [--scale F]`` writes a FASTA file
(and, for the structured workload, ``OUT.fa.boxes``).
"""
from __future__ import annotations

import argparse
import sys
from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
BASE_SEED = 0x4D6F54655800

DNA4 = b"ACGT"
PROT20 = b"ARNDCQEGHILKMFPSTWYV"

# Letter frequencies of the protein workload: one Dirichlet(alpha=2) draw over PROT20, fixed once and
# kept as integer weights out of 2**32 so that sampling stays integer-only.  (The draw was
# Gamma(2) = -ln(U1*U2) per letter from SplitMix64 seeded with BASE_SEED + 3 + 0x5EED,
# normalised and floored to 32-bit fixed point; the remainder went to the largest weight.)
PROT_DIRICHLET_W32 = [
    256742628, 272715792, 263472770, 86969794, 155231402, 345101415, 390606107,
    248880786, 76090129, 32618077, 440188179, 191763323, 302417674, 207097876,
    229021288, 46351697, 35100040, 231034516, 112674795, 370889008,
]


class SplitMix64:
    """Counter-based SplitMix64: output i = mix(seed + (i + 1) * golden)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.ctr = 0

    def draw(self, n: int) -> np.ndarray:
        idx = np.arange(self.ctr + 1, self.ctr + 1 + n, dtype=np.uint64)
        self.ctr += n
        with np.errstate(over="ignore"):
            z = self.seed + idx * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def one(self) -> int:
        return int(self.draw(1)[0])

    def uniform_int(self, lo: int, hi: int, n: int | None = None):
        """Integers in [lo, hi] (inclusive) by modulo reduction."""
        span = np.uint64(hi - lo + 1)
        if n is None:
            return lo + int(self.draw(1)[0] % span)
        return (self.draw(n) % span).astype(np.int64) + lo

    def symbols(self, n: int, alphabet: bytes, weights32: list[int] | None = None) -> np.ndarray:
        """n symbols; weights32 are integer weights summing to 2**32 (uniform if None)."""
        a = np.frombuffer(alphabet, dtype=np.uint8)
        r = self.draw(n) >> np.uint64(32)
        if weights32 is None:
            return a[(r * np.uint64(len(a))) >> np.uint64(32)]
        cum = np.cumsum(np.asarray(weights32, dtype=np.uint64))
        return a[np.searchsorted(cum, r, side="right")]

    def choose(self, n_total: int, n_pick: int) -> np.ndarray:
        """n_pick distinct indices from range(n_total) (partial Fisher-Yates)."""
        perm = np.arange(n_total, dtype=np.int64)
        for i in range(n_pick):
            j = i + int(self.draw(1)[0] % np.uint64(n_total - i))
            perm[i], perm[j] = perm[j], perm[i]
        return np.sort(perm[:n_pick])


def _weights(probs: list[float]) -> list[int]:
    w = [int(round(p * (1 << 32))) for p in probs]
    w[-1] += (1 << 32) - sum(w)
    return w


def _prot_weights() -> list[int]:
    assert sum(PROT_DIRICHLET_W32) == 1 << 32
    return PROT_DIRICHLET_W32


def substitute(rng: SplitMix64, word: np.ndarray, nsub: int, alphabet: bytes) -> np.ndarray:
    """Change nsub distinct positions of word to a different symbol of alphabet."""
    w = word.copy()
    if nsub <= 0:
        return w
    a = np.frombuffer(alphabet, dtype=np.uint8)
    for p in rng.choose(len(w), nsub):
        cur = int(np.nonzero(a == w[p])[0][0]) if np.any(a == w[p]) else 0
        step = 1 + rng.uniform_int(0, len(a) - 2)
        w[p] = a[(cur + step) % len(a)]
    return w


def edit_once(rng: SplitMix64, word: np.ndarray, alphabet: bytes) -> np.ndarray:
    """One substitution, insertion or deletion (probability 1/3 each)."""
    kind = rng.uniform_int(0, 2)
    if kind == 0:
        return substitute(rng, word, 1, alphabet)
    a = np.frombuffer(alphabet, dtype=np.uint8)
    if kind == 1:
        p = rng.uniform_int(0, len(word))
        c = a[rng.uniform_int(0, len(a) - 1)]
        return np.concatenate([word[:p], np.array([c], dtype=np.uint8), word[p:]])
    p = rng.uniform_int(0, len(word) - 1)
    return np.concatenate([word[:p], word[p + 1:]])


def plant(rng: SplitMix64, seq: np.ndarray, copy: np.ndarray) -> None:
    """Overwrite seq at a uniform offset in [0, len(seq) - len(copy)]."""
    if len(copy) > len(seq):
        return
    off = rng.uniform_int(0, len(seq) - len(copy))
    seq[off:off + len(copy)] = copy


@dataclass
class Workload:
    name: str
    alphabet: str          # "DNA" | "PROT"
    seqs: list[bytes]
    distance: int
    k: int
    e: int
    q: int
    Q: int = 100
    n: int = 1
    N: int = 10000
    boxes: list[tuple[int, int, int, int]] = field(default_factory=list)  # (gmin,gmax,len,err)
    long_mode: bool = False
    planted: list[bytes] = field(default_factory=list)

    @property
    def lens(self) -> list[int]:
        return [len(s) for s in self.seqs]

    def cells(self) -> int:
        """Algorithmic cells: sum over ordered pairs of (m - k + 1) * n."""
        L = np.array(self.lens, dtype=np.int64)
        tot = self.k
        for g in self.boxes:
            tot += g[0] + g[2]
        ok = L >= tot
        return int(np.sum(np.maximum(L[ok] - self.k + 1, 0)) * np.sum(L[ok]))

    def fasta(self) -> bytes:
        out = []
        for i, s in enumerate(self.seqs):
            out.append(b">seq%d\n" % i)
            for p in range(0, len(s), 60):
                out.append(s[p:p + 60] + b"\n")
        return b"".join(out)

    def boxes_text(self) -> str:
        lines = [str(len(self.boxes))]
        for g in self.boxes:
            lines += [str(x) for x in g]
        return "\n".join(lines) + "\n"

    def cli_args(self) -> list[str]:
        a = ["-a", self.alphabet, "-d", str(self.distance), "-k", str(self.k),
             "-e", str(self.e), "-q", str(self.q)]
        if self.Q != 100:
            a += ["-Q", str(self.Q)]
        if self.n != 1:
            a += ["-n", str(self.n)]
        if self.N != 10000:
            a += ["-N", str(self.N)]
        if self.long_mode:
            a += ["-L", "1"]
        return a


def _iid(rng, count, length, alphabet, weights=None):
    return [rng.symbols(length, alphabet, weights) for _ in range(count)]


def dna50_hamming(scale: float = 1.0) -> Workload:
    rng = SplitMix64(BASE_SEED + 1)
    nseq = max(2, int(round(50 * scale)))
    seqs = _iid(rng, nseq, 500, DNA4)
    motifs = [rng.symbols(10, DNA4) for _ in range(3)]
    for mo in motifs:
        for i in rng.choose(nseq, max(1, (nseq * 60) // 100)):
            plant(rng, seqs[i], substitute(rng, mo, rng.uniform_int(0, 1), DNA4))
    return Workload("50 DNA x 500 nt, Hamming k=10 e=1 q=50", "DNA",
                    [s.tobytes() for s in seqs], 0, 10, 1, 50,
                    planted=[m.tobytes() for m in motifs])


def dna1200_hamming(scale: float = 1.0) -> Workload:
    rng = SplitMix64(BASE_SEED + 2)
    nseq = max(2, int(round(1200 * scale)))
    w = _weights([0.295, 0.205, 0.205, 0.295])
    seqs = _iid(rng, nseq, 1000, DNA4, w)
    motifs = [rng.symbols(12, DNA4) for _ in range(5)]
    for mo in motifs:
        frac = rng.uniform_int(30, 70)
        for i in rng.choose(nseq, max(1, (nseq * frac) // 100)):
            plant(rng, seqs[i], substitute(rng, mo, rng.uniform_int(0, 2), DNA4))
    return Workload("1200 DNA x 1000 nt GC41%, Hamming k=12 e=2 q=25", "DNA",
                    [s.tobytes() for s in seqs], 0, 12, 2, 25,
                    planted=[m.tobytes() for m in motifs])


def prot500_edit(scale: float = 1.0) -> Workload:
    rng = SplitMix64(BASE_SEED + 3)
    nseq = max(2, int(round(500 * scale)))
    w = _prot_weights()
    lens = rng.uniform_int(200, 2000, nseq)
    seqs = [rng.symbols(int(L), PROT20, w) for L in lens]
    motifs = [rng.symbols(8, PROT20, w) for _ in range(4)]
    for mo in motifs:
        for i in rng.choose(nseq, max(1, (nseq * 40) // 100)):
            c = edit_once(rng, mo, PROT20) if rng.uniform_int(0, 1) else mo
            plant(rng, seqs[i], c)
    return Workload("500 PROT x U[200,2000], edit k=8 e=1 q=30", "PROT",
                    [s.tobytes() for s in seqs], 1, 8, 1, 30,
                    planted=[m.tobytes() for m in motifs])


def dna1000_structured(scale: float = 1.0) -> Workload:
    rng = SplitMix64(BASE_SEED + 0x4A)
    nseq = max(2, int(round(1000 * scale)))
    seqs = _iid(rng, nseq, 600, DNA4)
    b1, b2 = rng.symbols(8, DNA4), rng.symbols(8, DNA4)
    for i in rng.choose(nseq, max(1, nseq // 2)):
        gap = rng.uniform_int(10, 30)
        c1 = substitute(rng, b1, rng.uniform_int(0, 1), DNA4)
        c2 = substitute(rng, b2, rng.uniform_int(0, 1), DNA4)
        copy = np.concatenate([c1, rng.symbols(gap, DNA4), c2])
        plant(rng, seqs[i], copy)
    return Workload("1000 DNA x 600 nt, structured 8+[10,30]+8 e=1 q=40", "DNA",
                    [s.tobytes() for s in seqs], 0, 8, 1, 40, boxes=[(10, 30, 8, 1)],
                    planted=[b1.tobytes(), b2.tobytes()])


def dna4m_long(scale: float = 1.0) -> Workload:
    rng = SplitMix64(BASE_SEED + 0x4B)
    n = max(1000, int(round(4_000_000 * scale)))
    w = _weights([0.285, 0.215, 0.215, 0.285])
    seq = rng.symbols(n, DNA4, w)
    mo = rng.symbols(10, DNA4)
    copies = 100 if n >= 100_000 else max(1, n // 1000)
    seg = n // copies
    for c in range(copies):
        cp = substitute(rng, mo, rng.uniform_int(0, 2), DNA4)
        off = c * seg + rng.uniform_int(0, seg - len(cp))
        seq[off:off + len(cp)] = cp
    return Workload("1 DNA x 4 Mnt GC43%, Hamming k=10 e=2, long-sequence mode",
                    "DNA", [seq.tobytes()], 0, 10, 2, 0, long_mode=True,
                    planted=[mo.tobytes()])


def serving_job(job_id: int, scale: float = 1.0) -> Workload:
    rng = SplitMix64(BASE_SEED + 5 + (job_id << 20))
    nseq = max(2, int(round(rng.uniform_int(20, 400) * scale)))
    dna = rng.uniform_int(0, 1) == 0
    k = rng.uniform_int(8, 14)
    e = rng.uniform_int(0, 3)
    d = rng.uniform_int(0, 1)
    q = rng.uniform_int(20, 80)
    lens = rng.uniform_int(100, 2000, nseq)
    alpha = DNA4 if dna else PROT20
    seqs = [rng.symbols(int(L), alpha).tobytes() for L in lens]
    return Workload(f"job{job_id}", "DNA" if dna else "PROT", seqs, d, k, e, q)


# Workloads by name: sequences and sizes, distance, alphabet.
WORKLOADS = {"dna50": dna50_hamming, "dna1200": dna1200_hamming, "prot500": prot500_edit,
             "struct1000": dna1000_structured, "long4m": dna4m_long}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("workload", choices=sorted(WORKLOADS) + ["job"])
    ap.add_argument("out")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--job", type=int, default=0)
    a = ap.parse_args(argv)
    w = serving_job(a.job, a.scale) if a.workload == "job" else WORKLOADS[a.workload](a.scale)
    with open(a.out, "wb") as f:
        f.write(w.fasta())
    if w.boxes:
        with open(a.out + ".boxes", "w") as f:
            f.write(w.boxes_text())
    print(" ".join(w.cli_args()))
    return 0


if __name__ == "__main__":
    sys.exit(main())
