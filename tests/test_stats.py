"""Background/foreground statistics (-b/-u/-U, -I) of the motex command line, on the CPU.

- The oracle's restatement of the reference's MPFR p-value (oracle/stats.py: 200-bit values,
  every operation rounded upward in binomial.cc:39-74's order) is pinned against the exact
  rational binomial tail, to the 10 printed digits, wherever 200 bits carry those digits,
  and against closed forms.
- The command line's p-values (motex_amd/csrc/cli/stats.cpp, MPFR) equal the restatement's,
  including tiny tails where only 200-bit rounding decides the digits.
- The command line's writers, run by tests/cli_host_check.cpp with u/v from the oracle,
  write what the oracle's restatement of functions.cc:1613-2832 writes (masked: run date and
  timings).
"""
import os
import subprocess
from fractions import Fraction as F

import numpy as np
import pytest

from oracle import stats as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MASK = ("# Rundate:", "# Run on", "User time:")


def masked(text):
    return [l for l in text.splitlines() if not l.startswith(MASK)]


# ---------------------------------------------------------------------------------------
# the restatement itself
# ---------------------------------------------------------------------------------------

@pytest.mark.parametrize("x,N,p,want", [
    (0, 5, F(1, 3), "1.000000000"),            # no terms: 1 - 0
    (1, 2, F(1, 2), "7.500000000e-1"),         # 1 - (1/2)^2
    (10, 10, F(1, 2), "9.765625000e-4"),       # p^N
    (3, 5, F(1), "1.000000000"),               # p >= 1: the reference's cdf is 0
    (4, 4, F(1, 4), "3.906250000e-3"),         # (1/4)^4
    (1, 7, F(0), "0"),                         # p = 0: the i = 0 term is 1
])
def test_closed_forms(x, N, p, want):
    assert S.pvalue(x, N, p) == want


def grid():
    rng = np.random.default_rng(11)
    cases = []
    for _ in range(60):
        N = int(rng.integers(1, 120))
        x = int(rng.integers(0, N + 1))
        q = int(rng.integers(1, 100))
        cases.append((x, N, S.long_double(F(q, 100))))              # background quorum / size
        cases.append((x, N, F(float("%f" % rng.uniform(0, 1)))))    # a printed ratio
    return cases


def test_emulation_equals_exact_rational_tail():
    checked = 0
    for x, N, p in grid():
        ex = S.exact_tail(x, N, p)
        if ex < F(1, 10 ** 40):
            continue      # below 2^-133: the 200-bit cancellation 1 - cdf decides the digits
        assert S.out_str(S.mpfr_tail(x, N, p)) == S.out_str(ex), (x, N, p)
        # a dropped term changes the printed value (where it is not rounded away next to 1)
        if 0 < x < N and ex < F(999, 1000):
            assert S.out_str(S.exact_tail(x - 1, N, p)) != S.out_str(ex)
        checked += 1
    assert checked > 80


def test_long_double_rounding():
    assert S.long_double(F(1, 3)) == F(12297829382473034411, 2 ** 65)
    assert S.long_double(F(1, 2)) == F(1, 2)
    assert S.long_double(F(25, 100)) == F(1, 4)


def test_out_str_formats():
    assert S.out_str(F(1)) == "1.000000000"
    assert S.out_str(F(12345, 10)) == "1.234500000e3"
    assert S.out_str(F(1, 3)) == "3.333333334e-1"        # rounded up
    assert S.out_str(F(-1, 3)) == "-3.333333333e-1"      # toward +infinity
    assert S.out_str(F(99999999999, 10 ** 11)) == "1.000000000"


# ---------------------------------------------------------------------------------------
# the command line's sources, built on the host
# ---------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    from motex_amd import build as B
    cflags, lflags = B.mpfr_flags()
    if not cflags:
        pytest.skip("MPFR headers or libraries not installed")
    out = str(tmp_path_factory.mktemp("cli") / "cli_host_check")
    csrc = os.path.join(ROOT, "motex_amd", "csrc")
    srcs = [os.path.join(csrc, "cli", f) for f in
            ("app.cpp", "fasta.cpp", "switches.cpp", "writer.cpp", "writer_back.cpp", "stats.cpp")]
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", csrc, *cflags,
                    os.path.join(ROOT, "tests", "cli_host_check.cpp"), *srcs, *lflags, "-o", out], check=True)
    return out


def test_cli_pvalues_equal_the_restatement(host_check):
    rng = np.random.default_rng(5)
    lines, want = [], []
    for _ in range(50):
        N = int(rng.integers(1, 90))
        x = int(rng.integers(0, N + 1))
        a, b = int(rng.integers(1, 100)), int(rng.integers(1, 200))
        lines.append(f"{x} {N} q {a} {b}")
        want.append(S.pvalue(x, N, S.long_double(F(a, b))))
        r = "%f" % rng.uniform(0, 1)
        lines.append(f"{x} {N} d {r} 0")
        want.append(S.pvalue(x, N, F(float(r))))
    # tails far below 2^-200 (the reference prints 200-bit cancellation residue there)
    for x, N, a, b in [(60, 60, 1, 10), (80, 90, 1, 20), (45, 50, 3, 100)]:
        lines.append(f"{x} {N} q {a} {b}")
        want.append(S.pvalue(x, N, S.long_double(F(a, b))))
    r = subprocess.run([host_check, "tail"], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       check=True)
    assert r.stdout.split("\n")[:-1] == want


def rand_seqs(rng, n, lo, hi, alpha=b"ACGT"):
    a = np.frombuffer(alpha, np.uint8)
    return [a[rng.integers(0, len(a), rng.integers(lo, hi + 1))].tobytes() for _ in range(n)]


def write_fasta(path, seqs):
    with open(path, "wb") as f:
        for i, s in enumerate(seqs):
            f.write(b">s%d\n%s\n" % (i, s))


def run_host(host_check, cwd, u, v, args):
    uv = os.path.join(cwd, "uv.bin")
    np.concatenate([np.asarray(u, np.uint32), np.asarray(v, np.uint32)]).tofile(uv)
    r = subprocess.run([host_check, "run", uv] + args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.parametrize("boxes", [[], [(2, 4, 3, 1)]], ids=["single", "structured"])
def test_background_writers(host_check, oracle, tmp_path, boxes):
    rng = np.random.default_rng(9 if boxes else 8)
    k, e, q = 5, 1, 20
    bg_seqs = rand_seqs(rng, 30, 40, 90)
    fg_seqs = rand_seqs(rng, 25, 30, 80)
    write_fasta(tmp_path / "bg.fa", bg_seqs)
    write_fasta(tmp_path / "fg.fa", fg_seqs)
    base = ["-a", "DNA", "-d", "0", "-k", str(k), "-e", str(e), "-q", str(q)]
    if boxes:
        (tmp_path / "boxes.txt").write_text(f"{len(boxes)}\n" + "".join(f"{g[0]}\n{g[1]}\n{g[2]}\n{g[3]}\n" for g in boxes))
        base += ["-s", "boxes.txt"]
    # the background: an ordinary run over bg.fa
    bu, bv, boff, _ = oracle.all_pairs(bg_seqs, 0, k, e, boxes=boxes)
    run_host(host_check, tmp_path, bu, bv, ["-i", "bg.fa", "-o", "bg.motex"] + base)
    bg_text = (tmp_path / "bg.motex").read_text()
    # the foreground against it
    fu, fv, foff, _ = oracle.all_pairs(fg_seqs, 0, k, e, boxes=boxes)
    run_host(host_check, tmp_path, fu, fv, ["-i", "fg.fa", "-o", "fg.motex", "-b", "bg.motex", "-u", "un.motex",
                                            "-S", "fg.smile", "-U", "un.smile"] + base)
    sw = {"input_filename": "fg.fa", "output_filename": "fg.motex", "d": 0, "q": q, "k": k, "e": e,
          "boxes": boxes, "boxes_filename": "boxes.txt"}
    want_out, want_un = S.write_back(sw, fg_seqs, fu, fv, foff, bg_text)
    got_out = (tmp_path / "fg.motex").read_text()
    assert masked(got_out) == masked(want_out)
    assert masked((tmp_path / "un.motex").read_text()) == masked(want_un)
    # both kinds of line occur
    body = [l for l in masked(got_out) if l and not l.startswith("#")]
    assert any(l.split()[5:7] == ["0", "0"] for l in body)
    assert any(l.split()[5:7] != ["0", "0"] for l in body)
    # SMILE forms: the lines after the (shared) SMILE header
    s_out, s_un = S.write_back_smile(sw, fg_seqs, fu, fv, foff, bg_text, "ACGTN")
    def after_header(p):
        lines = p.read_text().splitlines(keepends=True)
        cut = next(i for i, l in enumerate(lines) if l.startswith("=" * 79)) + 1
        return masked("".join(lines[cut:]))
    assert after_header(tmp_path / "fg.smile") == masked(s_out)
    assert after_header(tmp_path / "un.smile") == masked(s_un)


def test_research_writer(host_check, oracle, tmp_path):
    rng = np.random.default_rng(12)
    k, e = 6, 1
    seqs = rand_seqs(rng, 20, 40, 120)
    write_fasta(tmp_path / "in.fa", seqs)
    # an unmatched-motif file as -u writes it: header block, motif lines, footer
    motifs = [s[i:i + k] for s in seqs[:6] for i in (0, 7)] + [b"ACGTAC", b"TTTTTT"]
    recs = [(m.decode(), int(rng.integers(1, 40)), 40, float("%f" % rng.uniform(0, 1)), int(rng.integers(1, 90)))
            for m in motifs]
    text = S.header({"input_filename": "old.fa", "output_filename": "old.motex", "d": 0, "q": 10, "k": k, "e": e}, 40)
    text += "".join(f"{m} {u} {n} {r:f} {v}\n" for m, u, n, r, v in recs)
    text += "\n#A total of 0 background motifs were read.\n"
    (tmp_path / "un.motex").write_text(text)
    # u/v of every motif against every input sequence
    xs = [r[0].encode() for r in S.read_motif_records(text)]
    off = np.zeros(len(xs) + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in xs])
    u = np.zeros(int(off[-1]), np.uint32)
    v = np.zeros(int(off[-1]), np.uint32)
    for i, x in enumerate(xs):
        for y in seqs:
            oracle.pair(0, x, y, k, e, u[int(off[i]):int(off[i + 1])], v[int(off[i]):int(off[i + 1])])
    run_host(host_check, tmp_path, u, v, ["-a", "DNA", "-i", "in.fa", "-o", "fore.motex", "-d", "0", "-k", str(k),
                                          "-e", str(e), "-q", "5", "-I", "un.motex"])
    sw = {"input_filename": "in.fa", "output_filename": "fore.motex", "d": 0, "q": 5, "k": k, "e": e}
    want = S.write_fore(sw, S.read_motif_records(text), u, v, off, len(seqs))
    got = (tmp_path / "fore.motex").read_text()
    assert masked(got) == masked(want)
    assert len([l for l in masked(got) if l and not l.startswith("#")]) >= 10
