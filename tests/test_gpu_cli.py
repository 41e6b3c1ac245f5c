"""End to end: the motex command line on the GPU produces the .motex and SMILE files the
reference writers would produce from the oracle's u/v (byte-identical apart from the run
date and timing lines)."""
import os
import subprocess

import numpy as np

import pytest

import motex_amd
from motex_amd import synth

pytestmark = pytest.mark.gpu

MASK = ("# Rundate:", "# Run on", "User time:")


def masked(path):
    return [l for l in open(path).read().splitlines() if not l.startswith(MASK)]


def run_both(tmp_path, oracle, w, smile=True):
    gdir = tmp_path / "gpu"
    odir = tmp_path / "oracle"
    gdir.mkdir()
    odir.mkdir()
    (gdir / "in.fa").write_bytes(w.fasta())
    args = [motex_amd.BIN_PATH, "-i", "in.fa", "-o", "out.motex"] + w.cli_args()
    if w.boxes:
        (gdir / "boxes.txt").write_text(w.boxes_text())
        args += ["-s", "boxes.txt"]
    if smile:
        args += ["-S", "out.smile"]
    print("[cli] " + " ".join(args[1:]), flush=True)
    r = subprocess.run(args, cwd=gdir, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    u, v, off, _ = oracle.all_pairs(w.seqs, w.distance, w.k, w.e, boxes=w.boxes)
    cwd = os.getcwd()
    os.chdir(odir)
    try:
        oracle.write("out.motex", w.seqs, u, v, off, alphabet=w.alphabet, d=w.distance, k=w.k,
                     e=w.e, q=w.q, Q=w.Q, n=w.n, N=w.N, boxes=w.boxes, input_filename="in.fa",
                     boxes_filename="boxes.txt", smile_path="out.smile" if smile else None)
    finally:
        os.chdir(cwd)
    assert masked(gdir / "out.motex") == masked(odir / "out.motex")
    if smile:
        assert masked(gdir / "out.smile") == masked(odir / "out.smile")
    return masked(gdir / "out.motex")


def test_dna50(tmp_path, oracle):
    w = synth.dna50_hamming()
    lines = run_both(tmp_path, oracle, w)
    motifs = [l.split()[0] for l in lines if l and not l.startswith("#")]
    # the planted 10-mers (in 60% of sequences, q=50) are reported
    for p in w.planted:
        assert p.decode() in motifs


def test_structured_dna1000_subset(tmp_path, oracle):
    w = synth.dna1000_structured(scale=0.03)
    w.q = 10
    run_both(tmp_path, oracle, w)


def test_protein_edit_subset(tmp_path, oracle):
    w = synth.prot500_edit(scale=0.02)
    w.q = 5
    run_both(tmp_path, oracle, w)


def test_quorum_zero_prints_every_factor(tmp_path, oracle):
    w = synth.dna50_hamming(scale=0.1)
    w.q, w.N = 0, 2000000000
    run_both(tmp_path, oracle, w, smile=False)


def test_long_mode_flag(tmp_path, oracle):
    # -L 1 on a 20-knt prefix of the long-sequence workload, every distinct factor printed
    w = synth.dna4m_long(scale=0.005)
    w.q, w.N = 0, 2000000000
    assert "-L" in w.cli_args()
    run_both(tmp_path, oracle, w, smile=False)


def test_reference_structured_example(tmp_path, oracle):
    # data/README step 6, structured: -d 0 -k 8 -e 1 -q 7 -s data/boxes.txt (on the first
    # 250 B. subtilis regions)
    from motex_amd import refdata
    w = synth.Workload("subtilis", "DNA", refdata.subtilis()[:250], 0, 8, 1, 7, boxes=refdata.boxes())
    run_both(tmp_path, oracle, w)


def test_hsapiens_genes(tmp_path, oracle):
    from motex_amd import refdata
    w = refdata.hsapiens(100)
    run_both(tmp_path, oracle, w)


def test_background_and_research_modes(tmp_path, oracle):
    """-b/-u/-U on the GPU against a background from an earlier run, then -I with the
    unmatched motifs, each equal to the oracle's writers on the oracle's u/v."""
    from oracle import stats as S
    bg = synth.dna50_hamming(scale=0.6)          # 30 x 500 nt, k=10 e=1
    fg = synth.dna50_hamming(scale=0.5)
    fg.seqs = [s[100:400] for s in fg.seqs]       # 25 x 300 nt, sharing the planted motifs
    for w in (bg, fg):
        w.k, w.e, w.q = 7, 1, 30
    (tmp_path / "bg.fa").write_bytes(bg.fasta())
    (tmp_path / "fg.fa").write_bytes(fg.fasta())
    base = ["-a", "DNA", "-d", "0", "-k", "7", "-e", "1", "-q", "30"]

    def motex(args):
        print("[cli] " + " ".join(args), flush=True)
        r = subprocess.run([motex_amd.BIN_PATH] + args, cwd=tmp_path, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr

    motex(["-i", "bg.fa", "-o", "bg.motex"] + base)
    motex(["-i", "fg.fa", "-o", "fg.motex", "-b", "bg.motex", "-u", "un.motex", "-S", "fg.smile",
           "-U", "un.smile"] + base)
    bg_text = (tmp_path / "bg.motex").read_text()
    fu, fv, foff, _ = oracle.all_pairs(fg.seqs, 0, 7, 1)
    sw = {"input_filename": "fg.fa", "output_filename": "fg.motex", "d": 0, "q": 30, "k": 7, "e": 1}
    want_out, want_un = S.write_back(sw, fg.seqs, fu, fv, foff, bg_text)
    assert masked(tmp_path / "fg.motex") == [l for l in want_out.splitlines() if not l.startswith(MASK)]
    assert masked(tmp_path / "un.motex") == [l for l in want_un.splitlines() if not l.startswith(MASK)]
    # -I: the unmatched motifs searched again in the background input
    motex(["-i", "bg.fa", "-o", "fore.motex", "-d", "0", "-k", "7", "-e", "1", "-q", "5", "-a", "DNA",
           "-I", "un.motex"])
    un_text = (tmp_path / "un.motex").read_text()
    recs = S.read_motif_records(un_text)
    assert recs
    xs = [r[0].encode() for r in recs]
    off = np.zeros(len(xs) + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in xs])
    u = np.zeros(int(off[-1]), np.uint32)
    v = np.zeros(int(off[-1]), np.uint32)
    for i, x in enumerate(xs):
        for y in bg.seqs:
            oracle.pair(0, x, y, 7, 1, u[int(off[i]):int(off[i + 1])], v[int(off[i]):int(off[i + 1])])
    sw2 = {"input_filename": "bg.fa", "output_filename": "fore.motex", "d": 0, "q": 5, "k": 7, "e": 1}
    want = S.write_fore(sw2, recs, u, v, off, len(bg.seqs))
    assert masked(tmp_path / "fore.motex") == [l for l in want.splitlines() if not l.startswith(MASK)]
