"""CPU restatement of MoTeX's background/foreground statistics -- test infrastructure only.

Imported by tests/ as the checker of the motex command line's -b/-u/-U and -I outputs; the
product never imports it.

- ``mpfr_tail``: the p-value of binomial.cc:25-74 (the MPFR build) evaluated the way MPFR
  evaluates it: every value a 200-bit binary number, every operation (mpfr_mul_ui,
  mpfr_pow_ui, mpfr_div, mpfr_mul, mpfr_add, mpfr_ui_sub) exact and then rounded upward, in
  the reference's order; MPFR rounds each operation correctly, so this is bit-for-bit what
  it computes.  ``exact_tail`` is the same sum in exact rational arithmetic: the two print the
  same 10 digits wherever 200 bits are enough, which pins the formula (tests/test_stats.py).
- ``out_str``: mpfr_out_str(f, 10, 10, x, round up) (mpfr-3.1.2 src/out_str.c: the leading
  digit, '.', nine more digits, then "e<exp>" unless the exponent is 0).
- ``long_double``: a value rounded to the 64-bit significand of x87 long double, as the
  reference's (long double) casts and divisions produce it.
- writers: functions.cc:1613-1905 / 1906-2241 (write_[structured_]motifs_back),
  2243-2466 / 2471-2745 (the SMILE forms), 2750-2832 (write_motifs_fore), with the .motex
  header of functions.cc:1010-1068 and the motif-line reading of motex.cc:883-925.
"""
from __future__ import annotations

from fractions import Fraction
from math import ceil, floor

PREC = 200          # ACC, motexdefs.h:82
LINE_SIZE = 512     # motexdefs.h:20


def _exp2(x: Fraction) -> int:
    """E with 2**(E-1) <= x < 2**E, for x > 0."""
    n, d = x.numerator, x.denominator
    E = n.bit_length() - d.bit_length()
    while Fraction(2) ** E <= x:
        E += 1
    while Fraction(2) ** (E - 1) > x:
        E -= 1
    return E


def round_up(x: Fraction, prec: int = PREC) -> Fraction:
    """x rounded toward +infinity to a prec-bit binary significand."""
    if x == 0:
        return Fraction(0)
    mag = abs(x)
    s = prec - _exp2(mag)
    scaled = mag * (Fraction(2) ** s)
    m = ceil(scaled) if x > 0 else floor(scaled)
    return (1 if x > 0 else -1) * Fraction(m) / (Fraction(2) ** s)


def long_double(x: Fraction) -> Fraction:
    """x rounded to nearest (ties to even) with a 64-bit significand."""
    if x == 0:
        return Fraction(0)
    mag = abs(x)
    s = 64 - _exp2(mag)
    scaled = mag * (Fraction(2) ** s)
    lo = floor(scaled)
    rem = scaled - lo
    m = lo + (1 if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and lo % 2) else 0)
    return (1 if x > 0 else -1) * Fraction(m) / (Fraction(2) ** s)


_lut_cache: dict[int, list[Fraction]] = {}


def _lut(n: int) -> list[Fraction]:
    """mpfr_fillTable (binomial.cc:25-37): LUT[0] = 1, LUT[i] = LUT[i-1] * i, rounded up."""
    lut = _lut_cache.get(0, [Fraction(1)])
    while len(lut) <= n:
        lut.append(round_up(lut[-1] * len(lut)))
    _lut_cache[0] = lut
    return lut


_pow_cache: dict[tuple[Fraction, int], tuple[list[Fraction], list[Fraction]]] = {}


def _powers(p: Fraction, N: int):
    """mpfr_pow_ui(p, i) and mpfr_pow_ui(1 - p, N - i), rounded up, for i = 0..N."""
    key = (p, N)
    if key not in _pow_cache:
        mp = round_up(p)
        om = round_up(1 - mp)
        a = [round_up(mp ** i) for i in range(N + 1)]
        b = [round_up(om ** (N - i)) for i in range(N + 1)]
        _pow_cache[key] = (a, b)
    return _pow_cache[key]


def mpfr_tail(x: int, N: int, p: Fraction) -> Fraction:
    """1 - mpfr_binomial_cdf_less_than(x, N, p) as the reference's MPFR build computes it
    (binomial.cc:39-74, then mpfr_ui_sub(., 1, ., RNDU))."""
    t = Fraction(0)
    if p < 1:
        lut = _lut(N)
        pa, pb = _powers(p, N)
        for i in range(min(x, N + 1)):
            a = pa[i]
            b = pb[i]
            pr = lut[N]
            pr = round_up(pr / lut[i])
            pr = round_up(pr / lut[N - i])
            pr = round_up(pr * a)
            pr = round_up(pr * b)
            t = round_up(t + pr)
    return round_up(1 - t)


def exact_tail(x: int, N: int, p: Fraction) -> Fraction:
    """P[X >= x] for X ~ Bin(N, p), exactly (0 terms beyond N; 1 for p >= 1... as the
    reference, whose cdf is 0 there)."""
    if p >= 1:
        return Fraction(1)
    from math import comb
    return 1 - sum(comb(N, i) * p ** i * (1 - p) ** (N - i) for i in range(min(x, N + 1)))


def out_str(v: Fraction) -> str:
    """mpfr_out_str(stream, 10, 10, v, MPFR_RNDU)."""
    if v == 0:
        return "0"
    mag = abs(v)
    e = 0
    while Fraction(10) ** e <= mag:
        e += 1
    while Fraction(10) ** (e - 1) > mag:
        e -= 1
    scaled = mag * Fraction(10) ** (10 - e)
    D = ceil(scaled) if v > 0 else floor(scaled)
    if D == 10 ** 10:
        D //= 10
        e += 1
    d = str(D)
    s = ("-" if v < 0 else "") + d[0] + "." + d[1:]
    return s + (f"e{e - 1}" if e - 1 != 0 else "")


def pvalue(x: int, N: int, p: Fraction) -> str:
    return out_str(mpfr_tail(x, N, p))


# ---------------------------------------------------------------------------------------
# writers
# ---------------------------------------------------------------------------------------

def c_lf(x: float) -> str:
    return "%f" % x


def header(sw: dict, num_seqs: int, exectime: float = 0.0, P: int = 1) -> str:
    """write_motex_header, functions.cc:1010-1068 (Rundate and timing lines are masked by the
    tests)."""
    out = ["####################################\n", "# Program: MoTeX\n", "# Rundate: -\n",
           f"# Input file: {sw['input_filename']}\n", f"# Output file: {sw['output_filename']}\n",
           f"# For {num_seqs} input sequences\n", f"#     d = {sw['d']}\n", f"#     q = {sw['q']}\n"]
    if sw.get("Q", 100) != 100:
        out.append(f"#     Q = {sw['Q']}\n")
    if sw.get("n", 1) != 1:
        out.append(f"#     n = {sw['n']}\n")
    if sw.get("N", 10000) != 10000:
        out.append(f"#     N = {sw['N']}\n")
    boxes = sw.get("boxes", [])
    if not boxes:
        out += [f"#     l = {sw['k']}\n", f"#     e = {sw['e']}\n"]
    else:
        out += [f"# Number of boxes: {len(boxes) + 1}\n", f"#     l0 = {sw['k']}\n", f"#     e0 = {sw['e']}\n"]
        for g, (gmin, gmax, bl, be) in enumerate(boxes):
            out += [f"#     g{g + 1} = [{gmin},{gmax}]\n", f"#     l{g + 1} = {bl}\n", f"#     e{g + 1} = {be}\n"]
        out.append(f"# Boxes input file: {sw['boxes_filename']}\n")
    out.append(f"# Run on {P} proc(s) in {exectime:f} secs\n")
    out.append("####################################\n\n")
    return "".join(out)


def _passes(sw, num, u, v) -> bool:
    r = (u / num) * 100
    return r >= sw["q"] and v >= sw.get("n", 1) and r <= sw.get("Q", 100) and v <= sw.get("N", 10000)


def _structs(boxes):
    from itertools import product
    return list(product(*[range(g[0], g[1] + 1) for g in boxes])) if boxes else [()]


def _motif(sw, seq: bytes, j: int, S) -> str:
    k = sw["k"]
    jj = j - k + 1
    out = seq[jj:jj + k].decode()
    off, ell = 0, k
    for g, (gmin, gmax, bl, be) in enumerate(sw.get("boxes", [])):
        off += S[g] + ell
        ell = bl
        out += "_" + seq[jj + off:jj + off + ell].decode()
    return out


def _fg_entries(sw, seqs, u, v, uv_off):
    """Foreground entries in the writers' scan order (i, j >= k-1, structure) that pass."""
    structs = _structs(sw.get("boxes", []))
    num = len(seqs)
    for i, s in enumerate(seqs):
        m = len(s)
        for j in range(sw["k"] - 1, m):
            for si, S in enumerate(structs):
                o = int(uv_off[i]) + si * m + j
                if _passes(sw, num, int(u[o]), int(v[o])):
                    yield _motif(sw, s, j, S), int(u[o]), int(v[o])


def _lines(text: str):
    """The reference's fgets(line, 512) loop: lines, with lines over 511 bytes split."""
    for raw in text.splitlines(keepends=True):
        while len(raw) > LINE_SIZE - 1:
            yield raw[:LINE_SIZE - 1]
            raw = raw[LINE_SIZE - 1:]
        yield raw


def _record(line: str):
    f = [t for t in line.split(" ") if t]
    def num(i, conv):
        try:
            return conv(f[i].strip()) if len(f) > i else 0
        except ValueError:
            return 0
    return (f[0] if f else ""), num(1, int), num(2, int), num(3, float), num(4, int)


def read_background(text: str, key_len: int):
    """(size, quorum, records, key -> record index) of a background .motex file."""
    it = _lines(text)
    size = quorum = 0
    line_num = 0
    for line in it:
        if line.startswith("\n"):
            break
        line_num += 1
        f = [t for t in line.split(" ") if t]
        if line_num == 6 and len(f) > 2:
            size = int(f[2]) if f[2].strip().lstrip("-").isdigit() else 0
        if line_num == 8 and len(f) > 3:
            rest = line.split("=", 1)[1].strip() if "=" in line else ""
            quorum = int(rest) if rest.lstrip("-").isdigit() else 0
    recs, index = [], {}
    for line in it:
        if line.startswith("\n"):
            break
        r = _record(line)
        index[r[0][:key_len]] = len(recs)
        recs.append(r)
    return size, quorum, recs, index


def _key_len(sw) -> int:
    return sw["k"] + sum(b[2] + 1 for b in sw.get("boxes", []))


def write_back(sw, seqs, u, v, uv_off, background: str):
    """(.motex text, unmatched text) of write_[structured_]motifs_back."""
    num = len(seqs)
    size, quorum, recs, index = read_background(background, _key_len(sw))
    assert size > 0 and quorum > 0
    p_un = long_double(Fraction(quorum, size))
    # both files carry the same header, naming the -o file (functions.cc:1662, 1670)
    out, un = [header(sw, num)], [header(sw, num)]
    pvalid = uvalid = 0
    for mot, uu, vv in _fg_entries(sw, seqs, u, v, uv_off):
        d = index.get(mot)
        if d is None:
            index[mot] = -1
            out.append(f"{mot} {uu} {num} {c_lf(uu / num)} {vv} 0 0 {pvalue(uu, num, p_un)}\n")
            un.append(f"{mot} {uu} {num} {c_lf(uu / num)} {vv}\n")
            uvalid += 1
        elif d != -1:
            r = recs[d]
            out.append(f"{mot} {uu} {num} {c_lf(uu / num)} {vv} {c_lf(r[3])} {r[4]} "
                       f"{pvalue(uu, num, Fraction(r[3]))}\n")
            index[mot] = -1
            pvalid += 1
    out.append(f"\n#A total of {len(recs)} background motifs were read.\n")
    out.append(f"#A total of {pvalid + uvalid} valid motifs were extracted:\n")
    out.append(f"#  {pvalid} of them were matched with background motifs\n")
    out.append(f"#  {uvalid} of them were not matched with any background motif.\n\n")
    un.append(f"\n#A total of {len(recs)} background motifs were read.\n")
    un.append(f"#A total of {pvalid + uvalid} valid motifs were extracted:\n")
    un.append(f"#  {uvalid} of them were not matched with any background motif.\n\n")
    return "".join(out), "".join(un)


def smile_id(sw, mot: str, astr: str) -> str:
    out = []
    for c in mot:
        a = astr.find(c)
        out.append("-" if (sw.get("boxes") and a < 0) else chr(ord("0") + (a if a >= 0 else len(astr))))
    return "".join(out)


def write_back_smile(sw, seqs, u, v, uv_off, background: str, astr: str):
    """(SMILE motifs after the header, unmatched SMILE motifs after the header, i.e. the
    lines the header is followed by) of write_[structuted_]motifs_back_smile."""
    num = len(seqs)
    _, _, recs, index = read_background(background, _key_len(sw))
    out, un = [], []
    pvalid = uvalid = 0
    for mot, uu, vv in _fg_entries(sw, seqs, u, v, uv_off):
        d = index.get(mot)
        if d is None:
            index[mot] = -1
            line = f"{mot} {smile_id(sw, mot, astr)} {uu}\t{vv}\n"
            un.append(line)
            out.append(line)
            uvalid += 1
        elif d != -1:
            out.append(f"{mot} {smile_id(sw, mot, astr)} {uu}\t{vv}\n")
            index[mot] = -1
            pvalid += 1
    out.append(f"Nb models: {pvalid + uvalid}\n")
    un.append(f"Nb models: {uvalid}\n")
    return "".join(out), "".join(un)


def read_motif_records(text: str):
    """The motif lines of an unmatched-motif file: after the header block, up to the next
    empty line (motex.cc:902-925, reading the header of this file)."""
    it = _lines(text)
    for line in it:
        if line.startswith("\n"):
            break
    recs = []
    for line in it:
        if line.startswith("\n"):
            break
        recs.append(_record(line))
    return recs


def write_fore(sw, records, u, v, uv_off, num_seqs: int) -> str:
    """write_motifs_fore (functions.cc:2750-2832): u, v are the rows of the re-searched
    motifs (one per record)."""
    out = [header(sw, num_seqs)]
    valid = 0
    for i, (mot, fu, fn, fr, fv) in enumerate(records):
        m = len(mot)
        if m == 0:
            continue
        o = int(uv_off[i]) + m - 1
        uu, vv = int(u[o]), int(v[o])
        if not _passes(sw, num_seqs, uu, vv):
            continue
        p = long_double(Fraction(uu, num_seqs))
        out.append(f"{mot} {fu} {fn} {c_lf(fr)} {fv} {c_lf(uu / num_seqs)} {vv} {pvalue(fu, fn, p)}\n")
        valid += 1
    out.append(f"\n#A total of {valid} valid motifs were extracted.\n\n")
    return "".join(out)
