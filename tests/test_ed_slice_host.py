"""Host check of the text-parallel edit recurrence (motex_amd/csrc/ed_slice.h, the step the
edit kernel runs for 32 texts at once) against the oracle's scalar per-pair edit operator
(orc_ed_pair, functions.cc:734-768 with MaxShift functions.cc:437-447).  The check program
tests/slice_check.cpp is built with g++ and run on random and low-complexity inputs; it also
asserts that no selected word needs more than 2k bits and that the truncated-column form
(slice_cell_masked + slice_pm, used for columns j > k) gives the same cells."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory, oracle):
    out = str(tmp_path_factory.mktemp("slice") / "slice_check")
    libdir = os.path.join(ROOT, "oracle", "_build")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "motex_amd", "csrc"),
                    os.path.join(ROOT, "tests", "slice_check.cpp"), "-L", libdir, "-loracle",
                    f"-Wl,-rpath,{libdir}", "-o", out], check=True)
    return out


# seed k e m nmin nmax sigma ntexts
CASES = [
    (1, 8, 1, 60, 20, 80, 20, 32),     # protein-like alphabet, the protein workload's k and e
    (2, 5, 2, 40, 5, 60, 4, 32),       # texts shorter than k mixed in
    (3, 3, 0, 30, 1, 40, 2, 17),       # e = 0, partial group
    (4, 12, 3, 50, 10, 90, 4, 32),
    (5, 14, 5, 40, 14, 60, 2, 32),     # largest k of the 32-bit word path
    (6, 1, 0, 20, 1, 30, 3, 8),        # k = 1
    (7, 10, 2, 100, 50, 200, 20, 32),
    (8, 14, 13, 30, 14, 80, 1, 32),    # one symbol: widest words, most ties
    (9, 14, 7, 60, 14, 120, 2, 32),
    (10, 8, 7, 40, 8, 100, 1, 32),
    (11, 2, 1, 30, 2, 50, 1, 32),
    (12, 14, 0, 40, 14, 70, 1, 32),
    (13, 4, 3, 40, 4, 90, 1, 32),
    (14, 6, 2, 40, 6, 90, 2, 32),
    (15, 15, 4, 60, 15, 120, 4, 32),
    (16, 20, 6, 80, 20, 160, 4, 32),
    (17, 20, 19, 40, 20, 90, 1, 32),   # one symbol, e = k - 1
    (18, 30, 8, 90, 30, 200, 4, 32),   # largest k of the text-parallel path
    (19, 30, 29, 60, 30, 120, 1, 32),  # one symbol: widest words, popcount up to k + 1
    (20, 30, 12, 70, 30, 140, 20, 24),
]


@pytest.mark.parametrize("case", CASES, ids=[f"k{c[1]}e{c[2]}s{c[6]}" for c in CASES])
def test_sliced_step_matches_oracle(checker, case):
    r = subprocess.run([checker, *map(str, case)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatched_columns=0" in r.stdout and "wide_state=0" in r.stdout
    assert "masked_form_differs=0" in r.stdout


def test_logic_tables_match_the_host_expressions(checker):
    """The device build selects, masks and compares with v_bitop3_b32 and the kBop3* truth
    tables of ed_slice.h; evaluated as the instruction does, each table must equal the plain
    expression the host build (and so the check above) uses, and the comparison tables must
    order 6-bit numbers exactly."""
    r = subprocess.run([checker, "tables"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "tables: 0 mismatches" in r.stdout
