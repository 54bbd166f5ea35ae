"""Host-side checks (no GPU): the C-ABI library and its exports, the
counter-based random stream, the restated libm, the CSV/encoding/tree-text
helpers against the reference's known answers, and the command line's
argument handling."""
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

import bbscite_amd as B
from bbscite_amd import _capi
from tests import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_known_answers.json")))
HEADER = os.path.join(ROOT, "include", "bbscite.h")
CLI = os.path.join(ROOT, "bbscite_amd", "bbscite")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \t\*]*?\b(bbs_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    """libbbscite.so loads without a GPU and exports every function that
    include/bbscite.h declares; the ctypes table covers all of them."""
    names = declared_functions()
    assert len(names) >= 36
    lib = _capi.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(_capi.EXPORTED), sorted(set(names) - set(_capi.EXPORTED))


# Random123 known-answer vectors for philox4x32-10 (counter, key -> block),
# as published with the Random123 library (Salmon et al., SC'11).
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(ctr, key, out):
    assert [int(x) for x in O.philox(ctr, key)] == out


def test_philox_against_rocrand():
    """rocRAND's host Philox4x32-10 (an independent implementation) on the
    counters the chain uses: (step, chain, job, purpose) with chain < 2^30."""
    exe = os.path.join(ROOT, "oracle", "build", "philox_rocrand")
    if not os.path.exists(exe):
        pytest.skip("oracle/build/philox_rocrand not built")
    rng = np.random.default_rng(7)
    cases = []
    for _ in range(200):
        c = [int(rng.integers(0, 2**32)), int(rng.integers(0, 2**16)),
             int(rng.integers(0, 2**32)), int(rng.integers(0, 8))]
        k = [int(rng.integers(0, 2**32)), int(rng.integers(0, 2**32))]
        cases.append((c, k))
    inp = "".join(" ".join(f"{x:x}" for x in c + k) + "\n" for c, k in cases)
    res = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True)
    lines = res.stdout.split("\n")
    for (c, k), line in zip(cases, lines):
        want = [int(t, 16) for t in line.split()]
        assert [int(x) for x in O.philox(c, k)] == want, (c, k)


def test_uniform_draws_are_unbiased_enough():
    """Move draws: v ~ U{0..N-2}, w ~ U{0..N-3} shifted past v, U[0,1)
    floats below 1; checked by chi-square over many steps."""
    n = 20
    p = B.default_params()
    probs = O.MoveProbs(0.25, 0.5, 0.75, 0.1)
    counts_v = np.zeros(n - 1)
    counts_type = np.zeros(4)
    for step in range(8000):
        m = O.draw_move(11, 0, 3, step, n, probs)
        assert 0 <= m.raw_v <= n - 2 and 0 <= m.raw_w <= n - 2 and m.raw_w != m.raw_v
        assert 0.0 <= m.raw_descendant_of_v < 1.0 and 0.0 <= m.acceptance_level < 1.0
        counts_v[m.raw_v] += 1
        counts_type[m.move_type] += 1
    e = 8000 / (n - 1)
    chi2 = float(((counts_v - e) ** 2 / e).sum())
    assert chi2 < 60.0  # 18 degrees of freedom, p ~ 1e-6
    assert np.all(np.abs(counts_type / 8000 - 0.25) < 0.03)
    del p


def test_normal_jump_moments():
    """The beta jump of ChangeBeta moves ~ Normal(0, sd/chi)."""
    probs = O.MoveProbs(1.0, 1.0, 1.0, 0.5)  # every move is a beta change
    x = np.array([O.draw_move(5, 1, 2, s, 10, probs).beta_jump for s in range(20000)])
    assert abs(x.mean()) < 0.02
    assert abs(x.std() - 0.5) < 0.02
    assert abs(((x / 0.5) ** 4).mean() - 3.0) < 0.25


def test_libm_restatement_sampled():
    """bbs_expf, bbs_logf and (float)log(1 - beta) against this host's glibc
    on every 65521st float bit pattern (every float in [0,1] is covered by
    the same program with stride 1; the device copies are checked on the GPU)."""
    exe = os.path.join(ROOT, "oracle", "build", "libm_exhaustive")
    res = subprocess.run([exe, "65521"], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    rows = dict((ln.split()[0], ln.split()[1:]) for ln in res.stdout.strip().split("\n"))
    for fn in ("expf", "logf", "log1mb"):
        n, bad = map(int, rows[fn])
        assert n > 10000 and bad == 0


# ---------------------------------------------------------------- host helpers


def test_parse_csv_reference_format():
    """src/main.cpp:52-144: digits 0..3 separated by spaces, one gene per
    line; 2 counts as mutated; the last line may lack its newline."""
    m = B.parse_csv("0 1 2\n3 0 1\n")
    assert m.tolist() == [[0, 1, 2], [3, 0, 1]]
    assert B.parse_csv("0 1 2\n3 0 1").tolist() == m.tolist()
    mut, kno = B.encode_matrix(m)
    # cell 1: gene 0 mutated (1), gene 1 known-not-mutated (0)
    assert int(mut[1][0]) == 0b01 and int(kno[1][0]) == 0b11
    # cell 0: gene 0 known 0, gene 1 unknown (3)
    assert int(mut[0][0]) == 0 and int(kno[0][0]) == 0b01
    # cell 2: 2 counts as mutated
    assert int(mut[2][0]) == 0b11 and int(kno[2][0]) == 0b11


@pytest.mark.parametrize("text,needle", [
    ("0 1\n0\n", "Inconsistent"),
    ("0 1\n0 x\n", "Illegal character"),
    ("", "Empty"),
    ("0 4\n", "Illegal character"),
    (" ".join(["0"] * 129) + "\n", "cells"),
    ("0\n" * 128, "genes"),
])
def test_parse_csv_errors(text, needle):
    with pytest.raises(_capi.BbsError) as e:
        B.parse_csv(text)
    assert needle in str(e.value)


def test_encode_matrix_wide_rows():
    """Bit g of a cell row is gene g, across both 64-bit words."""
    rng = np.random.default_rng(1)
    m = rng.choice(np.array([0, 1, 2, 3], np.int8), size=(127, 9))
    mut, kno = B.encode_matrix(m)
    for c in range(9):
        for g in range(127):
            mb = (int(mut[c][g >> 6]) >> (g & 63)) & 1
            kb = (int(kno[c][g >> 6]) >> (g & 63)) & 1
            assert kb == (m[g, c] != 3)
            assert mb == (m[g, c] in (1, 2))


def test_pruefer_and_matrix_known_answers():
    g = GOLD["pruefer_code_to_parent_vector"]
    assert B.pruefer_to_parent_vector(g["code"]).tolist() == g["expected_parent"]
    g = GOLD["parent_vector_to_matrix"]
    am, dm = B.parent_vector_to_matrix(g["parent_vector"])
    n = len(g["parent_vector"])
    for x in range(n):
        row = "".join(str((int(am[x][0]) >> b) & 1) for b in reversed(range(n)))
        assert row == g["am"][x]
        drow = "".join(str((int(dm[x][0]) >> b) & 1) for b in reversed(range(n)))
        assert drow == g["dm"][x]


def test_tree_text_known_answers():
    for key, fn in (("to_graphviz", B.tree_to_graphviz), ("to_newick", B.tree_to_newick)):
        g = GOLD[key]
        am, dm = B.parent_vector_to_matrix(g["parent_vector"])
        assert fn(am, dm) == g["expected"]


def test_generated_dataset_follows_the_recipe():
    """Truth = ancestor-or-self of the attachment node; noise flips only in
    the stated directions; entries in {0, 1, 3}; deterministic in the seed."""
    G_, C_ = 30, 120
    m0, par, att = B.generate_dataset(9, G_, C_, 0.0, 0.0, 0.0)
    am, dm = B.parent_vector_to_matrix(par)
    for c in range(C_):
        for g in range(G_):
            truth = (int(dm[att[c]][g >> 6]) >> (g & 63)) & 1
            assert m0[g, c] == truth
    m1, par1, att1 = B.generate_dataset(9, G_, C_, 0.0, 0.0, 1.0)
    assert np.all(m1 == 3) and par1.tolist() == par.tolist()
    m2, _, _ = B.generate_dataset(9, G_, C_, 0.0, 0.5, 0.0)
    ones = m0 == 1
    assert np.all(m2[~ones] == 0)  # no false positives without fd
    frac = (m2[ones] == 0).mean()
    assert 0.4 < frac < 0.6
    m3, _, _ = B.generate_dataset(9, G_, C_, 0.0, 0.5, 0.0)
    assert np.array_equal(m2, m3)


def test_initial_tree_matches_oracle():
    """bbs_initial_tree = the oracle's Pruefer-based initial state."""
    p = B.default_params(seed=21)
    for chain in range(4):
        am, dm = B.initial_tree(21, 0, chain, 40)
        t = O.Tree()
        O.lib().orc_initial_tree(O.C.byref(p), 0, chain, 40, O.C.byref(t))
        assert np.array_equal(dm, O.rows(t, "dm")) and np.array_equal(am, O.rows(t, "am"))


# ---------------------------------------------------------------- command line


def run_cli(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def test_cli_reports_missing_arguments():
    """src/Parameters.hpp:263-281: missing -i, -r and -l are each reported
    and the program quits with status 1."""
    r = run_cli()
    assert r.returncode == 1
    assert "Missing input file path" in r.stderr
    assert "Missing number of markov chains" in r.stderr
    assert "Missing markov chain length" in r.stderr
    assert "Quitting due to CLI argument errors." in r.stderr


def test_cli_rejects_bad_flags(tmp_path):
    f = tmp_path / "x.csv"
    f.write_text("0 1\n1 0\n")
    r = run_cli("-i", str(f), "-r", "1", "-l", "10", "-bogus")
    assert r.returncode == 1 and "Unknown parameter -bogus" in r.stderr
    r = run_cli("-i", str(f), "-r", "1", "-l", "10", "-e", "-1")
    assert r.returncode == 1 and "may not be negative" in r.stderr
    r = run_cli("-i", str(f), "-r", "1", "-l", "10", "-e", "0", "-move_probs", "0", "0", "0")
    assert r.returncode == 1 and "sum of the move type probabilities is zero" in r.stderr
    r = run_cli("-i", str(tmp_path / "missing.csv"), "-r", "1", "-l", "10")
    assert r.returncode == 1


def test_cli_warns_on_ignored_flags(tmp_path):
    f = tmp_path / "bad.csv"
    f.write_text("0 1\n1 x\n")
    r = run_cli("-i", str(f), "-r", "1", "-l", "10", "-n", "5", "-m", "7", "-t", "-s")
    assert "number of genes is inferred" in r.stderr
    assert "number of cells is inferred" in r.stderr
    assert r.returncode == 1 and "Illegal character" in r.stderr


def test_default_parameters_match_reference():
    """src/Parameters.hpp:39-45 defaults."""
    p = B.default_params()
    assert math.isclose(p.alpha_mean, 6.04e-5, rel_tol=1e-6)
    assert math.isclose(p.beta_mean, 0.4309, rel_tol=1e-6)
    assert math.isclose(p.beta_sd, 0.1, rel_tol=1e-6)
    assert math.isclose(p.gamma, 1.0)
    assert math.isclose(p.beta_jump_scaling_chi, 10.0)


def test_zero_chains_are_rejected_before_the_device():
    """bbs_create refuses n_chains = 0 with BBS_ERR_ARG (the reference would
    take the best of zero chains); the CLI refuses -r 0."""
    m = np.zeros((4, 3), np.int8)
    mut, kno = B.encode_matrix(m)
    with pytest.raises(B.BbsError) as e:
        B.Context(B.default_params(n_chains=0, n_steps=10), mut, kno, 4)
    assert e.value.code == _capi.ERR_ARG
    assert "chain" in str(e.value)


def test_cli_rejects_zero_chains(tmp_path):
    f = tmp_path / "x.csv"
    f.write_text("0 1\n1 0\n")
    r = run_cli("-i", str(f), "-r", "0", "-l", "10")
    assert r.returncode == 1 and "at least 1" in r.stderr
