"""The quality harness's host pieces (verification/__main__.py:54-163
restated): Newick parsing, best-attachment scoring in double precision, and
the two one-sided t-tests, checked against independent computations (a
brute-force textbook likelihood, scipy's one-sample t-test)."""
import json
import math
import os

import numpy as np
import pytest
from scipy import stats

import bbscite_amd as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_known_answers.json")))


def test_newick_known_answer_and_round_trip():
    g = GOLD["to_newick"]
    assert B.newick_to_parent_vector(g["expected"]).tolist() == g["parent_vector"]
    rng = np.random.default_rng(3)
    for G in (2, 5, 30, 127):
        code = rng.integers(0, G + 1, G - 1)
        par = B.pruefer_to_parent_vector(code)
        am, dm = B.parent_vector_to_matrix(par)
        text = B.tree_to_newick(am, dm)
        assert B.newick_to_parent_vector(text).tolist() == par.tolist()


def test_newick_relabels_and_rejects():
    # labels 10, 20, 30 -> 0, 1, 2 in sorted order (verification/lib.py:148-164)
    assert B.newick_to_parent_vector(" ( 10 , 20 ) 30 \n").tolist() == [2, 2, 2]
    for bad in ("(0,1", "(0,1)2)", "(0,2)1", "(0,0)1", "", "(a)1"):
        with pytest.raises(B.BbsError):
            B.newick_to_parent_vector(bad)


def _brute_double(parent, mut, kno, n_genes, alpha, beta):
    n = len(parent)

    def anc(x):
        out = {x}
        while x != n - 1:
            x = parent[x]
            out.add(x)
        return out

    L = {(0, 0): math.log(1 - alpha), (1, 0): math.log(alpha),
         (0, 1): math.log(beta), (1, 1): math.log(1 - beta)}
    total, att = 0.0, []
    for c in range(len(mut)):
        best, arg = None, None
        for v in range(n):
            a = anc(v)
            ll = 0.0
            cnt = {(0, 0): 0, (1, 0): 0, (0, 1): 0, (1, 1): 0}
            for g in range(n_genes):
                if (kno[c] >> g) & 1:
                    cnt[((mut[c] >> g) & 1, 1 if g in a else 0)] += 1
            ll = sum(cnt[k] * L[k] for k in ((0, 0), (1, 0), (0, 1), (1, 1)))
            if best is None or ll > best:
                best, arg = ll, v
        total += best
        att.append(arg)
    return total, att


@pytest.mark.parametrize("G,C,seed", [(4, 6, 1), (20, 30, 2), (70, 90, 3)])
def test_score_attachments_against_brute_force(G, C, seed):
    rng = np.random.default_rng(seed)
    m = rng.choice(np.array([0, 1, 2, 3], np.int8), p=[0.6, 0.2, 0.1, 0.1], size=(G, C))
    mut, kno = B.encode_matrix(m)
    par = B.pruefer_to_parent_vector(rng.integers(0, G + 1, G - 1))
    am, dm = B.parent_vector_to_matrix(par)
    ll, att = B.score_attachments(dm, mut, kno, 1e-3, 0.3)
    mi = [int(mut[c][0]) | (int(mut[c][1]) << 64) for c in range(C)]
    ki = [int(kno[c][0]) | (int(kno[c][1]) << 64) for c in range(C)]
    ref, ref_att = _brute_double(par.tolist(), mi, ki, G, 1e-3, 0.3)
    assert ll == pytest.approx(ref, rel=1e-12, abs=1e-9)
    assert att.tolist() == ref_att


@pytest.mark.parametrize("seed", range(6))
def test_tost_matches_scipy(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 80))
    d = rng.normal(rng.uniform(-2, 2), rng.uniform(0.1, 3), n)
    delta = float(rng.uniform(0.2, 3))
    mean, pl, pu = B.tost(d, delta)
    want_l = stats.ttest_1samp(d, -delta, alternative="greater").pvalue
    want_u = stats.ttest_1samp(d, delta, alternative="less").pvalue
    assert mean == pytest.approx(d.mean(), rel=1e-12)
    assert pl == pytest.approx(want_l, rel=1e-9, abs=1e-300)
    assert pu == pytest.approx(want_u, rel=1e-9, abs=1e-300)


def test_tost_degenerate_cases():
    assert B.tost([0.0] * 8, 1.0) == (0.0, 0.0, 0.0)
    assert B.tost([2.0] * 8, 1.0) == (2.0, 0.0, 1.0)
    with pytest.raises(B.BbsError):
        B.tost([1.0], 1.0)


QUALITY = os.path.join(ROOT, "bbscite_amd", "bbscite_quality")


def _write_case(tmp_path):
    """The reference's TreeScorer known answer as CSV + Newick files."""
    g = GOLD["logscore_tree"]
    rows = []
    for gene in range(g["n_genes"]):
        row = []
        for c in range(g["n_cells"]):
            if not (g["is_known"][c] >> gene) & 1:
                row.append("3")
            else:
                row.append("1" if (g["is_mutated"][c] >> gene) & 1 else "0")
        rows.append(" ".join(row))
    csv = tmp_path / "input.csv"
    csv.write_text("\n".join(rows) + "\n")
    am, dm = B.parent_vector_to_matrix(g["parent_vector"])
    tree = tmp_path / "tree.newick"
    tree.write_text(B.tree_to_newick(am, dm))
    return g, csv, tree


def test_quality_score_command_known_answer(tmp_path):
    """test/TreeScorer.cpp:90-92: 13 log(1-a) + 5 log(1-b) + log b + log a
    (best attachments), here in double precision."""
    import subprocess

    g, csv, tree = _write_case(tmp_path)
    a, b = g["alpha"], g["beta"]
    r = subprocess.run([QUALITY, "score", "-i", str(csv), "-t", str(tree), "-fd", str(a),
                        "-ad", str(b), "-compact"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    want = 13 * math.log(1 - a) + 5 * math.log(1 - b) + math.log(b) + math.log(a)
    assert float(r.stdout) == pytest.approx(want, rel=1e-14)
    r = subprocess.run([QUALITY, "score", "-i", str(csv), "-t", str(tree)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0 and "Attachments:" in r.stdout
    r = subprocess.run([QUALITY, "frobnicate"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_quality_harness_all_forms_equivalent():
    """The full harness on a small configuration: every form (full and
    incremental scoring, batched, served) is equivalent to the plain launch
    within one bit flip; here they are identical input by input."""
    import subprocess

    print("[quality] harness", flush=True)
    r = subprocess.run([QUALITY, "test", "-inputs", "6", "-genes", "20", "-cells", "24",
                        "-chains", "4", "-steps", "4000", "-seed", "9"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Every form is equivalent to the plain launch." in r.stdout
    for form in ("full vs incremental scoring", "batched vs plain", "served vs plain"):
        line = [ln for ln in r.stdout.splitlines() if ln.startswith(form)][0]
        assert "mean +0.0000" in line and "equivalent" in line
