"""The reference's end-to-end sanity check (test/Application.cpp:23-142),
restated: a planted 4-gene tree with three cells per node and a few
erroneous entries; 3 chains x 50,000 steps must find a tree that scores
like the planted one, and the reported best score must be that tree's
score (Catch::Approx, relative 100 float epsilons).

The oracle is checked on the CPU; on the GPU the plain chain search must
give the oracle's best bit for bit, and the command line must write the
reference's five output files for it."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import bbscite_amd as B
from tests import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_known_answers.json")))["application"]
SEEDS = [0, 1, 2]


def approx(a, b):
    """Catch::Approx default: |a - b| <= 100 eps_float * (scale + max(|a|, |b|)), scale 0."""
    return abs(a - b) <= 100 * np.finfo(np.float32).eps * max(abs(a), abs(b))


def data():
    mut = np.zeros((APP["n_cells"], 2), np.uint64)
    kno = np.zeros((APP["n_cells"], 2), np.uint64)
    mut[:, 0] = APP["is_mutated"]
    kno[:, 0] = APP["is_known"]
    return mut, kno


def params(seed):
    return B.default_params(alpha_mean=APP["alpha"], beta_mean=APP["beta"], beta_sd=APP["beta_sd"],
                            n_chains=APP["n_chains"], n_steps=APP["chain_length"], seed=seed)


def planted_score(s):
    am, dm = B.parent_vector_to_matrix(APP["planted_parent_vector"])
    t = O.tree_from_rows(am, dm, APP["beta"])
    return O.logscore_tree(s, t)


@pytest.mark.parametrize("seed", SEEDS)
def test_oracle_finds_the_planted_tree(seed):
    mut, kno = data()
    p = params(seed)
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, APP["n_genes"])
    best, best_score, _, _ = O.run(p, s)
    found = O.logscore_tree(s, best)
    assert approx(found, planted_score(s))
    assert approx(found, best_score)


def write_csv(path, mut, kno, n_genes):
    rows = []
    for g in range(n_genes):
        row = []
        for c in range(mut.shape[0]):
            k = (int(kno[c][0]) >> g) & 1
            m = (int(mut[c][0]) >> g) & 1
            row.append("3" if not k else ("1" if m else "0"))
        rows.append(" ".join(row))
    path.write_text("\n".join(rows) + "\n")


@pytest.mark.gpu
@pytest.mark.parametrize("seed", SEEDS)
def test_gpu_finds_the_oracle_best(seed):
    mut, kno = data()
    p = params(seed)
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, APP["n_genes"])
    o_best, o_score, o_chain, o_step = O.run(p, s)
    with B.Context(p, mut, kno, APP["n_genes"]) as ctx:
        ctx.run()
        am, dm, beta, score = ctx.best()
    assert np.float32(score).tobytes() == np.float32(o_score).tobytes()
    assert np.array_equal(dm, O.rows(o_best, "dm")) and np.array_equal(am, O.rows(o_best, "am"))
    assert approx(score, planted_score(s))


@pytest.mark.gpu
def test_cli_end_to_end(tmp_path):
    """bbscite -i data.csv -r 3 -l 50000 -seed 1: the five output files
    (src/main.cpp:165-214) hold the best tree of the same search."""
    mut, kno = data()
    csv = tmp_path / "planted.csv"
    write_csv(csv, mut, kno, APP["n_genes"])
    exe = os.path.join(ROOT, "bbscite_amd", "bbscite")
    r = subprocess.run([exe, "-i", str(csv), "-r", "3", "-l", "50000", "-fd", str(APP["alpha"]),
                        "-ad", str(APP["beta"]), "-sd", str(APP["beta_sd"]), "-seed", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Time elapsed:" in r.stdout
    base = str(tmp_path / "planted")
    with B.Context(params(1), mut, kno, APP["n_genes"]) as ctx:
        ctx.run()
        am, dm, beta, score = ctx.best()
    n = APP["n_genes"] + 1
    am_txt = open(base + "_ml0_am.txt").read().split("\n")
    dm_txt = open(base + "_ml0_dm.txt").read().split("\n")
    for i in range(n):
        assert am_txt[i] == "".join("1" if (int(am[i][0]) >> j) & 1 else "0" for j in range(n))
        assert dm_txt[i] == "".join("1" if (int(dm[i][0]) >> j) & 1 else "0" for j in range(n))
    assert open(base + "_ml0.newick").read() == B.tree_to_newick(am, dm)
    assert open(base + "_ml0.gv").read() == B.tree_to_graphviz(am, dm)
    assert math.isclose(float(open(base + "_ml0_beta.txt").read()), beta, rel_tol=1e-5)
