"""Pins the oracle's TreeScorer restatement (oracle/bbs_oracle.c).

* the reference's closed-form test (test/TreeScorer.cpp:33-92), as exact
  counts and as the log-likelihood;
* brute force: counts per (cell, attachment) computed by walking parent
  pointers, the textbook SCITE likelihood, on random data and on every
  labelled tree with 5 nodes;
* the beta prior against the formula in double precision.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

from tests import _oracle as O

GOLD = json.load(
    open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json"))
)


def _rows_from_masks(masks):
    a = np.zeros((len(masks), 2), np.uint64)
    for i, m in enumerate(masks):
        a[i, 0] = m & ((1 << 64) - 1)
        a[i, 1] = m >> 64
    return a


def test_closed_form_reference_case():
    g = GOLD["logscore_tree"]
    mut = _rows_from_masks(g["is_mutated"])
    kno = _rows_from_masks(g["is_known"])
    s = O.make_scorer(g["alpha"], g["beta"], g["prior_sd"], mut, kno, g["n_genes"])
    t = O.tree_from_parent(g["parent_vector"], g["beta"])
    score, counts, node = O.logscore_tree(s, t, with_prior=False, counts=True)
    tot = counts.astype(int).sum(axis=0)
    e = g["expected_counts_total"]
    assert tot.tolist() == [e["n00"], e["n01"], e["n10"], e["n11"]]
    a, b = np.float32(g["alpha"]), np.float32(g["beta"])
    true = 13 * math.log(1 - a) + 5 * math.log(1 - b) + math.log(b) + math.log(a)
    assert score == pytest.approx(true, rel=1e-6)
    with_prior = O.logscore_tree(s, t, with_prior=True)
    beta_score = O.lib().orc_logscore_beta(s, np.float32(g["beta"]))
    assert with_prior - beta_score == pytest.approx(true, rel=1e-6)


def _brute(parent, mut_masks, known_masks, n_genes, alpha, beta):
    """Textbook SCITE likelihood by walking parents; returns (T, counts)."""
    n = len(parent)

    def ancestors(x):
        out = {x}
        while x != n - 1:
            x = parent[x]
            out.add(x)
        return out

    f = np.float32
    L00 = f(math.log(1.0 - float(f(alpha))))
    L10 = f(math.log(f(alpha)))  # logf
    L01 = f(math.log(f(beta)))
    L11 = f(math.log(1.0 - float(f(beta))))
    T = f(0.0)
    all_counts = []
    for c in range(len(mut_masks)):
        best = None
        best_counts = None
        for a in range(n):
            anc = ancestors(a)
            cnt = [0, 0, 0, 0]
            for g in range(n_genes):
                if not (known_masks[c] >> g) & 1:
                    continue
                obs = (mut_masks[c] >> g) & 1
                true = 1 if g in anc else 0
                cnt[obs * 2 + true] += 1
            s = f(0.0)
            for k, L in zip(range(4), (L00, L01, L10, L11)):
                s = f(s + f(f(cnt[k]) * L))
            if best is None or best < s:
                best, best_counts = s, cnt
        T = f(T + best)
        all_counts.append(best_counts)
    return T, np.array(all_counts)


def _random_masks(rng, n_cells, n_genes, p1=0.3, p3=0.1):
    mut, kno = [], []
    for _ in range(n_cells):
        m = k = 0
        for g in range(n_genes):
            u = rng.random()
            if u < p3:
                continue
            k |= 1 << g
            if u < p3 + p1:
                m |= 1 << g
        mut.append(m)
        kno.append(k)
    return mut, kno


@pytest.mark.parametrize("n_genes,n_cells,seed", [(4, 6, 1), (20, 17, 2), (63, 64, 3), (94, 95, 4), (127, 128, 5)])
def test_scorer_against_brute_force(n_genes, n_cells, seed):
    rng = np.random.default_rng(seed)
    import random

    r = random.Random(seed)
    mut, kno = _random_masks(r, n_cells, n_genes)
    n = n_genes + 1
    alpha, beta = 6.04e-5, 0.4309
    s = O.make_scorer(alpha, beta, 0.1, _rows_from_masks(mut), _rows_from_masks(kno), n_genes)
    n_trees = 3 if n_genes > 60 else 10
    for _ in range(n_trees):
        code = rng.integers(0, n, n - 2)
        par = np.zeros(n, np.uint32)
        O.lib().orc_pruefer_to_parent(O._ptr(code.astype(np.uint32)), n - 2, O._ptr(par))
        t = O.tree_from_parent(par, beta)
        T, counts, _ = O.logscore_tree(s, t, with_prior=False, counts=True)
        bT, bcounts = _brute(par.tolist(), mut, kno, n_genes, alpha, beta)
        assert np.array_equal(counts.astype(int), bcounts)
        assert np.float32(T).tobytes() == bT.tobytes()


def test_scorer_exhaustive_5_nodes():
    """Every labelled tree on 5 nodes (all 125 Pruefer codes)."""
    import random

    r = random.Random(11)
    mut, kno = _random_masks(r, 9, 4, p1=0.45, p3=0.15)
    s = O.make_scorer(0.05, 0.3, 0.1, _rows_from_masks(mut), _rows_from_masks(kno), 4)
    seen = set()
    for code in itertools.product(range(5), repeat=3):
        par = np.zeros(5, np.uint32)
        O.lib().orc_pruefer_to_parent(O._ptr(np.array(code, np.uint32)), 3, O._ptr(par))
        seen.add(tuple(par.tolist()))
        t = O.tree_from_parent(par, 0.3)
        T, counts, _ = O.logscore_tree(s, t, with_prior=False, counts=True)
        bT, bcounts = _brute(par.tolist(), mut, kno, 4, 0.05, 0.3)
        assert np.array_equal(counts.astype(int), bcounts)
        assert np.float32(T).tobytes() == bT.tobytes()
    assert len(seen) == 125  # Cayley: 5^3 distinct labelled trees


def test_beta_prior_formula():
    """logscore_beta against the Beta(a, b) log-density in double."""
    for mean, sd in ((0.4309, 0.1), (0.2, 0.1), (0.25, 0.1)):
        s = O.make_scorer(6.04e-5, mean, sd, np.zeros((1, 2), np.uint64), np.zeros((1, 2), np.uint64), 2)
        f = np.float32
        # (float)(1 - mean) promoted to double, pow in double (src/TreeScorer.hpp:108-111)
        a = f(float(f(1) - f(mean)) * float(f(mean)) ** 2 / float(f(sd)) ** 2 - float(f(mean)))
        b = f(a * (f(1) / f(mean) - f(1)))
        assert s.bpriora == a and s.bpriorb == b
        assert s.prior_fallback == 0
        for beta in (0.05, 0.2, 0.4309, 0.7):
            exact = (
                math.lgamma(a + b)
                + (a - 1) * math.log(beta)
                + (b - 1) * math.log(1 - beta)
                - math.lgamma(a)
                - math.lgamma(b)
            )
            got = O.lib().orc_logscore_beta(s, beta)
            assert got == pytest.approx(exact, abs=2e-4 * max(1.0, abs(exact)))


def test_beta_prior_overflow_fallback():
    """Mean 0.3, sd 0.05 (the 94x95 chain configuration): tgammaf(a+b)
    overflows, the reference's prior is NaN; the oracle uses log-gamma."""
    s = O.make_scorer(1e-4, 0.3, 0.05, np.zeros((1, 2), np.uint64), np.zeros((1, 2), np.uint64), 2)
    assert s.prior_fallback == 1
    a, b = float(s.bpriora), float(s.bpriorb)
    assert math.isinf(float(np.float32(math.gamma(min(a + b, 170)))))
    v = O.lib().orc_logscore_beta(s, 0.3)
    exact = (math.lgamma(a + b) + (a - 1) * math.log(0.3) + (b - 1) * math.log(0.7)
             - math.lgamma(a) - math.lgamma(b))
    assert math.isfinite(v)
    assert v == pytest.approx(exact, abs=1e-3)
