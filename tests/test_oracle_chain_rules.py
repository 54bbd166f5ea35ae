"""Independent checks of the oracle's chain loop (src/Application.hpp:425-487).

The oracle's per-step mechanics (moves, realisation, scores) are pinned
elsewhere. These tests pin the rules that combine them, each against a
recomputation that does not use the oracle's loop:
  - best tracking: the first strictly greater proposal in chain-major order,
    chain 0's step 0 seeding it (src/Application.hpp:456), on a fixture
    where many proposals score exactly the same;
  - the Metropolis test with the subtree-swap Hastings ratio nd(v)/nd(w)
    when w is an ancestor of v (src/Application.hpp:443-455), recomputed
    from parent vectors;
  - the draw of the node pair (src/RawMoveDistribution.hpp:146-155).
"""
import ctypes as C

import numpy as np

from bbscite_amd import _capi as B
from tests import _oracle as O
from tests.test_oracle_tree import _brute_realize, _manual_move

_libm = C.CDLL("libm.so.6")
_libm.expf.restype = C.c_float
_libm.expf.argtypes = [C.c_float]


def _f32(x):
    return np.float32(x)


def _tie_fixture():
    """5 genes in two groups of identical data rows (genes 0-2 and 3-4), 6
    cells, no missing entries: relabelling genes inside a group leaves every
    score unchanged, so the best score is reached by many trees."""
    m = np.array([
        [1, 1, 0, 0, 1, 0],
        [1, 1, 0, 0, 1, 0],
        [1, 1, 0, 0, 1, 0],
        [0, 1, 1, 0, 0, 1],
        [0, 1, 1, 0, 0, 1],
    ], np.int8)
    mut, kno = B.encode_matrix(m)
    p = B.default_params(n_chains=6, n_steps=400, seed=17, alpha_mean=0.05, beta_mean=0.2,
                         prob_beta_change=0.0)
    return p, mut, kno, 5


def test_best_is_first_strict_maximum_in_chain_major_order():
    p, mut, kno, G = _tie_fixture()
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    traces, results = [], []
    for c in range(p.n_chains):
        res, tr = O.run_chain(p, s, c, c == 0, trace=True)
        traces.append(tr)
        results.append(res)
    # the reference's scan over the whole run (P = 1: chain-major)
    best, where = None, None
    for c, tr in enumerate(traces):
        for step, sc in enumerate(tr["proposed_score"]):
            if (c == 0 and step == 0) or sc > best:
                best, where = sc, (c, step)
    # non-vacuous: the maximum is reached again later, in the same chain and
    # in later chains, so a ">=" or "last wins" rule would move the record
    hits = [(c, st) for c, tr in enumerate(traces)
            for st in np.nonzero(tr["proposed_score"] == best)[0]]
    assert len(hits) >= 10 and len({c for c, _ in hits}) >= 3
    assert hits[0] == where
    tree, score, chain, step = O.run(p, s)
    assert (chain, step) == where and _f32(score) == _f32(best)
    assert _f32(O.logscore_tree(s, tree)) == _f32(best)
    # the per-chain records and the trace flags follow the same scan, each
    # chain starting from no record (chain 0's step 0 counts unconditionally)
    for c, (tr, res) in enumerate(zip(traces, results)):
        run_best, first = -np.inf, None
        for st, sc in enumerate(tr["proposed_score"]):
            new = (c == 0 and st == 0) or sc > run_best
            assert bool(tr["flags"][st] & 2) == new, (c, st)
            if new:
                run_best, first = sc, st
        assert res.best_step == first and _f32(res.best_score) == _f32(run_best)
    assert O.chain_major_best(results) == where[0]


def _ancestor_or_self(parent, x, y):
    n = len(parent)
    while True:
        if x == y:
            return True
        if y == n - 1:
            return False
        y = parent[y]


def _n_desc(parent, x):
    return sum(_ancestor_or_self(parent, x, y) for y in range(len(parent)))


def test_subtree_swap_acceptance_recomputed_from_parent_vectors():
    """Subtree swaps only: every proposal's score equals the score of the
    tree moved by hand on its parent vector, and every acceptance decision
    equals corr * expf((s' - s) * gamma) > u with corr = nd(v)/nd(w) when w
    is an ancestor of v in the current tree, else 1."""
    G, Cn = 12, 8
    m, _, _ = B.generate_dataset(21, G, Cn, 0.02, 0.3, 0.1)
    mut, kno = B.encode_matrix(m)
    p = B.default_params(n_chains=3, n_steps=400, seed=23, alpha_mean=0.02, beta_mean=0.3,
                         gamma=0.5, prob_beta_change=0.0, prob_prune_n_reattach=0.0,
                         prob_swap_nodes=0.0, prob_swap_subtrees=1.0)
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    probs = O.move_probs(p)
    n = G + 1
    related = sensitive = accepted = 0
    for c in range(p.n_chains):
        res, tr = O.run_chain(p, s, c, c == 0, trace=True)
        t0 = O.Tree()
        O.lib().orc_initial_tree(p, 0, c, G, t0)
        parent = O.parents(t0)
        parent[n - 1] = n - 1
        tree = O.tree_from_parent(parent, p.beta_mean)
        cur = _f32(O.logscore_tree(s, tree))
        for step in range(p.n_steps):
            raw = O.draw_move(p.seed, 0, c, step, n, probs)
            assert raw.move_type == 3
            mod = _brute_realize(parent, tree, raw)
            assert (tr["v"][step], tr["w"][step]) == (mod["v"], mod["w"])
            v, w = mod["v"], mod["w"]
            assert not _ancestor_or_self(parent, v, w)  # the realisation's swap
            m_obj = O.Mod(3, v, w, mod["parent_of_v"], mod["parent_of_w"],
                          mod["descendant_of_v"], mod["nondescendant_of_v"], p.beta_mean)
            nxt = _manual_move(parent, m_obj, tree)
            nxt_tree = O.tree_from_parent(nxt, p.beta_mean)
            prop = _f32(O.logscore_tree(s, nxt_tree))
            assert prop.tobytes() == _f32(tr["proposed_score"][step]).tobytes(), (c, step)
            e = _f32(_libm.expf(float(_f32((prop - cur) * _f32(p.gamma)))))
            if _ancestor_or_self(parent, w, v):
                related += 1
                corr = _f32(_f32(_n_desc(parent, v)) / _f32(_n_desc(parent, w)))
                inv = _f32(_f32(_n_desc(parent, w)) / _f32(_n_desc(parent, v)))
                if (_f32(corr * e) > raw.acceptance_level) != (_f32(inv * e) > raw.acceptance_level):
                    sensitive += 1
            else:
                corr = _f32(1.0)
            acc = bool(_f32(corr * e) > _f32(raw.acceptance_level))
            assert acc == bool(tr["flags"][step] & 1), (c, step)
            if acc:
                accepted += 1
                parent, tree, cur = nxt, nxt_tree, prop
        assert cur.tobytes() == _f32(res.final_score).tobytes()
        assert np.array_equal(O.rows(res.final_tree, "dm"), O.rows(tree, "dm"))
    # the ratio decided some steps: its inverse would have flipped them
    assert related > 100 and sensitive >= 5 and accepted > 50


def test_node_pair_draw_is_uniform_over_nonroot_pairs():
    """v ~ U{0..N-2}, then w ~ U{0..N-3} shifted past v: (v, w) is uniform
    over ordered pairs of distinct non-root nodes; the root N-1 is never
    drawn (src/RawMoveDistribution.hpp:146-155)."""
    n = 12
    probs = O.MoveProbs(0.25, 0.5, 0.75, 0.1)
    counts = np.zeros((n - 1, n - 1))
    draws = 33000
    for step in range(draws):
        mv = O.draw_move(29, 1, 2, step, n, probs)
        assert 0 <= mv.raw_v <= n - 2 and 0 <= mv.raw_w <= n - 2 and mv.raw_w != mv.raw_v
        counts[mv.raw_v, mv.raw_w] += 1
    cells = (n - 1) * (n - 2)
    e = draws / cells
    off = ~np.eye(n - 1, dtype=bool)
    chi2 = float(((counts[off] - e) ** 2 / e).sum())
    assert chi2 < 200.0  # 109 degrees of freedom, p ~ 1e-7
    # w given v: uniform over the other n - 2 non-root nodes for every v
    for v in range(n - 1):
        row = counts[v][off[v]]
        ev = row.sum() / (n - 2)
        assert float(((row - ev) ** 2 / ev).sum()) < 40.0  # 9 dof, p ~ 4e-6


def test_recorded_proposals_are_each_steps_move():
    """The oracle's per-step record of proposed trees, which the device's
    record is compared with row for row: each step's rows are the tree that
    the step's move makes of the current tree (the parent-vector
    restatement of test/MutationTree.cpp:697-735), a beta move proposes the
    current tree, and the current tree follows the accepted steps."""
    m, _, _ = B.generate_dataset(21, 24, 30, 0.05, 0.2, 0.1)
    mut, kno = B.encode_matrix(m)
    G = m.shape[0]
    n = G + 1
    p = B.default_params(n_chains=1, n_steps=600, seed=21, alpha_mean=0.05, beta_mean=0.2,
                         prob_beta_change=0.2)
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    res, tr, am, dm = O.run_chain_proposals(p, s, 0, True, n)
    t0 = O.Tree()
    O.lib().orc_initial_tree(C.byref(p), 0, 0, G, C.byref(t0))
    parent = O.parents(t0)
    accepted = [0, 0, 0, 0]
    for step in range(p.n_steps):
        rec = tr[step]
        if rec["move_type"] == 0:
            expected = parent
        else:
            mod = O.Mod()
            for k in ("move_type", "v", "w", "parent_of_v", "parent_of_w",
                      "descendant_of_v", "nondescendant_of_v"):
                setattr(mod, k, int(rec[k]))
            expected = _manual_move(parent, mod, O.tree_from_parent(parent))
        ref = O.tree_from_parent(expected)
        assert np.array_equal(am[step], O.rows(ref, "am")), step
        assert np.array_equal(dm[step], O.rows(ref, "dm")), step
        if rec["flags"] & 1:
            parent = expected
            accepted[int(rec["move_type"])] += 1
    final = O.tree_from_parent(parent)
    assert np.array_equal(O.rows(res.final_tree, "dm"), O.rows(final, "dm"))
    # non-vacuous: accepted moves of every kind change the current tree
    assert min(accepted) >= 5, accepted
