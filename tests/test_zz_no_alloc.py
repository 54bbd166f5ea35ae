"""bbs_score_trees_noalloc (staging in a fixed device array, no device
allocation per call) against bbs_score_trees, through the C-ABI: two calls
in a row, the second smaller, give bit-identical scores and counts; a call
whose staging exceeds the array is refused with BBS_ERR_SIZE and the next
call still works."""
import sys

import numpy as np
import pytest

from bbscite_amd import _capi as B
from tests import _oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _case(G, C, seed):
    m, _, _ = B.generate_dataset(seed, G, C, 1e-3, 0.25, 0.1)
    mut, kno = B.encode_matrix(m)
    p = B.default_params(n_chains=1, n_steps=1, alpha_mean=1e-3, beta_mean=0.25)
    return p, mut, kno


@pytest.mark.parametrize("G,C", [(94, 95), (127, 128), (20, 17)])
def test_noalloc_scoring_equals_allocating_scoring(G, C):
    p, mut, kno = _case(G, C, G + C)
    big, small = 3000, 700
    dm1 = B.random_trees(11 + G, G, big)
    dm2 = B.random_trees(12 + G, G, small)
    betas = np.random.default_rng(G).uniform(0.05, 0.95, big).astype(np.float32)
    with B.Context(p, mut, kno, G) as ctx:
        say(f"[noalloc] {G}x{C}: {big} trees with betas and counts")
        ll1, c1 = ctx.score_trees(dm1, betas=betas, counts=True, no_alloc=True)
        say(f"[noalloc] {G}x{C}: {small} trees, matrix-core and vector-unit scorers")
        ll2 = ctx.score_trees(dm2, with_prior=False, no_alloc=True)
        ll3 = ctx.score_trees(dm2[: small // 2], with_prior=True, valu_lanes=True,
                              no_alloc=True)
        say(f"[noalloc] {G}x{C}: the allocating entry point on the same inputs")
        rl1, rc1 = ctx.score_trees(dm1, betas=betas, counts=True)
        rl2 = ctx.score_trees(dm2, with_prior=False)
        rl3 = ctx.score_trees(dm2[: small // 2], with_prior=True, valu_lanes=True)
    assert np.array_equal(_bits(ll1), _bits(rl1))
    assert np.array_equal(c1, rc1)
    assert np.array_equal(_bits(ll2), _bits(rl2))
    assert np.array_equal(_bits(ll3), _bits(rl3))
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    for t in range(0, small, 97):
        tree = O.tree_from_rows(np.zeros_like(dm2[t]), dm2[t], p.beta_mean)
        assert _bits(ll2[t]) == _bits(O.logscore_tree(s, tree, with_prior=False)), t


def test_noalloc_rejects_calls_beyond_its_staging():
    G, C = 94, 95
    p, mut, kno = _case(G, C, 5)
    n = B.NOALLOC_SCRATCH_BYTES // ((G + 1) * 16) + 1  # the rows alone overflow
    dm = B.random_trees(3, G, n)
    with B.Context(p, mut, kno, G) as ctx:
        say(f"[noalloc] {n} trees: refused")
        with pytest.raises(B.BbsError) as e:
            ctx.score_trees(dm, no_alloc=True)
        assert e.value.code == B.ERR_SIZE and "staging" in str(e.value)
        say("[noalloc] a call that fits afterwards")
        ok = ctx.score_trees(dm[:100], no_alloc=True)
        ref = ctx.score_trees(dm[:100])
    assert np.array_equal(_bits(ok), _bits(ref))
