"""The measured workloads, as synthetic inputs (one definition shared by
bench.py and the tests).

Data generation follows the recipe of verification/lib.py:92-98,199-230
with the library's own Philox generator (bbs_generate_dataset): a uniform
Pruefer tree over genes + root, uniform cell attachment, truth =
ancestor-or-self, then 0->1 with probability fd, 1->0 with probability ad,
then missing (3) with the given probability.

  plain_64x63   63 genes x 64 cells, seed 1, -ad 0.2, -fd 1e-4, 10 % missing;
                8 chains x 100,000 steps, fixed error rates (-e 0)
  plain_94x95   94 genes x 95 cells, seed 2, -ad 0.3, -fd 1e-4, 15 % missing;
                1,024 chains, beta moves -e 0.1, beta prior mean 0.3 sd 0.05
  scoring       one 94 x 95 matrix of iid entries (P(1) = 0.3, P(3) = 0.1,
                else 0; numpy generator seed 3) and the uniform random trees
                of seed 3
  batch_mix     256 datasets, dataset i from seed 4 + i with genes ~ U[16,94]
                and cells ~ U[16,95] (sizes from a numpy generator, seed 4),
                noise as plain_64x63, 64 chains each
  served_stream Poisson arrivals at 50 jobs/s (numpy generator, seed 5);
                genes U[16,127], cells U[16,128], chains U[8,256], steps
                log-uniform in [1e4, 1e6]; job i's data from seed 1000 + i,
                noise as plain_64x63
"""
from __future__ import annotations

import numpy as np

from . import _capi as B


def plain_64x63(n_chains=8, n_steps=100_000):
    """(params, is_mutated, is_known, n_genes) of the 63 x 64 workload."""
    G, C = 63, 64
    m, _, _ = B.generate_dataset(1, G, C, 1e-4, 0.2, 0.1)
    mut, kno = B.encode_matrix(m)
    p = B.default_params(n_chains=n_chains, n_steps=n_steps, seed=1, alpha_mean=1e-4,
                         beta_mean=0.2, prob_beta_change=0.0)
    return p, mut, kno, G


def plain_94x95(n_chains=1024, n_steps=1_000_000):
    """(params, is_mutated, is_known, n_genes) of the 94 x 95 workload."""
    G, C = 94, 95
    m, _, _ = B.generate_dataset(2, G, C, 1e-4, 0.3, 0.15)
    mut, kno = B.encode_matrix(m)
    p = B.default_params(n_chains=n_chains, n_steps=n_steps, seed=2, alpha_mean=1e-4,
                         beta_mean=0.3, beta_sd=0.05, prob_beta_change=0.1)
    return p, mut, kno, G


SCORING_GENES, SCORING_CELLS, SCORING_TREE_SEED = 94, 95, 3


def scoring_matrix():
    """(is_mutated, is_known) of the bulk-scoring matrix."""
    rng = np.random.default_rng(3)
    u = rng.random((SCORING_GENES, SCORING_CELLS))
    m = np.zeros((SCORING_GENES, SCORING_CELLS), np.int8)
    m[u < 0.3] = 1
    m[(u >= 0.3) & (u < 0.4)] = 3
    return B.encode_matrix(m)


def batch_mix(steps, chains=64, n_datasets=256):
    """[(params, is_mutated, is_known, n_genes)] of the 256-dataset mix (the
    first n_datasets of it)."""
    rng = np.random.default_rng(4)
    jobs = []
    for i in range(n_datasets):
        G = int(rng.integers(16, 95))
        Cc = int(rng.integers(16, 96))
        m, _, _ = B.generate_dataset(4 + i, G, Cc, 1e-4, 0.2, 0.1)
        mut, kno = B.encode_matrix(m)
        p = B.default_params(n_chains=chains, n_steps=steps, seed=4 + i,
                             alpha_mean=1e-4, beta_mean=0.2)
        jobs.append((p, mut, kno, G))
    return jobs


def served_stream(n_jobs, seed=5, max_steps=None):
    """[(arrival_s, (params, is_mutated, is_known, n_genes))] of the first
    n_jobs jobs of the served stream; max_steps caps each job's steps (for
    parity checks that replay every chain on the CPU)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n_jobs):
        t += rng.exponential(1.0 / 50.0)
        G = int(rng.integers(16, 128))
        Cc = int(rng.integers(16, 129))
        R = int(rng.integers(8, 257))
        S = int(10 ** rng.uniform(4.0, 6.0))
        if max_steps is not None:
            S = min(S, max_steps)
        m, _, _ = B.generate_dataset(1000 + i, G, Cc, 1e-4, 0.2, 0.1)
        mut, kno = B.encode_matrix(m)
        p = B.default_params(n_chains=R, n_steps=S, seed=1000 + i, alpha_mean=1e-4,
                             beta_mean=0.2)
        jobs.append((t, (p, mut, kno, G)))
    return jobs
