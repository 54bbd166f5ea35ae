"""Oracle parity of the measured workloads themselves, through the C-ABI.

Every chain that is compared is replayed on the CPU oracle from the same
seeds: best score and step, final score, beta and tree (bit for bit), and
each dataset's or job's global best in chain-major order
(src/Application.hpp:456, P = 1).
  - 63 x 64 (seed 1): the whole run, 8 chains x 100,000 steps;
  - 94 x 95 (seed 2): 1,024 chains x 20,000 steps in two windows, 31
    chains replayed;
  - the 256-dataset mix: every chain of every dataset, 300 steps, batched
    (incremental and full scoring);
  - the served stream: its first 40 jobs (steps capped at 400), submitted
    while others run, every chain replayed.
"""
import sys

import numpy as np
import pytest

from bbscite_amd import _capi as B
from bbscite_amd import workloads as W
from tests import _oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _same_f(a, b):
    return _bits(np.float32(a)) == _bits(np.float32(b))


def _check_chain(tag, got, ref):
    """got = (best_score, best_step, final_score, final_beta) of one chain."""
    bs, st, fs, fb = got
    assert st == ref.best_step, (tag, st, ref.best_step)
    if ref.best_step != 0xFFFFFFFF:
        assert _same_f(bs, ref.best_score), (tag, bs, ref.best_score)
    assert _same_f(fs, ref.final_score), (tag, fs, ref.final_score)
    assert _same_f(fb, ref.final_tree.beta), (tag, fb, ref.final_tree.beta)


def _check_best(tag, got, results):
    """got = (am, dm, beta, score) of a run's global best."""
    b = O.chain_major_best(results)
    assert b is not None, tag
    ref = results[b]
    assert _same_f(got[3], ref.best_score), (tag, got[3], ref.best_score)
    assert _same_f(got[2], ref.best_tree.beta), tag
    assert np.array_equal(got[0], O.rows(ref.best_tree, "am")), tag
    assert np.array_equal(got[1], O.rows(ref.best_tree, "dm")), tag
    return b


def test_plain_64x63_whole_run():
    p, mut, kno, G = W.plain_64x63()
    say(f"[64x63] GPU {p.n_chains} chains x {p.n_steps} steps")
    with B.Context(p, mut, kno, G) as ctx:
        ctx.run()
        res_gpu = ctx.chain_results()
        am, dm = ctx.chain_states()
        best = ctx.best()
    say("[64x63] oracle replay")
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    ref = O.run_chains(p, s)
    for c in range(p.n_chains):
        _check_chain(c, [x[c] for x in res_gpu], ref[c])
        assert np.array_equal(am[c], O.rows(ref[c].final_tree, "am")), c
        assert np.array_equal(dm[c], O.rows(ref[c].final_tree, "dm")), c
    _check_best("64x63", best, ref)


def test_plain_94x95_windows():
    p, mut, kno, G = W.plain_94x95(n_chains=1024, n_steps=20_000)
    say(f"[94x95] GPU {p.n_chains} chains x {p.n_steps} steps in windows 6000 + 14000")
    with B.Context(p, mut, kno, G) as ctx:
        ctx.begin()
        ctx.advance(6_000)
        ctx.advance(14_000)
        res_gpu = ctx.chain_results()
        am, dm = ctx.chain_states()
    chains = list(range(16)) + list(range(83, 1024, 67))
    say(f"[94x95] oracle replay of {len(chains)} chains")
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    ref = O.run_chain_list(p, s, chains)
    for c, r in zip(chains, ref):
        _check_chain(c, [x[c] for x in res_gpu], r)
        assert np.array_equal(dm[c], O.rows(r.final_tree, "dm")), c
        assert np.array_equal(am[c], O.rows(r.final_tree, "am")), c
    # beta moves were proposed and accepted along the way
    assert len({float(x) for x in res_gpu[3][chains]}) > len(chains) // 2


@pytest.mark.parametrize("scoring", ["auto", "full"])
def test_batch_mix_every_chain(scoring):
    jobs = W.batch_mix(300)
    say(f"[batch mix] GPU 256 datasets x 64 chains x 300 steps, scoring {scoring}")
    with B.Batch(jobs, B.batch_config(waves_per_group=4, n_streams=16, scoring=scoring)) as b:
        b.run()
        got = [(b.chain_results(j), b.best(j)) for j in range(len(jobs))]
    say("[batch mix] oracle replay")
    for j, (p, mut, kno, G) in enumerate(jobs):
        s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
        ref = O.run_chains(p, s)
        for c in range(p.n_chains):
            _check_chain((j, c), [x[c] for x in got[j][0]], ref[c])
        _check_best(j, got[j][1], ref)


def test_served_stream_first_jobs():
    stream = W.served_stream(40, max_steps=400)
    say("[served] 40 jobs, slices of about 0.5 ms, submitted while others run")
    cfg = B.server_config(n_streams=16, round_cost=200_000, round_ms=0.5)
    results, ids = {}, {}
    with B.Server(cfg) as srv:
        nxt, calls = 0, 0
        while len(results) < len(stream):
            if nxt < len(stream) and (calls % 2 == 0 or nxt < 4):
                ids[srv.submit(*stream[nxt][1])] = nxt
                nxt += 1
            srv.step()
            calls += 1
            for r in srv.poll():
                results[ids[r.id]] = r
            assert calls < 200000
        _, _, n_slices = srv.stats()
    assert n_slices > 2 * len(stream)  # jobs ran over several slices each
    say("[served] oracle replay")
    for i, (_, (p, mut, kno, G)) in enumerate(stream):
        s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
        ref = O.run_chains(p, s)
        r = results[i]
        b = _check_best(i, (r.rows("am"), r.rows("dm"), r.best_beta, r.best_score), ref)
        assert r.best_chain == b and r.best_step == ref[b].best_step, i


def test_server_rejects_job_beyond_resident_set():
    """A job with more chains than the resident set would overrun its lane's
    tables: rejected with BBS_ERR_BUDGET; one that fits runs normally."""
    say("[served] resident-set limit")
    m, _, _ = B.generate_dataset(3, 10, 12, 1e-3, 0.2, 0.1)
    mut, kno = B.encode_matrix(m)
    big = B.default_params(n_chains=41, n_steps=20, seed=3, alpha_mean=1e-3, beta_mean=0.2)
    fit = B.default_params(n_chains=40, n_steps=20, seed=3, alpha_mean=1e-3, beta_mean=0.2)
    with B.Server(B.server_config(max_resident_chains=40)) as srv:
        with pytest.raises(B.BbsError) as e:
            srv.submit(big, mut, kno, 10)
        assert e.value.code == B.ERR_BUDGET
        jid = srv.submit(fit, mut, kno, 10)
        out = []
        while not out:
            srv.step()
            out = srv.poll()
    assert out[0].id == jid
    with B.Context(fit, mut, kno, 10) as ctx:
        ctx.run()
        _, _, _, score = ctx.best()
    assert _same_f(out[0].best_score, score)
