"""The serving loop (continuous batching): jobs submitted at different
times, admitted as capacity frees up and advanced slice by slice (one lane
per register layout) alongside whatever else is resident, must give exactly
the result of the same job run alone through the plain path (bbs_create +
bbs_run)."""
import numpy as np
import pytest

import bbscite_amd as B
from bbscite_amd import _capi


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def say(msg):
    print(msg, flush=True)


def test_server_rejects_bad_configuration():
    """Checked before any device is touched."""
    for kw in ({"waves_per_group": 5}, {"n_streams": 0}, {"round_cost": 0},
               {"hbm_budget_bytes": 1024}, {"admission": 2}, {"aging_s": -1.0}):
        with pytest.raises(B.BbsError) as e:
            B.Server(B.server_config(**kw))
        assert e.value.code == _capi.ERR_ARG


def _job(i, G, C, chains, steps, p_beta=0.1):
    m, _, _ = B.generate_dataset(300 + i, G, C, 1e-4, 0.2, 0.1)
    mut, kno = B.encode_matrix(m)
    p = B.default_params(n_chains=chains, n_steps=steps, seed=500 + i, alpha_mean=1e-4,
                         beta_mean=0.2 + 0.02 * (i % 3), prob_beta_change=p_beta)
    return p, mut, kno, G


JOBS = [  # G, C, chains, steps: every register layout family, uneven lengths
    (16, 16, 8, 900), (94, 95, 12, 700), (40, 80, 9, 1300), (127, 128, 5, 400),
    (63, 64, 16, 1100), (20, 120, 7, 800), (100, 33, 10, 600), (2, 5, 3, 500),
    (64, 65, 6, 1000), (30, 96, 11, 750),
]


@pytest.mark.gpu
@pytest.mark.parametrize("admission", [_capi.ADMIT_ARRIVAL, _capi.ADMIT_LEAST_WORK])
def test_served_results_equal_alone(admission):
    """Small slices (many windows per job) and a small chain budget (jobs
    queue and are admitted as others finish); some jobs arrive while others
    run. Each result equals the job's plain run bit for bit, in either
    admission order."""
    say(f"[serve] mixed jobs, admission {admission}")
    jobs = [_job(i, *spec) for i, spec in enumerate(JOBS)]
    cfg = B.server_config(round_cost=300_000, max_resident_chains=40, n_streams=3,
                          waves_per_group=2, hbm_budget_bytes=64 << 20,
                          admission=admission)
    results = {}
    ids = {}
    with B.Server(cfg) as srv:
        for i in range(4):
            ids[srv.submit(*jobs[i])] = i
        calls = 0
        nxt = 4
        while len(results) < len(jobs):
            if nxt < len(jobs) and calls % 3 == 2:
                ids[srv.submit(*jobs[nxt])] = nxt
                nxt += 1
            res, pend, ready = srv.step()
            calls += 1
            for r in srv.poll():
                results[ids[r.id]] = r
            assert calls < 20000
        chain_steps, busy, n_slices = srv.stats()
    assert chain_steps == sum(p.n_chains * p.n_steps for p, *_ in jobs)
    assert n_slices > len(jobs)
    say("[serve] plain runs")
    for i, (p, mut, kno, G) in enumerate(jobs):
        r = results[i]
        assert r.submit_s <= r.admit_s <= r.done_s
        with B.Context(p, mut, kno, G) as ctx:
            ctx.run()
            am, dm, beta, score = ctx.best()
            bs, st, _, _ = ctx.chain_results()
        assert _bits(r.best_score) == _bits(score), i
        assert _bits(r.best_beta) == _bits(beta), i
        assert np.array_equal(r.rows("am"), am) and np.array_equal(r.rows("dm"), dm), i
        assert r.best_step == st[r.best_chain]


@pytest.mark.gpu
@pytest.mark.parametrize("admission", [_capi.ADMIT_ARRIVAL, _capi.ADMIT_LEAST_WORK])
def test_admission_order(admission):
    """Room for one job at a time and four jobs of one layout pending,
    submitted longest first: arrival order admits them as submitted, least
    work first admits them shortest first."""
    say(f"[serve] admission order {admission}")
    steps = [800, 600, 400, 200]
    jobs = [_job(i, 30, 40, 16, n) for i, n in enumerate(steps)]
    cfg = B.server_config(max_resident_chains=16, n_streams=1, admission=admission)
    done = []
    with B.Server(cfg) as srv:
        ids = {srv.submit(*j): i for i, j in enumerate(jobs)}
        for _ in range(100000):
            srv.step()
            done += srv.poll()
            if len(done) == len(jobs):
                break
    assert len(done) == len(jobs)
    order = [ids[r.id] for r in sorted(done, key=lambda r: r.admit_s)]
    assert order == ([0, 1, 2, 3] if admission == _capi.ADMIT_ARRIVAL else [3, 2, 1, 0])
    # one resident at a time: each admission follows the previous completion
    by_admit = sorted(done, key=lambda r: r.admit_s)
    for a, b in zip(by_admit, by_admit[1:]):
        assert a.done_s <= b.admit_s


@pytest.mark.gpu
def test_served_lanes_with_default_configuration():
    """The default server (a resident set larger than the device runs at
    once, 16 streams): several jobs per register layout share a lane, and
    a lane with few chains may switch to the pair-score table between
    slices. Every result equals the job's plain run bit for bit."""
    say("[serve] lanes, default configuration")
    specs = [(94, 95, 40, 600), (90, 90, 24, 900), (94, 95, 8, 1500), (30, 40, 64, 700),
             (28, 33, 16, 1200), (127, 128, 12, 500), (120, 100, 6, 800), (60, 60, 32, 650)]
    jobs = [_job(50 + i, *spec) for i, spec in enumerate(specs)]
    cfg = B.server_config(n_streams=16, round_cost=400_000, round_ms=2.0)
    results = {}
    ids = {}
    with B.Server(cfg) as srv:
        for i, j in enumerate(jobs):
            ids[srv.submit(*j)] = i
        calls = 0
        while len(results) < len(jobs):
            srv.step()
            calls += 1
            for r in srv.poll():
                results[ids[r.id]] = r
            assert calls < 100000
    for i, (p, mut, kno, G) in enumerate(jobs):
        r = results[i]
        with B.Context(p, mut, kno, G) as ctx:
            ctx.run()
            am, dm, beta, score = ctx.best()
        assert _bits(r.best_score) == _bits(score), i
        assert _bits(r.best_beta) == _bits(beta), i
        assert np.array_equal(r.rows("dm"), dm), i


@pytest.mark.gpu
def test_server_rejects_jobs_beyond_its_memory():
    say("[serve] budget")
    with B.Server(B.server_config(hbm_budget_bytes=1 << 20)) as srv:
        with pytest.raises(B.BbsError) as e:
            srv.submit(*_job(0, 127, 128, 200, 10))
        assert e.value.code == _capi.ERR_BUDGET
        jid = srv.submit(*_job(1, 10, 12, 4, 50))
        while True:
            srv.step()
            out = srv.poll()
            if out:
                assert out[0].id == jid
                break
