"""GPU parity of the HIP path (through the C-ABI) against the CPU oracle.

Bar: bit-exact for every integer/bit result (counts, moves, trees, flags)
and for the float scores too (the reference's float order is reproduced, so
the 1e-4 tolerance of the specification is met with zero difference).
"""
import math
import sys

import numpy as np
import pytest

from bbscite_amd import _capi as B
from tests import _oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dataset(seed, G, C, ad=0.2, fd=1e-4, missing=0.1):
    m, par, att = B.generate_dataset(seed, G, C, fd, ad, missing)
    mut, kno = B.encode_matrix(m)
    return m, mut, kno


def _iid_matrix(seed, G, C, p1=0.3, p3=0.1):
    rng = np.random.default_rng(seed)
    u = rng.random((G, C))
    m = np.zeros((G, C), np.int8)
    m[u < p1] = 1
    m[(u >= p1) & (u < p1 + p3)] = 3
    return m


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


SHAPES = [(2, 1), (4, 6), (20, 17), (31, 64), (32, 65), (33, 97), (63, 64), (64, 65),
          (94, 95), (95, 96), (96, 33), (127, 128)]


@pytest.mark.parametrize("G,C", SHAPES)
def test_score_trees_bit_exact(G, C):
    say(f"[score] G={G} C={C}")
    m = _iid_matrix(G * 1000 + C, G, C)
    if C > 2:
        m[:, 1] = 3  # a cell without any known entry
    mut, kno = B.encode_matrix(m)
    n_trees = 64
    dm = B.random_trees(7 + G, G, n_trees)
    rng = np.random.default_rng(G)
    betas = rng.uniform(0.05, 0.95, n_trees).astype(np.float32)
    p = B.default_params(n_chains=1, n_steps=1)
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    with B.Context(p, mut, kno, G) as ctx:
        ll, cnt = ctx.score_trees(dm, with_prior=False, counts=True)  # lane per cell
        ll_mfma = ctx.score_trees(dm, with_prior=False)  # matrix-core counts, wave per tree
        ll_lanes = ctx.score_trees(dm, with_prior=False, valu_lanes=True)  # lane per tree
        llb = ctx.score_trees(dm, betas=betas, with_prior=True)
        llb_cells = ctx.score_trees(dm, betas=betas, with_prior=True, cell_lanes=True)
        llb_lanes = ctx.score_trees(dm, betas=betas, with_prior=True, valu_lanes=True)
    assert np.array_equal(_bits(ll), _bits(ll_mfma))
    assert np.array_equal(_bits(ll), _bits(ll_lanes))
    assert np.array_equal(_bits(llb), _bits(llb_cells))
    assert np.array_equal(_bits(llb), _bits(llb_lanes))
    for t in range(n_trees):
        tree = O.tree_from_rows(np.zeros_like(dm[t]), dm[t], p.beta_mean)  # scorer reads dm
        T, ocnt, _ = O.logscore_tree(s, tree, with_prior=False, counts=True)
        assert _bits(ll[t]) == _bits(T), (t, ll[t], T)
        assert np.array_equal(cnt[t], ocnt), t
        tree.beta = float(betas[t])
        Tb = O.logscore_tree(s, tree, with_prior=True)
        assert _bits(llb[t]) == _bits(Tb), (t, llb[t], Tb)


@pytest.mark.parametrize("G", [2, 16, 63, 64, 94, 127])
def test_device_random_trees_match_host(G):
    say(f"[trees] G={G}")
    m = _iid_matrix(3, G, 8)
    mut, kno = B.encode_matrix(m)
    p = B.default_params(n_chains=4, n_steps=1, seed=99)
    with B.Context(p, mut, kno, G) as ctx:
        dm_dev = ctx.random_trees_device(5, B.PURPOSE_TREE_SAMPLE, 40)
        am_i, dm_i = ctx.random_trees_device(99, B.PURPOSE_INIT_TREE, 4, with_am=True)
    assert np.array_equal(dm_dev, B.random_trees(5, G, 40))
    for c in range(4):
        am_h, dm_h = B.initial_tree(99, 0, c, G)
        assert np.array_equal(am_i[c], am_h) and np.array_equal(dm_i[c], dm_h)
        t = O.Tree()
        O.lib().orc_initial_tree(p, 0, c, G, t)
        assert np.array_equal(O.rows(t, "am"), am_h)
        assert np.array_equal(O.rows(t, "dm"), dm_h)


def test_device_libm_matches_glibc():
    say("[libm]")
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.uniform(0, 1, 400000).astype(np.float32),
        rng.uniform(-110, 90, 400000).astype(np.float32),
        rng.integers(0, 2**32, 400000, dtype=np.uint64).astype(np.uint32).view(np.float32),
        np.array([0.0, 1.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, 88.72, -103.97], np.float32),
    ])
    m = _iid_matrix(1, 4, 4)
    mut, kno = B.encode_matrix(m)
    with B.Context(B.default_params(n_chains=1, n_steps=1), mut, kno, 4) as ctx:
        e, l, l1 = ctx.eval_libm(x)
    with np.errstate(all="ignore"):
        ref_e = np.exp(x.astype(np.float32))
    # numpy's float32 exp/log are not glibc's; compare with the oracle's
    # glibc through ctypes on a subset, and exactly through libm.so.6 below.
    import ctypes

    libm = ctypes.CDLL("libm.so.6")
    libm.expf.restype = ctypes.c_float
    libm.expf.argtypes = [ctypes.c_float]
    libm.logf.restype = ctypes.c_float
    libm.logf.argtypes = [ctypes.c_float]
    libm.log.restype = ctypes.c_double
    libm.log.argtypes = [ctypes.c_double]
    sub = np.arange(0, len(x), 7)
    for i in sub:
        xi = float(x[i])
        ge = np.float32(libm.expf(xi))
        gl = np.float32(libm.logf(xi))
        gl1 = np.float32(libm.log(1.0 - xi)) if not math.isnan(xi) else np.float32(np.nan)
        for got, ref in ((e[i], ge), (l[i], gl), (l1[i], gl1)):
            if math.isnan(ref):
                assert math.isnan(got), (hex(x[i:i+1].view(np.uint32)[0]), got, ref)
            else:
                assert _bits(got) == _bits(ref), (hex(x[i:i+1].view(np.uint32)[0]), got, ref)
    del ref_e


def _chain_case(G, C, seed, n_chains, n_steps, p_beta, beta_mean=0.25, beta_sd=0.1,
                ad=0.2, fd=1e-4, missing=0.1, alpha=None):
    m, mut, kno = _dataset(seed, G, C, ad=ad, fd=fd, missing=missing)
    p = B.default_params(
        n_chains=n_chains, n_steps=n_steps, seed=seed,
        alpha_mean=alpha if alpha is not None else max(fd, 1e-6),
        beta_mean=beta_mean, beta_sd=beta_sd, prob_beta_change=p_beta,
    )
    return p, mut, kno


CHAIN_CASES = [
    # G, C, seed, chains, steps, p_beta, beta_mean, beta_sd
    (4, 15, 1, 3, 400, 0.0, 0.25, 0.1),
    (16, 20, 2, 4, 300, 0.2, 0.25, 0.1),
    (63, 64, 3, 4, 200, 0.1, 0.2, 0.1),
    (64, 33, 4, 2, 150, 0.1, 0.25, 0.1),
    (94, 95, 5, 3, 150, 0.1, 0.3, 0.05),
    (127, 128, 6, 2, 100, 0.1, 0.4309, 0.1),
    (40, 80, 7, 2, 150, 0.1, 0.25, 0.1),
    (20, 96, 8, 2, 150, 0.0, 0.25, 0.1),
    (64, 65, 9, 2, 120, 0.1, 0.3, 0.1),
    (3, 70, 10, 3, 200, 0.1, 0.25, 0.1),
    (2, 1, 11, 2, 100, 0.1, 0.25, 0.1),
    # half the moves change beta: rounds of the speculative chain with
    # several beta proposals, accepted and rejected ones
    (20, 70, 12, 3, 300, 0.5, 0.3, 0.1),
    # over 96 cells (two full cell slots; the half table), one and two node
    # slots, many beta moves
    (100, 110, 13, 3, 200, 0.1, 0.3, 0.1),
    (50, 97, 14, 2, 250, 0.4, 0.25, 0.1),
    (10, 120, 15, 2, 300, 0.2, 0.3, 0.1),
    # the most genes with two cells; nine entries in ten missing
    (127, 2, 17, 2, 150, 0.1, 0.3, 0.1),
    (30, 40, 18, 3, 200, 0.2, 0.25, 0.1, 0.9),
]


MODES = ["full", "incremental", "half", "split2", "split4", "spec2", "spec4", "spec8"]


def _set_mode(ctx, mode):
    """full: every pair rescored; incremental: the pair-score table, one
    wave per chain; half: the table of cells 0..63 only, the others rescored
    every step (over 96 cells); split2 / split4: the table with each chain
    split over 2 or 4 waves by cells; spec2 / spec4 / spec8: the table with
    2, 4 or 8 waves per chain proposing consecutive steps at once."""
    try:
        ctx.set_scoring_mode(mode if mode in ("full", "half") else "incremental")
        if mode != "full":
            ctx.set_speculation(int(mode[4:]) if mode.startswith("spec") else 1)
            ctx.set_chain_split(int(mode[5:]) if mode.startswith("split") else 1)
    except B.BbsError as e:
        if mode != "full" and e.code == -2:
            pytest.skip(f"no {mode} scorer for this size (LDS of a workgroup, or a wave "
                        "without cells)")
        raise


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("case", CHAIN_CASES, ids=lambda c: f"G{c[0]}C{c[1]}")
def test_chain_trace_replay(case, mode):
    G, C, seed, R, S, pb, bm, bsd = case[:8]
    missing = case[8] if len(case) > 8 else 0.1
    say(f"[chain] G={G} C={C} chains={R} steps={S} {mode}")
    p, mut, kno = _chain_case(G, C, seed, R, S, pb, beta_mean=bm, beta_sd=bsd, missing=missing)
    with B.Context(p, mut, kno, G) as ctx:
        _set_mode(ctx, mode)
        ctx.set_trace(R)
        ctx.set_proposal_trace(R)
        ctx.run()
        trace = ctx.trace(R)
        prop_am, prop_dm = ctx.proposal_trace(R)
        bs, st, fs, fb = ctx.chain_results()
        am_f, dm_f = ctx.chain_states()
        best_am, best_dm, best_beta, best_score = ctx.best()
    # the untraced kernel (which realises only what each move type uses)
    # gives the same chains as the traced one checked below
    with B.Context(p, mut, kno, G) as ctx:
        _set_mode(ctx, mode)
        ctx.run()
        plain = ctx.chain_results(), ctx.chain_states(), ctx.best()
    for x, y in zip(plain[0], (bs, st, fs, fb)):
        assert np.array_equal(_bits(x) if x.dtype == np.float32 else x,
                              _bits(y) if y.dtype == np.float32 else y)
    assert np.array_equal(plain[1][0], am_f) and np.array_equal(plain[1][1], dm_f)
    assert np.array_equal(plain[2][1], best_dm) and _bits(plain[2][3]) == _bits(best_score)
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    for c in range(R):
        res, otr, oam, odm = O.run_chain_proposals(p, s, c, c == 0, G + 1)
        # every step's proposed tree, accepted or not, matrix for matrix
        # (the reference's move test checks the full matrices after each
        # update, test/MutationTree.cpp:697-735)
        for name, got, want in (("am", prop_am[c], oam), ("dm", prop_dm[c], odm)):
            bad = np.nonzero((got != want).any(axis=(1, 2)))[0]
            assert len(bad) == 0, (c, name, bad[:5], trace[c][bad[:3]])
        for field in ("move_type", "v", "w", "parent_of_v", "parent_of_w",
                      "descendant_of_v", "nondescendant_of_v", "flags"):
            bad = np.nonzero(trace[c][field] != otr[field])[0]
            assert len(bad) == 0, (c, field, bad[:5], trace[c][bad[:3]], otr[bad[:3]])
        for field in ("new_beta", "proposed_score"):
            bad = np.nonzero(_bits(trace[c][field]) != _bits(otr[field]))[0]
            assert len(bad) == 0, (c, field, bad[:5], trace[c][field][bad[:3]], otr[field][bad[:3]])
        assert _bits(fs[c]) == _bits(res.final_score)
        assert np.array_equal(am_f[c], O.rows(res.final_tree, "am"))
        assert np.array_equal(dm_f[c], O.rows(res.final_tree, "dm"))
        assert st[c] == res.best_step
        if res.best_step != 0xFFFFFFFF:
            assert _bits(bs[c]) == _bits(res.best_score)
    t, sc, ch, stp = O.run(p, s)
    assert _bits(best_score) == _bits(sc)
    assert np.array_equal(best_am, O.rows(t, "am"))
    assert np.array_equal(best_dm, O.rows(t, "dm"))
    assert _bits(best_beta) == _bits(t.beta)
    # every move type occurs
    types = set(np.unique(trace["move_type"]).tolist())
    assert {1, 2, 3} <= types and (pb == 0 or 0 in types)


@pytest.mark.parametrize("G,C,R", [(63, 64, 512), (63, 64, 1024), (94, 95, 256)])
def test_automatic_speculation_with_many_chains(G, C, R):
    """With hundreds of chains the default runs 8, 4 or 2 speculative waves
    per chain (up to two waves per SIMD); every chain's results and final
    tree equal one wave per chain."""
    say(f"[auto spec] G={G} C={C} chains={R}")
    p, mut, kno = _chain_case(G, C, 16, R, 300, 0.1, beta_mean=0.3, beta_sd=0.05)
    out = []
    for spec in (0, 1):
        with B.Context(p, mut, kno, G) as ctx:
            ctx.set_speculation(spec)
            ctx.run()
            out.append((ctx.chain_results(), ctx.chain_states(), ctx.best()))
    (r0, s0, b0), (r1, s1, b1) = out
    for x, y in zip(r0, r1):
        assert np.array_equal(_bits(x) if x.dtype == np.float32 else x,
                              _bits(y) if y.dtype == np.float32 else y)
    assert np.array_equal(s0[0], s1[0]) and np.array_equal(s0[1], s1[1])
    assert np.array_equal(b0[1], b1[1]) and _bits(b0[3]) == _bits(b1[3])


@pytest.mark.parametrize("G,C", [(94, 95), (63, 64), (127, 96), (30, 128), (5, 7), (120, 110)])
def test_incremental_equals_full_over_long_runs(G, C):
    """Thousands of steps with many accepted moves of every kind: the
    pair-score table stays exact, so both scoring modes give identical
    chains (scores, best steps, final trees, betas)."""
    say(f"[incremental] G={G} C={C}")
    p, mut, kno = _chain_case(G, C, 12, 16, 4000, 0.1, beta_mean=0.3, beta_sd=0.05)
    out = {}
    for mode in MODES:
        with B.Context(p, mut, kno, G) as ctx:
            try:
                _set_mode(ctx, mode)
            except pytest.skip.Exception:
                continue
            ctx.begin()
            for n in (1000, 3000):
                ctx.advance(n)
            out[mode] = (ctx.chain_results(), ctx.chain_states(), ctx.best())
    r0, s0, b0 = out["full"]
    assert len(out) >= 2
    for mode, (r1, s1, b1) in out.items():
        for x, y in zip(r0, r1):
            assert np.array_equal(_bits(x) if x.dtype == np.float32 else x,
                                  _bits(y) if y.dtype == np.float32 else y), mode
        for x, y in zip(s0, s1):
            assert np.array_equal(x, y), mode
        assert np.array_equal(b0[0], b1[0]) and np.array_equal(b0[1], b1[1]), mode
        assert _bits(b0[3]) == _bits(b1[3]), mode


def test_advance_windows_equal_one_run():
    say("[windows]")
    p, mut, kno = _chain_case(30, 40, 9, 8, 300, 0.1)
    with B.Context(p, mut, kno, 30) as ctx:
        ctx.set_trace(3)
        ctx.set_proposal_trace(3)
        ctx.run()
        one = ctx.chain_results(), ctx.chain_states(), ctx.best()
        one_tr = ctx.trace(3), ctx.proposal_trace(3)
    with B.Context(p, mut, kno, 30) as ctx:
        ctx.set_trace(3)
        ctx.set_proposal_trace(3)
        ctx.begin()
        for n in (1, 99, 100, 100):
            ctx.advance(n)
        split = ctx.chain_results(), ctx.chain_states(), ctx.best()
        split_tr = ctx.trace(3), ctx.proposal_trace(3)
    for a, b in zip(one[0], split[0]):
        assert np.array_equal(a, b)
    for a, b in zip(one[1], split[1]):
        assert np.array_equal(a, b)
    assert np.array_equal(one[2][0], split[2][0]) and one[2][3] == split[2][3]
    # the records of every step land at the same place across windows
    assert np.array_equal(one_tr[0], split_tr[0])
    for a, b in zip(one_tr[1], split_tr[1]):
        assert np.array_equal(a, b)


def test_caller_initial_states():
    """bbs_set_initial_states with and without scores: same as the oracle
    started from the same trees."""
    say("[init states]")
    G, C, R = 12, 10, 3
    p, mut, kno = _chain_case(G, C, 4, R, 120, 0.1)
    trees = [B.parent_vector_to_matrix(B.pruefer_to_parent_vector(
        np.random.default_rng(c).integers(0, G + 1, G - 1))) for c in range(R)]
    am = np.stack([t[0] for t in trees])
    dm = np.stack([t[1] for t in trees])
    beta = np.full(R, 0.3, np.float32)
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    for give_score in (False, True):
        with B.Context(p, mut, kno, G) as ctx:
            score = None
            if give_score:
                score = np.array([O.logscore_tree(s, O.tree_from_rows(am[c], dm[c], 0.3))
                                  for c in range(R)], np.float32)
            ctx.set_initial_states(am, dm, beta, score)
            ctx.run()
            bs, st, fs, fb = ctx.chain_results()
        for c in range(R):
            res = O.run_chain(p, s, c, c == 0, init=O.tree_from_rows(am[c], dm[c], 0.3))
            assert _bits(fs[c]) == _bits(res.final_score)
            assert st[c] == res.best_step


def test_score_trees_degenerate_beta():
    """beta exactly 0 or 1 gives infinite log-probabilities; NaN pairs at
    node 0 stick in the maximum as in the reference (all three scorers)."""
    G, C = 10, 12
    m = _iid_matrix(5, G, C)
    mut, kno = B.encode_matrix(m)
    dm = B.random_trees(2, G, 70)
    betas = np.array([0.0, 1.0] * 35, np.float32)
    p = B.default_params(n_chains=1, n_steps=1)
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    with B.Context(p, mut, kno, G) as ctx:
        a = ctx.score_trees(dm, betas=betas, with_prior=False)
        b = ctx.score_trees(dm, betas=betas, with_prior=False, cell_lanes=True)
        c = ctx.score_trees(dm, betas=betas, with_prior=False, valu_lanes=True)
    for t in range(70):
        tree = O.tree_from_rows(np.zeros_like(dm[t]), dm[t], float(betas[t]))
        ref = O.logscore_tree(s, tree, with_prior=False)
        for got in (a[t], b[t], c[t]):
            if math.isnan(ref):
                assert math.isnan(got), t
            else:
                assert _bits(got) == _bits(ref), (t, got, ref)


def test_bench_score_random_head_matches_oracle():
    say("[bench score head]")
    G, C = 94, 95
    m = _iid_matrix(3, G, C)
    mut, kno = B.encode_matrix(m)
    p = B.default_params(n_chains=1, n_steps=1)
    with B.Context(p, mut, kno, G) as ctx:
        ms, head = ctx.bench_score_random(3, 4096, 2, with_prior=False, n_head=32)
        ms2, head2 = ctx.bench_score_random(3, 4096, 1, with_prior=False, n_head=32,
                                            cell_lanes=True)
        ms3, head3 = ctx.bench_score_random(3, 4096, 1, with_prior=False, n_head=32,
                                            valu_lanes=True)
    assert (ms > 0).all()
    assert np.array_equal(_bits(head), _bits(head2))
    assert np.array_equal(_bits(head), _bits(head3))
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    dm = B.random_trees(3, G, 32)
    rate, ll = O.bench_score(s, dm, p.beta_mean, 1)
    assert np.array_equal(_bits(head), _bits(ll))


def test_errors_are_reported():
    m = _iid_matrix(1, 4, 4)
    mut, kno = B.encode_matrix(m)
    with pytest.raises(B.BbsError) as e:
        B.Context(B.default_params(n_chains=1, n_steps=1, beta_mean=1.5), mut, kno, 4)
    assert e.value.code == B.ERR_ARG
    with B.Context(B.default_params(n_chains=2, n_steps=10), mut, kno, 4) as ctx:
        with pytest.raises(B.BbsError) as e:
            ctx.best()
        assert e.value.code == B.ERR_STATE
        with pytest.raises(B.BbsError):
            ctx.set_trace(5)
        with pytest.raises(B.BbsError) as e:
            ctx.set_proposal_trace(3)
        assert e.value.code == B.ERR_ARG
        with pytest.raises(B.BbsError) as e:
            ctx.proposal_trace(0)
        assert e.value.code == B.ERR_STATE
        # the half table serves datasets of over 96 cells only
        with pytest.raises(B.BbsError) as e:
            ctx.set_scoring_mode("half")
        assert e.value.code == B.ERR_SIZE
        ctx.begin()
        with pytest.raises(B.BbsError):
            ctx.advance(11)
        # waves per chain: 0, 1, 2 or 4; a split needs a cell for every wave
        with pytest.raises(B.BbsError) as e:
            ctx.set_chain_split(3)
        assert e.value.code == B.ERR_ARG
        ctx.set_chain_split(4)  # 4 cells, one per wave
        with pytest.raises(B.BbsError) as e:
            ctx.set_speculation(3)
        assert e.value.code == B.ERR_ARG
        ctx.set_speculation(8)
    m3 = _iid_matrix(2, 4, 3)
    mut3, kno3 = B.encode_matrix(m3)
    with B.Context(B.default_params(n_chains=2, n_steps=10), mut3, kno3, 4) as ctx:
        with pytest.raises(B.BbsError) as e:
            ctx.set_chain_split(4)  # 3 cells cannot feed 4 waves
        assert e.value.code == B.ERR_SIZE
        ctx.set_chain_split(2)


def _mix(n_jobs, seed0, steps, chains_range=(2, 6)):
    rng = np.random.default_rng(seed0)
    jobs = []
    for i in range(n_jobs):
        G = int(rng.integers(2, 128))
        C = int(rng.integers(1, 129))
        R = int(rng.integers(*chains_range))
        m, _, _ = B.generate_dataset(seed0 + i, G, C, 1e-3, 0.2, 0.1)
        mut, kno = B.encode_matrix(m)
        p = B.default_params(n_chains=R, n_steps=steps, seed=seed0 + 100 + i,
                             alpha_mean=1e-3, beta_mean=0.2 + 0.01 * (i % 5),
                             prob_beta_change=0.1 * (i % 2))
        jobs.append((p, mut, kno, G))
    return jobs


PERSISTENT_HANG = (
    "the persistent graph-captured batched kernel hung on its first GPU run "
    "(no output for 180 s); its cause has not been found, so that path is "
    "switched off (bbs_batch_run rejects persistent = 1) and stays off the GPU"
)


@pytest.mark.skip(reason=PERSISTENT_HANG)
def test_batch_equals_plain_runs():
    """(persistent form) every job of a mixed batch gives the plain result."""


@pytest.mark.skip(reason=PERSISTENT_HANG)
def test_batch_configs_agree_and_autotune():
    """(persistent form) launch shapes agree; autotuner picks one."""


def test_batch_persistent_mode_is_rejected():
    jobs = _mix(1, 5, 10)
    with B.Batch(jobs, B.batch_config()) as b:
        pass
    cfg = B.batch_config()
    cfg.persistent = 1
    with B.Batch(jobs, cfg) as b:
        with pytest.raises(B.BbsError) as e:
            b.run()
        assert e.value.code == B.ERR_STATE


@pytest.mark.parametrize("scoring", ["auto", "full"])
def test_batched_items_equal_plain_runs(scoring):
    """Every job of a mixed batch (one wave per item, per-layout launches,
    incremental or full scoring) gives exactly the plain bbs_run result, on
    every rerun."""
    say(f"[batched items] {scoring}")
    jobs = _mix(12, 40, 150)
    with B.Batch(jobs, B.batch_config(waves_per_group=2, n_streams=4, scoring=scoring)) as b:
        b.run()
        batched = [(b.chain_results(j), b.best(j)) for j in range(len(jobs))]
        b.run()
        again = [(b.chain_results(j), b.best(j)) for j in range(len(jobs))]
    say("[batched items] plain runs")
    for j, (p, mut, kno, G) in enumerate(jobs):
        with B.Context(p, mut, kno, G) as ctx:
            ctx.run()
            plain = ctx.chain_results(), ctx.best()
        for got in (batched[j], again[j]):
            for x, y in zip(got[0], plain[0]):
                assert np.array_equal(_bits(x) if x.dtype == np.float32 else x,
                                      _bits(y) if y.dtype == np.float32 else y), j
            assert np.array_equal(got[1][0], plain[1][0]) and np.array_equal(got[1][1], plain[1][1])
            assert _bits(got[1][3]) == _bits(plain[1][3])


def test_batched_largest_layouts_equal_plain_runs():
    """The 128-cell layouts with two node slots (70 KB tables) and the
    paired-slot layout with four gene words: few items, so the batched form
    takes the tables; every job equals its plain run."""
    say("[batched largest layouts]")
    jobs = []
    for i, (G, C, R) in enumerate(((127, 128, 3), (100, 110, 2), (120, 96, 3), (65, 97, 2))):
        m, _, _ = B.generate_dataset(70 + i, G, C, 1e-3, 0.2, 0.1)
        mut, kno = B.encode_matrix(m)
        p = B.default_params(n_chains=R, n_steps=300, seed=80 + i, alpha_mean=1e-3,
                             beta_mean=0.25, prob_beta_change=0.1)
        jobs.append((p, mut, kno, G))
    with B.Batch(jobs, B.batch_config(waves_per_group=2, n_streams=4)) as b:
        b.run()
        got = [(b.chain_results(j), b.best(j)) for j in range(len(jobs))]
    for j, (p, mut, kno, G) in enumerate(jobs):
        with B.Context(p, mut, kno, G) as ctx:
            ctx.run()
            plain = ctx.chain_results(), ctx.best()
        for x, y in zip(got[j][0], plain[0]):
            assert np.array_equal(_bits(x) if x.dtype == np.float32 else x,
                                  _bits(y) if y.dtype == np.float32 else y), j
        assert np.array_equal(got[j][1][1], plain[1][1]) and _bits(got[j][1][3]) == _bits(plain[1][3])


def test_batched_items_configs_agree_and_autotune():
    say("[batched configs]")
    jobs = _mix(6, 77, 60)
    ref = None
    # the last two: every layout with the two-waves-per-SIMD register budget,
    # and none
    for wpg, nst, order, sc, tw in ((1, 1, 0, "auto", None), (4, 2, 1, "auto", None),
                                    (2, 4, 0, "full", None), (4, 1, 0, "full", None),
                                    (2, 16, 0, "auto", None), (4, 4, 0, "auto", 0xFFFFFF),
                                    (1, 4, 1, "auto", 0)):
        with B.Batch(jobs, B.batch_config(wpg, nst, order, scoring=sc,
                                          two_wave_layouts=tw)) as b:
            b.run()
            res = [b.chain_results(j)[2] for j in range(len(jobs))]
        if ref is None:
            ref = res
        else:
            for x, y in zip(ref, res):
                assert np.array_equal(_bits(x), _bits(y))
    cfg, ms = B.batch_autotune(jobs, probe_steps=20)
    assert cfg.waves_per_group in (1, 2) and cfg.n_streams in (2, 16) and ms > 0
    assert cfg.two_wave_layouts & ~0xFFFFFF == 0


def test_batched_graph_replays_equal_eager():
    """graph = 1: the first run captures initial trees, cleared maxima and
    every layout launch on four streams in a hipGraph; later runs replay it.
    Every replay, with eager batched and plain runs in between, gives the
    eager batched result exactly."""
    say("[batched graph] eager")
    jobs = _mix(8, 61, 120)
    with B.Batch(jobs, B.batch_config(waves_per_group=2, n_streams=4)) as b:
        b.run()
        eager = [(b.chain_results(j), b.best(j)) for j in range(len(jobs))]
    with B.Batch(jobs, B.batch_config(waves_per_group=2, n_streams=4, graph=True)) as g:
        for rep in range(4):
            say(f"[batched graph] replay {rep}")
            g.run()
            got = [(g.chain_results(j), g.best(j)) for j in range(len(jobs))]
            for j in range(len(jobs)):
                for x, y in zip(got[j][0], eager[j][0]):
                    assert np.array_equal(_bits(x) if x.dtype == np.float32 else x,
                                          _bits(y) if y.dtype == np.float32 else y), (rep, j)
                assert np.array_equal(got[j][1][0], eager[j][1][0]), (rep, j)
                assert _bits(got[j][1][3]) == _bits(eager[j][1][3]), (rep, j)
            # eager work between replays: another batch and a plain run
            with B.Batch(jobs[:2], B.batch_config()) as other:
                other.run()
            p, mut, kno, G = jobs[rep]
            with B.Context(p, mut, kno, G) as ctx:
                ctx.run()
