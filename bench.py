"""Benchmark of the MI355X bbSCITE chain search (one GPU).

Workload (the plain method at the 96-bit limit): 94 genes x 95 cells,
synthetic data from seed 2 (uniform Pruefer tree, uniform attachment,
-ad 0.3, -fd 1e-4, 15 % missing), 1,024 chains with beta-change moves
(beta prior mean 0.3, sd 0.05, -e 0.1 with the default move mix), one
wavefront per chain. One bench step = every chain advances by
--chain-steps steps (default 100,000); the default warmup 1 + steps 9 runs
exactly 1,024 chains x 1,000,000 steps.

Printed: one JSON line with the chain steps/s of the timed steps, the
roofline fraction of the chain kernel against the gfx950 int32 VALU peak
(HIP-event kernel time on the launch stream), an oracle check of the timed
run, the 63 x 64 workload, the bulk-scoring throughput (2^20 random trees
against a 94 x 95 matrix), the 256-dataset batch, the served stream, and
the CPU oracle (built -O3 -march=native on this host) timed on the host
cores of the same machine.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The batched form and the server spread their launches over up to 16
# streams; HIP runs them at once only on as many hardware queues as the
# process asks for (read when HIP initialises). The environment may already
# hold a smaller value (4, HIP's default), so it is set, not defaulted.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

from bbscite_amd import _capi as B  # noqa: E402
from bbscite_amd import workloads as WL  # noqa: E402

G, C = 94, 95
N = G + 1
W32 = (G + 31) // 32
# Algorithmic integer work per scored tree / chain step: per (cell, node)
# pair 2 AND + 2 popcount per 32-bit gene word, plus 2 subtractions.
ALG_INT_OPS = N * C * (4 * W32 + 2)  # 126,350 at 94 x 95
# gfx950 int32 VALU peak: 256 CUs x 64 lanes x 2.4 GHz.
P_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12
# MI355X dense int8 MFMA peak (published), in tera-ops (2 per multiply-add)
P_INT8_MFMA_TOPS = 5000.0
# HBM bytes of one 100,000-step chain launch at 94 x 95 with 1,024 chains,
# from rocprofv3 FETCH_SIZE (x 2.0) and WRITE_SIZE (x 1.0), the factors a
# plain 1 GiB device copy gives those counters: 2,589 KiB x 2 + 6,304 KiB.
TRAFFIC_100K_WINDOW = (2589 * 2 + 6304) * 1024
FPGA_KSTEPS = 2938.25  # reference's published chain steps/s at 95 x 94 (kilo)
# Served slices of about 50 ms: fewer relaunches than the library's 20 ms
# default; a newly admitted job waits at most one slice for its lane.
SERVE_ROUND_MS = 50.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def chain_kernel_name(genes, cells):
    """Template instance the chain launch picks for this size (the layout
    rule of shape_of in bbs_kernels.hip: node slots, cell slots with the
    paired half slot for 65..96 cells, 32-bit gene words)."""
    ns = 1 if genes + 1 <= 64 else 2
    cs = 1 if cells <= 64 else (3 if cells <= 96 else 2)
    w = max(1, (genes + 31) // 32)
    if ns == 2:
        w = max(w, 2)
    return f"chain_kernel<{ns},{cs},{w}>"


def make_context(n_chains, n_steps):
    p, mut, kno, _ = WL.plain_94x95(n_chains, n_steps)
    return p, mut, kno


def host_cpu():
    """The host CPU as lscpu names it, and the cores this process may use."""
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.lower().startswith("model name"):
                model = line.split(":", 1)[1].strip()
    except Exception:  # pragma: no cover
        pass
    return {"model": model, "nproc": os.cpu_count(),
            "usable_cores": len(os.sched_getaffinity(0))}


_NATIVE = None


def oracle_for_timing():
    """The oracle built -O3 -march=native -ffp-contract=off for this host (in
    a temporary directory, so it never reaches another machine); its results
    equal the -O2 build's bit for bit. Falls back to the -O2 build."""
    global _NATIVE
    from tests import _oracle as O

    if _NATIVE is None:
        out = os.path.join(tempfile.mkdtemp(prefix="bbs_oracle_"), "native")
        r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native",
                            f"NATIVE_OUT={out}"], capture_output=True, text=True)
        if r.returncode == 0:
            _NATIVE = (os.path.join(out, "liboracle.so"),
                       "gcc -O3 -march=native -ffp-contract=off -fopenmp")
        else:
            log("[bench] native oracle build failed, timing the -O2 build:", r.stderr[-500:])
            _NATIVE = (O.LIB_PATH, "gcc -O2 -ffp-contract=off -fopenmp")
        O.use_library(_NATIVE[0])
    return O, _NATIVE[1]


def cpu_baseline(p, mut, kno, genes, label, seconds_budget=20.0):
    """The CPU oracle's chain loop on the host cores (bounded sample), one
    chain per thread; the 1-core rate is measured first."""
    try:
        O, build = oracle_for_timing()
    except Exception as e:  # pragma: no cover
        log("cpu baseline unavailable:", e)
        return None
    cores = min(16, len(os.sched_getaffinity(0)))
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, genes)
    # calibrate on one core, then size the sample to ~seconds_budget of CPU work
    r1, _ = O.bench_chains(p, s, 1, 400, 1)
    r1, _ = O.bench_chains(p, s, 1, max(400, int(r1 * 2.0)), 1)
    steps_total = max(cores * 200, int(r1 * seconds_budget))
    n_chains = cores * 4
    n_steps = max(100, steps_total // n_chains)
    rate, _ = O.bench_chains(p, s, n_chains, n_steps, cores)
    return {
        "value": rate,
        "unit": "chain steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"oracle chain loop at {label}, {n_chains} chains x {n_steps} steps "
                  f"(one chain per thread; 1-core rate {r1:.0f} steps/s)",
        "one_core_value": r1,
        "build": build,
        "cpu": host_cpu(),
    }


def scoring_matrix():
    return WL.scoring_matrix()


def scoring_bench(reps=5, peak=P_INT32_TOPS):
    """2^20 uniform random trees (seed 3) against one 94 x 95 iid matrix."""
    mut, kno = scoring_matrix()
    p = B.default_params(n_chains=1, n_steps=1)
    n_trees = 1 << 20
    with B.Context(p, mut, kno, G) as ctx:
        ms, ll_gpu = ctx.bench_score_random(3, n_trees, reps, with_prior=False, n_head=n_trees)
        ms_v, ll_v = ctx.bench_score_random(3, n_trees, 3, with_prior=False, n_head=n_trees,
                                            valu_lanes=True)
    k_ms = float(np.median(ms))
    trees_s = n_trees / (k_ms / 1e3)
    v_ms = float(np.median(ms_v))
    v_trees_s = n_trees / (v_ms / 1e3)
    # The pair counts run on the matrix cores (int8 MFMA over 96 x 96 x 96
    # tiles, two matrices); what bounds the kernel is the float part of each
    # (cell, node) pair on the vector unit: 2 cvt, 2 sub, 4 mul, 3 add, 1 max.
    fp_ops = N * C * 12
    achieved = fp_ops * trees_s / 1e12
    macs = 2 * 96 * 96 * 96
    return {
        "trees_per_s": trees_s,
        "kernel_ms": k_ms,
        "roofline": {"bound": "valu-fp32", "achieved": achieved, "peak": peak,
                     "unit": "TOP/s", "frac": achieved / peak,
                     "alg_valu_ops_per_tree": fp_ops,
                     "mfma_int8": {"achieved_tops": 2 * macs * trees_s / 1e12,
                                   "peak_tops": P_INT8_MFMA_TOPS,
                                   "frac": 2 * macs * trees_s / 1e12 / P_INT8_MFMA_TOPS}},
        "alg_int_tops": ALG_INT_OPS * trees_s / 1e12,
        "counted_bits_per_s": trees_s * 4 * N ** 3,
        # the same trees through the vector-unit scorer (one tree per lane,
        # AND + popcount over 32-bit words): its integer-op roofline
        "valu_scorer": {"kernel": "score_lanes_kernel", "trees_per_s": v_trees_s,
                        "kernel_ms": v_ms,
                        "alg_int_tops": ALG_INT_OPS * v_trees_s / 1e12,
                        "peak": peak,
                        "frac": ALG_INT_OPS * v_trees_s / 1e12 / peak,
                        "scores_identical_to_matrix_core_scorer":
                            bool(np.array_equal(ll_gpu.view(np.uint32), ll_v.view(np.uint32)))},
        "_ll_gpu": ll_gpu,
    }


def cpu_scoring_baseline(ll_gpu, seconds_budget=5.0):
    """The oracle's CPUTreeScorer-structured scorer (node-outer, cell-inner,
    popcounts over 64-bit words, TreeScorer semantics, no prior) over the
    first trees of the same seed-3 stream, parallel over trees on the host
    cores (bounded sample); every CPU score is compared bit for bit with the
    GPU's score of the same tree."""
    try:
        O, build = oracle_for_timing()
    except Exception as e:  # pragma: no cover
        log("cpu scoring baseline unavailable:", e)
        return None
    cores = min(16, len(os.sched_getaffinity(0)))
    mut, kno = scoring_matrix()
    p = B.default_params(n_chains=1, n_steps=1)
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    r1, _ = O.bench_score(s, B.random_trees(3, G, 256), p.beta_mean, 1)
    r1, _ = O.bench_score(s, B.random_trees(3, G, max(256, int(r1))), p.beta_mean, 1)
    n = int(min(1 << 20, max(cores * 256, r1 * cores * seconds_budget)))
    rate, ll = O.bench_score(s, B.random_trees(3, G, n), p.beta_mean, cores)
    mismatches = int(np.count_nonzero(ll.view(np.uint32) != ll_gpu[:n].view(np.uint32)))
    return {
        "value": rate,
        "unit": "trees/s",
        "cores": cores,
        "kind": "port",
        "sample": f"first {n} trees of the seed-3 stream at 94x95 (1-core rate {r1:.0f} trees/s)",
        "one_core_value": r1,
        "counted_bits_per_s": rate * 4 * N ** 3,
        "trees_compared_with_gpu": n,
        "score_mismatches_vs_gpu": mismatches,
        "build": build,
        "cpu": host_cpu(),
    }


def batch_mix_jobs(steps, chains=64, n_datasets=256):
    return WL.batch_mix(steps, chains, n_datasets)


def batched_bench(steps_batched=200_000, steps_plain=2_000, probe_steps=20_000):
    """The 256-dataset uneven mix, plain launches vs the batched
    per-layout launches. Plain launches run one dataset each, serially. At
    the plain step count the batch is also timed as an eager run and as a
    replay of its hipGraph capture."""
    jobs_plain = batch_mix_jobs(steps_plain)
    t_plain = 0.0
    bests_plain = []
    for p, mut, kno, G in jobs_plain:
        with B.Context(p, mut, kno, G) as ctx:
            t_plain += ctx.run() / 1e3
            bests_plain.append(ctx.best()[3])
    plain_steps = sum(p.n_chains * p.n_steps for p, *_ in jobs_plain)
    cfg, probe_ms = B.batch_autotune(batch_mix_jobs(probe_steps), probe_steps)
    log(f"[bench] batch autotune: wpg={cfg.waves_per_group} streams={cfg.n_streams} "
        f"order={cfg.order} two-wave layouts {two_wave_codes(cfg.two_wave_layouts)} "
        f"({probe_ms:.1f} ms at {probe_steps} steps)")
    with B.Batch(jobs_plain, cfg) as b:
        b.run()
        same = all(np.float32(b.best(j)[3]).tobytes() == np.float32(bests_plain[j]).tobytes()
                   for j in range(len(jobs_plain)))
        eager_short = min(b.run() for _ in range(3))
    # the same short batch captured once in a hipGraph and replayed
    gcfg = B.batch_config(cfg.waves_per_group, cfg.n_streams, cfg.order, graph=True,
                          two_wave_layouts=cfg.two_wave_layouts)
    with B.Batch(jobs_plain, gcfg) as g:
        g.run()
        graph_short = min(g.run() for _ in range(3))
        same_graph = all(np.float32(g.best(j)[3]).tobytes() == np.float32(bests_plain[j]).tobytes()
                         for j in range(len(jobs_plain)))
    jobs = batch_mix_jobs(steps_batched)
    # datasets whose timed batched result is checked against a plain run of
    # the same step count: the smallest, the largest and two in between
    by_size = sorted(range(len(jobs)), key=lambda j: (jobs[j][3] + 1) * jobs[j][1].shape[0])
    sample = [by_size[0], by_size[len(by_size) // 3], by_size[2 * len(by_size) // 3],
              by_size[-1]]
    with B.Batch(jobs, cfg) as b:
        ms = b.run()
        total = b.total_steps
        timed_best = {j: b.best(j) for j in sample}
    same_timed = True
    for j in sample:
        p, mut, kno, Gj = jobs[j]
        with B.Context(p, mut, kno, Gj) as ctx:
            ctx.run()
            am, dm, beta, score = ctx.best()
        tb = timed_best[j]
        same_timed &= (np.float32(tb[3]).tobytes() == np.float32(score).tobytes()
                       and np.float32(tb[2]).tobytes() == np.float32(beta).tobytes()
                       and np.array_equal(tb[1], dm) and np.array_equal(tb[0], am))
    log(f"[bench] timed batch vs plain runs of datasets {sample}: identical={same_timed}")
    rate_b = total / (ms / 1e3)
    rate_p = plain_steps / t_plain
    # the reference algorithm's int work of one step of each dataset,
    # N C (4W + 2), over the batched run
    ops = sum(p.n_chains * p.n_steps * (G + 1) * int(mut.shape[0]) * (4 * ((G + 31) // 32) + 2)
              for p, mut, kno, G in jobs)
    return {
        "workload": f"256 datasets (genes U[16,94], cells U[16,95]), 64 chains each; "
                    f"batched {steps_batched} steps, plain {steps_plain} steps per chain",
        "batched_steps_per_s": rate_b,
        "plain_steps_per_s": rate_p,
        "speedup": rate_b / rate_p,
        "batched_ms": ms,
        "best_ll_identical_to_plain": same,
        "best_ll_identical_to_plain_steps": steps_plain,
        "timed_batch_identical_to_plain": {"datasets": sample, "steps": steps_batched,
                                           "identical": bool(same_timed)},
        "plain_rate_note": f"plain launches timed at {steps_plain} steps per chain (kernel "
                           "time; it grows linearly with steps), the batch at "
                           f"{steps_batched}",
        "alg_int_tops": ops / (ms / 1e3) / 1e12,
        "roofline_frac_nominal": ops / (ms / 1e3) / 1e12 / P_INT32_TOPS,
        "short_run": {"steps": steps_plain, "eager_ms": eager_short,
                      "graph_replay_ms": graph_short,
                      "graph_best_ll_identical_to_plain": same_graph},
        "autotuned": {"waves_per_group": cfg.waves_per_group,
                      "n_streams": cfg.n_streams, "order": cfg.order,
                      "two_wave_layouts": two_wave_codes(cfg.two_wave_layouts)},
    }


def served_stream_jobs(n_jobs, seed=5):
    return WL.served_stream(n_jobs, seed)


ADMISSION_NAMES = {B.ADMIT_ARRIVAL: "arrival order", B.ADMIT_LEAST_WORK: "least work first"}


def serving_bench(n_jobs, latencies_path=None, deadline_s=None, admission=None, aging_s=0.0):
    """Continuous batching on the first n_jobs of the served stream, in real
    time: each job is submitted at its arrival time; latency = completion -
    arrival. With deadline_s the run stops after that many seconds and
    reports the jobs completed by then (complete: false)."""
    jobs = served_stream_jobs(n_jobs)
    log(f"[bench] serving {n_jobs} jobs")
    lat = []
    arrival = {}
    records = []
    kw = {} if admission is None else {"admission": admission, "aging_s": aging_s}
    cfg = B.server_config(hbm_budget_bytes=4 << 30, n_streams=16, round_ms=SERVE_ROUND_MS, **kw)
    with B.Server(cfg) as srv:
        t0 = srv.now()
        nxt = 0
        done = 0
        complete = True
        while done < n_jobs:
            now = srv.now() - t0
            if deadline_s is not None and now > deadline_s:
                complete = False
                log(f"[bench] serving deadline {deadline_s:.0f} s reached with {done} jobs done")
                break
            while nxt < n_jobs and jobs[nxt][0] <= now:
                jid = srv.submit(*jobs[nxt][1])
                arrival[jid] = jobs[nxt][0]
                nxt += 1
            res, pend, ready = srv.step()
            for r in srv.poll():
                lat.append(r.done_s - t0 - arrival[r.id])
                records.append((r.id, arrival[r.id], r.admit_s - t0, r.done_s - t0, r.n_genes,
                                r.n_cells, r.n_chains, r.n_steps))
                done += 1
                if done % 250 == 0:
                    log(f"[bench] served {done}/{n_jobs} jobs at {srv.now() - t0:.1f} s "
                        f"({res} resident, {pend} pending)")
            if res == 0 and pend == 0 and nxt < n_jobs and done < n_jobs:
                time.sleep(max(0.0, jobs[nxt][0] - (srv.now() - t0)))
        end = srv.now() - t0
        chain_steps, busy, rounds = srv.stats()
    span = jobs[-1][0]
    offered = sum(p.n_chains * p.n_steps for _, (p, *_r) in jobs)
    lat = np.array(lat)
    if latencies_path:
        with open(latencies_path, "w") as f:
            f.write("job,arrival_s,admit_s,done_s,genes,cells,chains,steps\n")
            for rec in sorted(records):
                f.write(",".join(f"{x:.6f}" if isinstance(x, float) else str(x) for x in rec)
                        + "\n")
    return {
        "workload": f"first {n_jobs} jobs of the served stream (Poisson 50 jobs/s, seed 5; "
                    "genes U[16,127], cells U[16,128], chains U[8,256], steps log-uniform "
                    "[1e4,1e6])",
        "jobs": n_jobs,
        "jobs_completed": done,
        "complete": complete,
        "subset_of_10000_job_stream": n_jobs < 10000,
        "admission": ADMISSION_NAMES[cfg.admission],
        "aging_s": cfg.aging_s,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p90_ms": float(np.percentile(lat, 90) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "max_ms": float(lat.max() * 1e3),
        "aggregate_steps_per_s": chain_steps / end,
        "busy_steps_per_s": chain_steps / busy,
        "offered_steps_per_s": offered / span,
        "arrival_span_s": span,
        "makespan_s": end,
        "slices": rounds,
    }


def serve_vs_batch(n_jobs, cap, round_ms=None):
    """The first n_jobs of the served stream with steps capped at `cap`, all
    present at once: as one batch (each layout one launch, 16 streams) and
    through the server (submitted together), to separate the serving loop's
    own cost from the kernels'."""
    jobs = [j for _, j in WL.served_stream(n_jobs, max_steps=cap)]
    total = sum(p.n_chains * p.n_steps for p, *_ in jobs)
    cfg = B.batch_config(waves_per_group=1, n_streams=16)
    with B.Batch(jobs, cfg) as b:
        ms = b.run()
    out = {"jobs": n_jobs, "steps_cap": cap, "chain_steps": total,
           "batch_steps_per_s": total / (ms / 1e3)}
    kw = {"hbm_budget_bytes": 4 << 30, "n_streams": 16,
          "round_ms": SERVE_ROUND_MS if round_ms is None else round_ms}
    with B.Server(B.server_config(**kw)) as srv:
        t0 = srv.now()
        for j in jobs:
            srv.submit(*j)
        done = 0
        while done < n_jobs:
            srv.step()
            done += len(srv.poll())
        end = srv.now() - t0
        steps, busy, slices = srv.stats()
    out.update({"served_steps_per_s": steps / end, "served_slices": slices,
                "round_ms": round_ms})
    log(f"[bench] {n_jobs} jobs capped at {cap} steps: batch {out['batch_steps_per_s']:.3e}, "
        f"served {out['served_steps_per_s']:.3e} steps/s ({slices} slices)")
    return out


def oracle_check(p, mut, kno, steps, snap):
    """The first window of the timed run replayed on the CPU oracle: chains
    0..n-1 after `steps` steps (best score and step, final score, beta and
    tree), compared bit for bit."""
    from tests import _oracle as O

    (bs, st, fs, fb), am, dm = snap
    q = B.default_params(n_chains=p.n_chains, n_steps=steps, seed=p.seed,
                         alpha_mean=p.alpha_mean, beta_mean=p.beta_mean, beta_sd=p.beta_sd,
                         prob_beta_change=p.prob_beta_change)
    s = O.make_scorer(q.alpha_mean, q.beta_mean, q.beta_sd, mut, kno, G)
    t0 = time.perf_counter()
    ref = O.run_chains(q, s, count=len(bs))
    bad = 0
    for c, r in enumerate(ref):
        same = (st[c] == r.best_step
                and np.float32(bs[c]).tobytes() == np.float32(r.best_score).tobytes()
                and np.float32(fs[c]).tobytes() == np.float32(r.final_score).tobytes()
                and np.float32(fb[c]).tobytes() == np.float32(r.final_tree.beta).tobytes()
                and np.array_equal(am[c], O.rows(r.final_tree, "am"))
                and np.array_equal(dm[c], O.rows(r.final_tree, "dm")))
        bad += 0 if same else 1
    log(f"[bench] oracle check: {len(bs)} chains x {steps} steps, {bad} mismatches "
        f"({time.perf_counter() - t0:.1f} s)")
    return {"chains": len(bs), "steps": steps, "mismatches": bad,
            "compared": "best score and step, final score, beta and ancestor rows"}


def plain_64x63_bench(with_cpu):
    """The 63 x 64 workload (seed 1, 8 chains x 100,000 steps, -e 0) as
    specified, and the same data with 1,024 chains (a filled device)."""
    out = {}
    for label, chains, steps in (("8_chains", 8, 100_000), ("1024_chains", 1024, 20_000)):
        p, mut, kno, Gs = WL.plain_64x63(chains, steps)
        with B.Context(p, mut, kno, Gs) as ctx:
            t0 = time.perf_counter()
            ms = ctx.run()
            wall = time.perf_counter() - t0
            _, _, _, best = ctx.best()
        out[label] = {"chains": chains, "steps": steps, "kernel_ms": ms,
                      "steps_per_s": chains * steps / wall,
                      "kernel_steps_per_s": chains * steps / (ms / 1e3),
                      "alg_int_tops": 64 * 64 * (4 * 2 + 2) * chains * steps / (ms / 1e3) / 1e12,
                      "best_log_likelihood": best}
        log(f"[bench] 63x64 {label}: {out[label]['steps_per_s']:.3e} steps/s")
    out["kernel"] = chain_kernel_name(63, 64)
    out["fpga_ref"] = "bbSCITE64 on Stratix 10, 4,349.70 ksteps/s at 64x63 (paper Table 2)"
    if with_cpu:
        p, mut, kno, Gs = WL.plain_64x63()
        out["cpu_baseline"] = cpu_baseline(p, mut, kno, Gs, "63x64", seconds_budget=10.0)
    return out


def few_chains_bench():
    """8 chains (the plain launch with fewer chains than SIMDs, as the
    command line runs it): one wave per chain, against 8 waves per chain
    proposing consecutive steps at once (speculation) and against each chain's
    cells split over two waves; the automatic choice is speculation up to 96
    cells and the split above."""
    out = {}
    modes = (("one_wave", 1, 1), ("speculation_8", 8, 1), ("split_2", 1, 2))
    for G, C, steps in ((63, 64, 100_000), (94, 95, 50_000), (127, 128, 30_000)):
        m, _, _ = B.generate_dataset(2, G, C, 1e-4, 0.3, 0.15)
        mut, kno = B.encode_matrix(m)
        p = B.default_params(n_chains=8, n_steps=steps, seed=2, alpha_mean=1e-4, beta_mean=0.3,
                             beta_sd=0.05, prob_beta_change=0.1)
        row = {"chains": 8, "steps": steps}
        best = {}
        for name, spec, split in modes:
            with B.Context(p, mut, kno, G) as ctx:
                try:
                    ctx.set_speculation(spec)
                except B.BbsError:  # 8 waves' LDS over a workgroup's: 4
                    spec = 4
                    ctx.set_speculation(spec)
                    row["speculation_8_ran_with"] = 4
                ctx.set_chain_split(split)
                ms = ctx.run()
                best[name] = np.float32(ctx.best()[3]).tobytes()
            row[f"{name}_steps_per_s"] = 8 * steps / (ms / 1e3)
        row["best_ll_identical"] = len(set(best.values())) == 1
        out[f"{G}x{C}"] = row
        log(f"[bench] {G}x{C} 8 chains: " + ", ".join(
            f"{n} {row[n + '_steps_per_s']:.3e}" for n, _, _ in modes) + " steps/s")
    return out


def full_device_split_bench(steps=20_000):
    """Waves per chain with every SIMD busy: one full round of chains (four
    per CU, or three where the split form's LDS allows no more) with one
    wave per chain, against each chain's cells split over two and four
    waves (two and four waves per SIMD)."""
    out = []
    for G, C, chains in ((80, 82, 1024), (94, 95, 768), (64, 96, 1024), (90, 70, 1024)):
        m, _, _ = B.generate_dataset(7, G, C, 1e-4, 0.3, 0.15)
        mut, kno = B.encode_matrix(m)
        p = B.default_params(n_chains=chains, n_steps=steps, seed=7, alpha_mean=1e-4,
                             beta_mean=0.3, beta_sd=0.05, prob_beta_change=0.1)
        row = {"genes": G, "cells": C, "chains": chains, "steps": steps}
        best = set()
        for split in (1, 2, 4):
            with B.Context(p, mut, kno, G) as ctx:
                ctx.set_speculation(1)
                ctx.set_chain_split(split)
                ctx.run()  # first launch of the instance
                ms = ctx.run()
                best.add(np.float32(ctx.best()[3]).tobytes())
            row[f"split_{split}_ms"] = round(ms, 2)
            row[f"split_{split}_steps_per_s"] = chains * steps / (ms / 1e3)
        row["best_ll_identical"] = len(best) == 1
        out.append(row)
        log(f"[bench] {G}x{C} {chains} chains: " + ", ".join(
            f"split {k} {row[f'split_{k}_ms']:.1f} ms" for k in (1, 2, 4)))
    return out


def scoring_modes_bench(steps=10_000):
    """Plain launches over 96 cells, 1,024 and 2,048 chains: full rescoring,
    the half table (cells 0..63), the whole table and the automatic choice,
    per launch of `steps` steps (the second of two runs)."""
    out = []
    for G, C, chains in ((127, 128, 1024), (100, 110, 1024), (120, 100, 1024), (127, 128, 2048)):
        m, _, _ = B.generate_dataset(7, G, C, 1e-4, 0.3, 0.15)
        mut, kno = B.encode_matrix(m)
        p = B.default_params(n_chains=chains, n_steps=steps, seed=7, alpha_mean=1e-4,
                             beta_mean=0.3, beta_sd=0.05, prob_beta_change=0.1)
        row = {"genes": G, "cells": C, "chains": chains, "steps": steps}
        best = set()
        for mode in ("full", "half", "incremental", "auto"):
            with B.Context(p, mut, kno, G) as ctx:
                ctx.set_scoring_mode(mode)
                ctx.run()
                row[f"{mode}_ms"] = round(ctx.run(), 2)
                best.add(np.float32(ctx.best()[3]).tobytes())
        row["best_ll_identical"] = len(best) == 1
        out.append(row)
        log(f"[bench] {G}x{C} {chains} chains: " + ", ".join(
            f"{k} {row[k + '_ms']:.1f} ms" for k in ("full", "half", "incremental", "auto")))
    return out


def layout_code(genes, cells):
    """Register layout (node slots, cell slots, gene words) as the library's
    shape_of picks it, as the code NS*100 + CS*10 + W."""
    ns = 1 if genes + 1 <= 64 else 2
    cs = 1 if cells <= 64 else (3 if cells <= 96 else 2)
    w = max(1, (genes + 31) // 32)
    if ns == 2:
        w = max(w, 2)
    return ns * 100 + cs * 10 + w


def two_wave_codes(mask):
    """Layout codes (NS*100 + CS*10 + W) whose bit is set in a
    two_wave_layouts mask."""
    return [ns * 100 + cs * 10 + w for ns in (1, 2) for cs in (1, 2, 3) for w in (1, 2, 3, 4)
            if mask & B.layout_bit(ns, cs, w)]


def layout_profile(steps, only=None, wpg=1, reps=2, served=0, streams=2):
    """The 256-dataset mix split by register layout: each layout's datasets
    (all 64 chains each) as one batch (its one or two launches on their own
    streams), timed alone with either
    register budget (one or two waves per SIMD), then the whole mix with each
    layout's faster budget. Per layout: items, kernel time, chain steps/s and
    the reference's int work N C (4W + 2) per step as a fraction of the
    nominal int32 VALU peak. served > 0: the first `served` jobs of the
    served stream instead, steps capped at `steps`."""
    jobs = ([j for _, j in WL.served_stream(served, max_steps=steps)] if served
            else WL.batch_mix(steps))
    groups = {}
    for j, (p, mut, kno, Gj) in enumerate(jobs):
        groups.setdefault(layout_code(Gj, mut.shape[0]), []).append(j)
    rows = []
    mask = 0
    for code in sorted(groups):
        if only is not None and code != only:
            continue
        sub = [jobs[j] for j in groups[code]]
        bit = B.layout_bit(code // 100, code // 10 % 10, code % 10)
        ops = sum(p.n_chains * p.n_steps * (Gj + 1) * int(mut.shape[0]) * (4 * ((Gj + 31) // 32) + 2)
                  for p, mut, kno, Gj in sub)
        row = {"layout": code, "datasets": len(sub), "items": sum(p.n_chains for p, *_ in sub),
               "mean_nodes": float(np.mean([Gj + 1 for *_, Gj in sub])),
               "mean_cells": float(np.mean([m.shape[0] for _, m, _, _ in sub])),
               "share_of_alg_work": None}
        for key, m, sc in (("wps1", 0, "auto"), ("wps2", bit, "auto"), ("full", 0, "full")):
            cfg = B.batch_config(waves_per_group=wpg, n_streams=streams, two_wave_layouts=m,
                                 scoring=sc)
            with B.Batch(sub, cfg) as b:
                ms = min(b.run() for _ in range(reps))
                total = b.total_steps
            row[key] = {"ms": ms, "steps_per_s": total / (ms / 1e3),
                        "alg_int_tops": ops / (ms / 1e3) / 1e12,
                        "frac_nominal": ops / (ms / 1e3) / 1e12 / P_INT32_TOPS}
            log(f"[layout] {code} ({key}): {row['datasets']} datasets, "
                f"{ms:.1f} ms, {total / (ms / 1e3):.3e} steps/s, "
                f"frac {row[key]['frac_nominal']:.3f}")
        if row["wps2"]["ms"] < row["wps1"]["ms"]:
            mask |= bit
        rows.append((row, ops))
    all_ops = sum(o for _, o in rows)
    for row, o in rows:
        row["share_of_alg_work"] = o / all_ops
    out = {"steps": steps, "waves_per_group": wpg, "streams": streams,
           "layouts": [r for r, _ in rows],
           "faster_two_wave_layouts": two_wave_codes(mask)}
    if only is None:
        cfg = B.batch_config(waves_per_group=wpg, n_streams=streams, two_wave_layouts=mask)
        with B.Batch(jobs, cfg) as b:
            ms = min(b.run() for _ in range(reps))
            total = b.total_steps
        out["whole_mix"] = {"ms": ms, "steps_per_s": total / (ms / 1e3),
                            "alg_int_tops": all_ops / (ms / 1e3) / 1e12,
                            "frac_nominal": all_ops / (ms / 1e3) / 1e12 / P_INT32_TOPS}
        log(f"[layout] whole mix, {streams} streams: {ms:.1f} ms, "
            f"frac {out['whole_mix']['frac_nominal']:.3f}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=9)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--chain-steps", type=int, default=100_000)
    ap.add_argument("--chains", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-score", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the 63 x 64 leg")
    ap.add_argument("--batch-steps", type=int, default=200_000)
    ap.add_argument("--serve-jobs", type=int, default=200,
                    help="jobs of the served stream to run (0: skip serving)")
    ap.add_argument("--layout-profile", action="store_true",
                    help="only time the 256-dataset mix layout by layout (one JSON line)")
    ap.add_argument("--layout", type=int, default=None,
                    help="with --layout-profile: only this layout code (NS*100+CS*10+W)")
    ap.add_argument("--layout-steps", type=int, default=20_000)
    ap.add_argument("--layout-served", type=int, default=0, metavar="N",
                    help="with --layout-profile: the first N served jobs instead of the "
                         "256-dataset mix")
    ap.add_argument("--layout-streams", type=int, default=2)
    ap.add_argument("--layout-wpg", type=int, default=1,
                    help="with --layout-profile: waves per workgroup")
    ap.add_argument("--serve-only", action="store_true",
                    help="only run the served stream (--serve-jobs jobs), one JSON line")
    ap.add_argument("--serve-latencies", default=None, metavar="PATH",
                    help="with the served stream: write each job's arrival, admission and "
                         "completion times (s) and size to PATH as CSV")
    ap.add_argument("--batch-shapes", default=None, metavar="WPG:STREAMS,...",
                    help="only time the 256-dataset batch at --batch-steps steps with each "
                         "given launch shape (waves per group : streams), no autotuning")
    ap.add_argument("--serve-admission", choices=("arrival", "least-work"), default=None,
                    help="server admission order (default: the library's)")
    ap.add_argument("--serve-aging", type=float, default=0.0, metavar="S",
                    help="with --serve-admission least-work: aging time constant (s)")
    ap.add_argument("--serve-deadline", type=float, default=None, metavar="S",
                    help="with --serve-only: stop after S seconds and report what completed")
    ap.add_argument("--serve-vs-batch", type=int, default=0, metavar="N",
                    help="only compare the first N served jobs (steps capped at "
                         "--serve-cap) run as one batch and through the server")
    ap.add_argument("--serve-cap", type=int, default=20_000)
    ap.add_argument("--round-ms", type=float, default=None,
                    help="server slice length (default: the library's)")
    ap.add_argument("--scoring-modes", action="store_true",
                    help="only time the scoring modes over 96 cells (one JSON line)")
    ap.add_argument("--split-full-device", action="store_true",
                    help="only time waves per chain with every SIMD busy (one JSON line)")
    ap.add_argument("--valu-peak-only", action="store_true",
                    help="only run the int32 VALU probe (for its counters)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch

        n = a.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        log(f"[bench] copied {a.copy_calibration} MiB")
        return
    if a.scoring_modes:
        print(json.dumps({"scoring_modes": scoring_modes_bench()}), flush=True)
        return
    if a.split_full_device:
        print(json.dumps({"split_full_device": full_device_split_bench()}), flush=True)
        return
    if a.valu_peak_only:
        m0, _, _ = B.generate_dataset(1, 4, 4, 1e-4, 0.2, 0.1)
        mu0, kn0 = B.encode_matrix(m0)
        with B.Context(B.default_params(n_chains=1, n_steps=1), mu0, kn0, 4) as c0:
            ops, insts = c0.measure_valu_peak()
        print(json.dumps({"int32_ops_per_s": ops, "valu_wave_insts_per_s": insts}), flush=True)
        return
    if a.batch_shapes:
        jobs = batch_mix_jobs(a.batch_steps)
        res = []
        for shape in a.batch_shapes.split(","):
            wpg, nst = (int(x) for x in shape.split(":"))
            with B.Batch(jobs, B.batch_config(waves_per_group=wpg, n_streams=nst)) as b:
                ms = b.run()
                total = b.total_steps
            res.append({"waves_per_group": wpg, "n_streams": nst, "ms": ms,
                        "steps_per_s": total / (ms / 1e3)})
            log(f"[bench] batch wpg={wpg} streams={nst}: {ms:.1f} ms, "
                f"{total / (ms / 1e3):.3e} steps/s")
        print(json.dumps({"steps": a.batch_steps, "shapes": res}), flush=True)
        return
    if a.serve_vs_batch:
        print(json.dumps(serve_vs_batch(a.serve_vs_batch, a.serve_cap, a.round_ms)), flush=True)
        return
    if a.serve_only:
        adm = {None: None, "arrival": B.ADMIT_ARRIVAL,
               "least-work": B.ADMIT_LEAST_WORK}[a.serve_admission]
        print(json.dumps(serving_bench(a.serve_jobs, a.serve_latencies, a.serve_deadline,
                                       adm, a.serve_aging)), flush=True)
        return
    if a.layout_profile:
        print(json.dumps(layout_profile(a.layout_steps, a.layout, wpg=a.layout_wpg,
                                        served=a.layout_served, streams=a.layout_streams)),
              flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(chains shard across GPUs with one final argmax exchange)",
              file=sys.stderr)
        sys.exit(2)

    # calibrate the int32 VALU probe on this device (reported beside the
    # nominal peak, which is the roofline's denominator)
    m0, _, _ = B.generate_dataset(1, 4, 4, 1e-4, 0.2, 0.1)
    mu0, kn0 = B.encode_matrix(m0)
    with B.Context(B.default_params(n_chains=1, n_steps=1), mu0, kn0, 4) as c0:
        peak_ops, peak_insts = c0.measure_valu_peak()
    peak = P_INT32_TOPS
    log(f"[bench] int32 VALU probe {peak_ops / 1e12:.2f} TOP/s "
        f"({peak_insts / 1e9:.1f} G wave-instr/s); nominal peak {peak:.2f}")

    S = a.chain_steps
    total = (a.warmup + a.steps) * S
    p, mut, kno = make_context(a.chains, total)
    log(f"[bench] 94x95, {a.chains} chains, {a.warmup}+{a.steps} windows of {S} steps")
    ctx = B.Context(p, mut, kno, G)
    ctx.begin()
    snap = None
    n_check = min(16, a.chains)
    for i in range(a.warmup):
        ms = ctx.advance(S)
        log(f"[bench] warmup window {i}: {ms:.1f} ms")
        if i == 0:
            # the chains after the first window, for the oracle check below
            r = ctx.chain_results()
            st = ctx.chain_states()
            snap = ([x[:n_check].copy() for x in r], st[0][:n_check].copy(),
                    st[1][:n_check].copy())
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(a.steps):
        ms = ctx.advance(S)  # blocks until the window's kernel has finished
        kernel_ms.append(ms)
        log(f"[bench] window {i}: {ms:.1f} ms")
    wall = time.perf_counter() - t0
    steps_done = a.steps * S * a.chains
    value = steps_done / wall
    k_rate = steps_done / (sum(kernel_ms) / 1e3)
    achieved = ALG_INT_OPS * k_rate / 1e12
    _, _, best_beta, best_score = ctx.best()
    ctx.close()
    check = None if a.no_cpu or snap is None else oracle_check(p, mut, kno, S, snap)

    small = None if a.no_small else plain_64x63_bench(not a.no_cpu)
    few = None if a.no_small else few_chains_bench()
    score = None if a.no_score else scoring_bench(peak=peak)
    if score:
        log(f"[bench] scoring: {score['trees_per_s']:.3e} trees/s")
    ll_gpu = score.pop("_ll_gpu") if score else None
    batched = None if a.no_batch else batched_bench(a.batch_steps)
    serving = serving_bench(a.serve_jobs) if a.serve_jobs > 0 else None
    serving_lw = (serving_bench(a.serve_jobs, admission=B.ADMIT_LEAST_WORK)
                  if a.serve_jobs > 0 else None)
    if serving:
        log(f"[bench] served {serving['jobs']} jobs: p50 {serving['p50_ms']:.0f} ms, "
            f"p99 {serving['p99_ms']:.0f} ms, {serving['aggregate_steps_per_s']:.3e} steps/s "
            f"(offered {serving['offered_steps_per_s']:.3e})")
    if serving_lw:
        log(f"[bench] served, least work first: p50 {serving_lw['p50_ms']:.0f} ms, "
            f"p99 {serving_lw['p99_ms']:.0f} ms, "
            f"{serving_lw['aggregate_steps_per_s']:.3e} steps/s")
    if batched:
        log(f"[bench] batched {batched['batched_steps_per_s']:.3e} steps/s, plain "
            f"{batched['plain_steps_per_s']:.3e} steps/s, x{batched['speedup']:.2f}")
    cpu = None if a.no_cpu else cpu_baseline(p, mut, kno, G, "94x95")
    if score is not None and not a.no_cpu:
        score["cpu_baseline"] = cpu_scoring_baseline(ll_gpu)

    out = {
        "metric": "chain steps/s at 94x95, batched persistent vs plain launch; p99 job latency",
        "value": value,
        "unit": "chain steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / (FPGA_KSTEPS * 1e3),
        "baseline_ref": "bbSCITE96 on Stratix 10, 2,938.25 ksteps/s at 95x94 (paper Table 2)",
        "dtype": "u32+f32",
        "data": "synthetic",
        "config": {
            "workload": "plain chain kernel, 94 genes x 95 cells, 1024 chains, "
                        f"{a.warmup + a.steps} x {S} steps per chain",
            "chains": a.chains,
            "chain_steps_per_bench_step": S,
            "genes": G,
            "cells": C,
        },
        "roofline": {
            "bound": "valu-int32",
            "achieved": achieved,
            "peak": peak,
            "unit": "TOP/s",
            "frac": achieved / peak,
            "peak_nominal": P_INT32_TOPS,
            "peak_probe_measured": peak_ops / 1e12,
            "frac_of_probe_peak": achieved / (peak_ops / 1e12),
            "traffic": TRAFFIC_100K_WINDOW if S == 100_000 and a.chains == 1024 else None,
            "alg_int_ops_per_step": ALG_INT_OPS,
            "kernel": chain_kernel_name(G, C),
        },
        "kernel_steps_per_s": k_rate,
        "best_log_likelihood": best_score,
        "best_beta": best_beta,
        "oracle_check": check,
        "plain_64x63": small,
        "few_chains": few,
        "scoring": score,
        "batched": batched,
        "serving_p99_ms": serving["p99_ms"] if serving else None,
        "serving": serving,
        "serving_least_work_first": serving_lw,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
