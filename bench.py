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
(HIP-event kernel time on the launch stream), the bulk-scoring throughput
(2^20 random trees against a 94 x 95 matrix) and the CPU oracle timed on
the host cores of the same machine.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The server runs one lane (stream) per register layout, on up to 16
# streams; HIP runs them at once only on as many hardware queues as the
# process asks for (read when HIP initialises, 4 by default).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

from bbscite_amd import _capi as B  # noqa: E402

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
    m, _, _ = B.generate_dataset(2, G, C, 1e-4, 0.3, 0.15)
    mut, kno = B.encode_matrix(m)
    p = B.default_params(
        n_chains=n_chains, n_steps=n_steps, seed=2, alpha_mean=1e-4, beta_mean=0.3,
        beta_sd=0.05, prob_beta_change=0.1,
    )
    return p, mut, kno


def cpu_baseline(p, mut, kno, seconds_budget=20.0):
    """The CPU oracle's chain loop on the host cores (bounded sample)."""
    try:
        from tests import _oracle as O
    except Exception as e:  # pragma: no cover
        log("cpu baseline unavailable:", e)
        return None
    cores = min(16, len(os.sched_getaffinity(0)))
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    # calibrate on one core, then size the sample to ~seconds_budget of CPU work
    r1, _ = O.bench_chains(p, s, 1, 200, 1)
    steps_total = max(cores * 200, int(r1 * seconds_budget))
    n_chains = cores * 4
    n_steps = max(100, steps_total // n_chains)
    rate, _ = O.bench_chains(p, s, n_chains, n_steps, cores)
    return {
        "value": rate,
        "unit": "chain steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"oracle chain loop at 94x95, {n_chains} chains x {n_steps} steps "
                  f"(one chain per thread; 1-core rate {r1:.0f} steps/s)",
        "one_core_value": r1,
    }


def scoring_matrix():
    """One 94 x 95 matrix of iid entries, P(1) = 0.3, P(3) = 0.1, else 0
    (numpy generator, seed 3), encoded as per-cell rows."""
    rng = np.random.default_rng(3)
    u = rng.random((G, C))
    m = np.zeros((G, C), np.int8)
    m[u < 0.3] = 1
    m[(u >= 0.3) & (u < 0.4)] = 3
    return B.encode_matrix(m)


def scoring_bench(reps=5, peak=P_INT32_TOPS):
    """2^20 uniform random trees (seed 3) against one 94 x 95 iid matrix."""
    mut, kno = scoring_matrix()
    p = B.default_params(n_chains=1, n_steps=1)
    n_trees = 1 << 20
    with B.Context(p, mut, kno, G) as ctx:
        ms, _ = ctx.bench_score_random(3, n_trees, reps, with_prior=False, n_head=0)
        ms_v, _ = ctx.bench_score_random(3, n_trees, 3, with_prior=False, n_head=0,
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
                        "frac": ALG_INT_OPS * v_trees_s / 1e12 / peak},
    }


def cpu_scoring_baseline(seconds_budget=5.0):
    """The oracle's CPUTreeScorer-structured scorer (node-outer, cell-inner,
    popcounts over 64-bit words, TreeScorer semantics, no prior) over the
    first trees of the same seed-3 stream, parallel over trees on the host
    cores (bounded sample)."""
    try:
        from tests import _oracle as O
    except Exception as e:  # pragma: no cover
        log("cpu scoring baseline unavailable:", e)
        return None
    cores = min(16, len(os.sched_getaffinity(0)))
    mut, kno = scoring_matrix()
    p = B.default_params(n_chains=1, n_steps=1)
    s = O.make_scorer(p.alpha_mean, p.beta_mean, p.beta_sd, mut, kno, G)
    r1, _ = O.bench_score(s, B.random_trees(3, G, 256), p.beta_mean, 1)
    n = int(min(1 << 20, max(cores * 256, r1 * cores * seconds_budget)))
    rate, _ = O.bench_score(s, B.random_trees(3, G, n), p.beta_mean, cores)
    return {
        "value": rate,
        "unit": "trees/s",
        "cores": cores,
        "kind": "port",
        "sample": f"first {n} trees of the seed-3 stream at 94x95 (1-core rate {r1:.0f} trees/s)",
        "one_core_value": r1,
        "counted_bits_per_s": rate * 4 * N ** 3,
    }


def config4_jobs(steps, chains=64, n_datasets=256):
    """256 datasets, dataset i from seed 4 + i with genes ~ U[16, 94] and
    cells ~ U[16, 95] (sizes drawn from a seeded numpy generator, seed 4),
    noise as the 63 x 64 configuration (-ad 0.2, -fd 1e-4, 10 % missing)."""
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


def batched_bench(steps_batched=200_000, steps_plain=2_000, probe_steps=2_000):
    """Config 4: the 256-dataset uneven mix, plain launches vs the batched
    per-layout launches. Plain launches run one dataset each, serially. At
    the plain step count the batch is also timed as an eager run and as a
    replay of its hipGraph capture."""
    jobs_plain = config4_jobs(steps_plain)
    t_plain = 0.0
    bests_plain = []
    for p, mut, kno, G in jobs_plain:
        with B.Context(p, mut, kno, G) as ctx:
            t_plain += ctx.run() / 1e3
            bests_plain.append(ctx.best()[3])
    plain_steps = sum(p.n_chains * p.n_steps for p, *_ in jobs_plain)
    cfg, probe_ms = B.batch_autotune(jobs_plain, probe_steps)
    log(f"[bench] batch autotune: wpg={cfg.waves_per_group} streams={cfg.n_streams} "
        f"order={cfg.order} ({probe_ms:.1f} ms at {probe_steps} steps)")
    with B.Batch(jobs_plain, cfg) as b:
        b.run()
        same = all(np.float32(b.best(j)[3]).tobytes() == np.float32(bests_plain[j]).tobytes()
                   for j in range(len(jobs_plain)))
        eager_short = min(b.run() for _ in range(3))
    # the same short batch captured once in a hipGraph and replayed
    gcfg = B.batch_config(cfg.waves_per_group, cfg.n_streams, cfg.order, graph=True)
    with B.Batch(jobs_plain, gcfg) as g:
        g.run()
        graph_short = min(g.run() for _ in range(3))
        same_graph = all(np.float32(g.best(j)[3]).tobytes() == np.float32(bests_plain[j]).tobytes()
                         for j in range(len(jobs_plain)))
    jobs = config4_jobs(steps_batched)
    with B.Batch(jobs, cfg) as b:
        ms = b.run()
        total = b.total_steps
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
        "alg_int_tops": ops / (ms / 1e3) / 1e12,
        "roofline_frac_nominal": ops / (ms / 1e3) / 1e12 / P_INT32_TOPS,
        "short_run": {"steps": steps_plain, "eager_ms": eager_short,
                      "graph_replay_ms": graph_short,
                      "graph_best_ll_identical_to_plain": same_graph},
        "autotuned": {"waves_per_group": cfg.waves_per_group,
                      "n_streams": cfg.n_streams, "order": cfg.order},
    }


def config5_jobs(n_jobs, seed=5):
    """The served stream: Poisson arrivals at 50 jobs/s; genes U[16,127],
    cells U[16,128], chains U[8,256], steps log-uniform in [1e4, 1e6]
    (numpy generator, seed 5); data as the 63 x 64 configuration (-ad 0.2,
    -fd 1e-4, 10 % missing) from seed 1000 + i."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n_jobs):
        t += rng.exponential(1.0 / 50.0)
        G = int(rng.integers(16, 128))
        Cc = int(rng.integers(16, 129))
        R = int(rng.integers(8, 257))
        S = int(10 ** rng.uniform(4.0, 6.0))
        m, _, _ = B.generate_dataset(1000 + i, G, Cc, 1e-4, 0.2, 0.1)
        mut, kno = B.encode_matrix(m)
        p = B.default_params(n_chains=R, n_steps=S, seed=1000 + i, alpha_mean=1e-4,
                             beta_mean=0.2)
        jobs.append((t, (p, mut, kno, G)))
    return jobs


def serving_bench(n_jobs):
    """Continuous batching on the first n_jobs of the served stream, in real
    time: each job is submitted at its arrival time; latency = completion -
    arrival."""
    jobs = config5_jobs(n_jobs)
    lat = []
    arrival = {}
    with B.Server(B.server_config(hbm_budget_bytes=4 << 30, n_streams=16)) as srv:
        t0 = srv.now()
        nxt = 0
        done = 0
        while done < n_jobs:
            now = srv.now() - t0
            while nxt < n_jobs and jobs[nxt][0] <= now:
                jid = srv.submit(*jobs[nxt][1])
                arrival[jid] = jobs[nxt][0]
                nxt += 1
            res, pend, ready = srv.step()
            for r in srv.poll():
                lat.append(r.done_s - t0 - arrival[r.id])
                done += 1
            if res == 0 and pend == 0 and nxt < n_jobs:
                time.sleep(max(0.0, jobs[nxt][0] - (srv.now() - t0)))
        end = srv.now() - t0
        chain_steps, busy, rounds = srv.stats()
    span = jobs[-1][0]
    offered = sum(p.n_chains * p.n_steps for _, (p, *_r) in jobs)
    lat = np.array(lat)
    return {
        "workload": f"first {n_jobs} jobs of the served stream (Poisson 50 jobs/s, seed 5; "
                    "genes U[16,127], cells U[16,128], chains U[8,256], steps log-uniform "
                    "[1e4,1e6])",
        "jobs": n_jobs,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "max_ms": float(lat.max() * 1e3),
        "aggregate_steps_per_s": chain_steps / end,
        "busy_steps_per_s": chain_steps / busy,
        "offered_steps_per_s": offered / span,
        "arrival_span_s": span,
        "makespan_s": end,
        "slices": rounds,
    }


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
    ap.add_argument("--batch-steps", type=int, default=200_000)
    ap.add_argument("--serve-jobs", type=int, default=200,
                    help="jobs of the served stream to run (0: skip serving)")
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
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported "
              "(chains shard across GPUs with one final argmax exchange)",
              file=sys.stderr)
        sys.exit(2)

    # calibrate the int32 VALU roofline on this device (use it if above the
    # nominal 256 CUs x 64 lanes x 2.4 GHz)
    m0, _, _ = B.generate_dataset(1, 4, 4, 1e-4, 0.2, 0.1)
    mu0, kn0 = B.encode_matrix(m0)
    with B.Context(B.default_params(n_chains=1, n_steps=1), mu0, kn0, 4) as c0:
        peak_ops, peak_insts = c0.measure_valu_peak()
    peak = max(P_INT32_TOPS, peak_ops / 1e12)
    log(f"[bench] measured int32 VALU peak {peak_ops / 1e12:.2f} TOP/s "
        f"({peak_insts / 1e9:.1f} G wave-instr/s); using {peak:.2f}")

    S = a.chain_steps
    total = (a.warmup + a.steps) * S
    p, mut, kno = make_context(a.chains, total)
    log(f"[bench] 94x95, {a.chains} chains, {a.warmup}+{a.steps} windows of {S} steps")
    ctx = B.Context(p, mut, kno, G)
    ctx.begin()
    for i in range(a.warmup):
        ms = ctx.advance(S)
        log(f"[bench] warmup window {i}: {ms:.1f} ms")
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

    score = None if a.no_score else scoring_bench(peak=peak)
    if score:
        log(f"[bench] scoring: {score['trees_per_s']:.3e} trees/s")
    batched = None if a.no_batch else batched_bench(a.batch_steps)
    serving = serving_bench(a.serve_jobs) if a.serve_jobs > 0 else None
    if serving:
        log(f"[bench] served {serving['jobs']} jobs: p50 {serving['p50_ms']:.0f} ms, "
            f"p99 {serving['p99_ms']:.0f} ms, {serving['aggregate_steps_per_s']:.3e} steps/s "
            f"(offered {serving['offered_steps_per_s']:.3e})")
    if batched:
        log(f"[bench] batched {batched['batched_steps_per_s']:.3e} steps/s, plain "
            f"{batched['plain_steps_per_s']:.3e} steps/s, x{batched['speedup']:.2f}")
    cpu = None if a.no_cpu else cpu_baseline(p, mut, kno)
    if score is not None and not a.no_cpu:
        score["cpu_baseline"] = cpu_scoring_baseline()

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
            "peak_measured": peak_ops / 1e12,
            "traffic": TRAFFIC_100K_WINDOW if S == 100_000 and a.chains == 1024 else None,
            "alg_int_ops_per_step": ALG_INT_OPS,
            "kernel": chain_kernel_name(G, C),
        },
        "kernel_steps_per_s": k_rate,
        "best_log_likelihood": best_score,
        "best_beta": best_beta,
        "scoring": score,
        "batched": batched,
        "serving_p99_ms": serving["p99_ms"] if serving else None,
        "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
