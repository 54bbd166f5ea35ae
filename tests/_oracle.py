"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker of the HIP path.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "build", "liboracle.so")

from bbscite_amd._capi import Params, StepTrace, TRACE_DTYPE  # noqa: E402

MAX_NODES = 128


class Row(C.Structure):
    _fields_ = [("w", C.c_uint64 * 2)]


class Tree(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_uint32),
        ("am", Row * MAX_NODES),
        ("dm", Row * MAX_NODES),
        ("beta", C.c_float),
    ]


class Mod(C.Structure):
    _fields_ = [
        ("move_type", C.c_uint32),
        ("v", C.c_uint32),
        ("w", C.c_uint32),
        ("parent_of_v", C.c_uint32),
        ("parent_of_w", C.c_uint32),
        ("descendant_of_v", C.c_uint32),
        ("nondescendant_of_v", C.c_uint32),
        ("new_beta", C.c_float),
    ]


class Scorer(C.Structure):
    _fields_ = [
        ("n_cells", C.c_uint32),
        ("n_genes", C.c_uint32),
        ("is_mutated", Row * 128),
        ("is_known", Row * 128),
        ("L00", C.c_float),
        ("L10", C.c_float),
        ("bpriora", C.c_float),
        ("bpriorb", C.c_float),
        ("lg_ab", C.c_float),
        ("lg_a", C.c_float),
        ("lg_b", C.c_float),
        ("prior_fallback", C.c_int),
    ]


class RawMove(C.Structure):
    _fields_ = [
        ("move_type", C.c_uint32),
        ("raw_v", C.c_uint32),
        ("raw_w", C.c_uint32),
        ("raw_descendant_of_v", C.c_float),
        ("raw_nondescendant_of_v", C.c_float),
        ("beta_jump", C.c_float),
        ("acceptance_level", C.c_float),
    ]


class MoveProbs(C.Structure):
    _fields_ = [
        ("t_beta", C.c_float),
        ("t_prune", C.c_float),
        ("t_swap", C.c_float),
        ("beta_jump_sd", C.c_float),
    ]


class ChainResult(C.Structure):
    _fields_ = [
        ("best_score", C.c_float),
        ("best_step", C.c_uint32),
        ("best_tree", Tree),
        ("final_score", C.c_float),
        ("final_tree", Tree),
    ]


_P = C.c_void_p
_u32 = C.c_uint32
_SIGS = {
    "orc_pruefer_to_parent": (C.c_int, [_P, _u32, _P]),
    "orc_parent_to_tree": (None, [_P, _u32, C.c_float, C.POINTER(Tree)]),
    "orc_is_parent": (C.c_int, [C.POINTER(Tree), _u32, _u32]),
    "orc_get_parent": (_u32, [C.POINTER(Tree), _u32]),
    "orc_is_ancestor": (C.c_int, [C.POINTER(Tree), _u32, _u32]),
    "orc_n_descendants": (_u32, [C.POINTER(Tree), _u32]),
    "orc_n_ancestors": (_u32, [C.POINTER(Tree), _u32]),
    "orc_select": (_u32, [C.POINTER(Tree), _u32, _u32, C.c_int]),
    "orc_realize": (Mod, [C.POINTER(Tree), C.POINTER(RawMove)]),
    "orc_update": (None, [C.POINTER(Tree), C.POINTER(Mod), C.POINTER(Tree)]),
    "orc_to_graphviz": (C.c_int64, [C.POINTER(Tree), C.c_char_p, C.c_uint64]),
    "orc_to_newick": (C.c_int64, [C.POINTER(Tree), C.c_char_p, C.c_uint64]),
    "orc_scorer_init": (
        None,
        [C.POINTER(Scorer), C.c_float, C.c_float, C.c_float, _u32, _u32, _P, _P],
    ),
    "orc_logscore_beta": (C.c_float, [C.POINTER(Scorer), C.c_float]),
    "orc_logscore_tree": (C.c_float, [C.POINTER(Scorer), C.POINTER(Tree), C.c_int, _P, _P]),
    "orc_move_probs": (MoveProbs, [C.POINTER(Params)]),
    "orc_initial_tree": (None, [C.POINTER(Params), _u32, _u32, _u32, C.POINTER(Tree)]),
    "orc_run_chain": (
        None,
        [
            C.POINTER(Params),
            C.POINTER(Scorer),
            _u32,
            _u32,
            C.c_int,
            C.POINTER(Tree),
            C.c_float,
            C.POINTER(ChainResult),
            _P,
        ],
    ),
    "orc_run_chain_proposals": (
        None,
        [
            C.POINTER(Params),
            C.POINTER(Scorer),
            _u32,
            _u32,
            C.c_int,
            C.POINTER(Tree),
            C.c_float,
            C.POINTER(ChainResult),
            _P,
            _P,
            _P,
        ],
    ),
    "orc_run": (
        None,
        [
            C.POINTER(Params),
            C.POINTER(Scorer),
            _u32,
            C.POINTER(Tree),
            C.POINTER(C.c_float),
            C.POINTER(_u32),
            C.POINTER(_u32),
        ],
    ),
    "orc_run_chains": (
        None,
        [C.POINTER(Params), C.POINTER(Scorer), _u32, _u32, _u32, C.c_int, _P],
    ),
    "orc_philox": (None, [_P, _P, _P]),
    "orc_draw_move": (RawMove, [C.c_uint64, _u32, _u32, _u32, _u32, MoveProbs]),
    "orc_normal": (C.c_double, [_u32, _u32]),
    "orc_bench_chains": (
        C.c_double,
        [C.POINTER(Params), C.POINTER(Scorer), _u32, _u32, C.c_int, C.POINTER(C.c_float)],
    ),
    "orc_bench_score": (
        C.c_double,
        [C.POINTER(Scorer), _P, _u32, C.c_float, C.c_int, _P],
    ),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def use_library(path):
    """Bind a different build of the same oracle (bench.py's CPU legs load
    the host-tuned -O3 -march=native build this way)."""
    global _lib, LIB_PATH
    _lib = None
    LIB_PATH = path
    return lib()


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ------------------------------------------------------------ conversions


def tree_from_parent(parent, beta=0.42):
    parent = np.ascontiguousarray(parent, np.uint32)
    t = Tree()
    lib().orc_parent_to_tree(_ptr(parent), len(parent), beta, C.byref(t))
    return t


def tree_from_rows(am, dm, beta):
    am = np.asarray(am, np.uint64)
    dm = np.asarray(dm, np.uint64)
    t = Tree()
    t.n_nodes = am.shape[0]
    t.beta = beta
    for i in range(am.shape[0]):
        t.am[i].w[0] = int(am[i, 0])
        t.am[i].w[1] = int(am[i, 1])
        t.dm[i].w[0] = int(dm[i, 0])
        t.dm[i].w[1] = int(dm[i, 1])
    return t


def rows(t: Tree, which="dm"):
    n = t.n_nodes
    src = getattr(t, which)
    out = np.zeros((n, 2), np.uint64)
    for i in range(n):
        out[i, 0] = src[i].w[0]
        out[i, 1] = src[i].w[1]
    return out


def bit(row: Row, i):
    return (row.w[i >> 6] >> (i & 63)) & 1


def parents(t: Tree):
    return [lib().orc_get_parent(C.byref(t), i) for i in range(t.n_nodes)]


def make_scorer(alpha, beta, sd, is_mutated, is_known, n_genes):
    is_mutated = np.ascontiguousarray(is_mutated, np.uint64)
    is_known = np.ascontiguousarray(is_known, np.uint64)
    s = Scorer()
    lib().orc_scorer_init(
        C.byref(s), alpha, beta, sd, is_mutated.shape[0], n_genes, _ptr(is_mutated), _ptr(is_known)
    )
    return s


def logscore_tree(s: Scorer, t: Tree, with_prior=True, counts=False):
    cnt = np.zeros((s.n_cells, 4), np.uint8) if counts else None
    node = np.zeros(s.n_cells, np.uint32) if counts else None
    v = lib().orc_logscore_tree(C.byref(s), C.byref(t), int(with_prior), _ptr(cnt), _ptr(node))
    return (v, cnt, node) if counts else v


def move_probs(p: Params) -> MoveProbs:
    return lib().orc_move_probs(C.byref(p))


def run_chain(p: Params, s: Scorer, chain, is_first, job=0, init=None, init_score=math.nan,
              trace=False):
    res = ChainResult()
    tr = np.zeros(p.n_steps, TRACE_DTYPE) if trace else None
    lib().orc_run_chain(
        C.byref(p), C.byref(s), job, chain, int(is_first),
        C.byref(init) if init is not None else None, init_score, C.byref(res), _ptr(tr),
    )
    return (res, tr) if trace else res


def run_chain_proposals(p: Params, s: Scorer, chain, is_first, n_nodes, job=0):
    """orc_run_chain with every step's proposed tree: (result, trace, am,
    dm), am/dm [n_steps][n_nodes][2] uint64 rows."""
    res = ChainResult()
    tr = np.zeros(p.n_steps, TRACE_DTYPE)
    am = np.zeros((p.n_steps, n_nodes, 2), np.uint64)
    dm = np.zeros((p.n_steps, n_nodes, 2), np.uint64)
    lib().orc_run_chain_proposals(
        C.byref(p), C.byref(s), job, chain, int(is_first), None, math.nan, C.byref(res),
        _ptr(tr), _ptr(am), _ptr(dm),
    )
    return res, tr, am, dm


def run(p: Params, s: Scorer, job=0):
    t = Tree()
    sc = C.c_float()
    ch = _u32()
    st = _u32()
    lib().orc_run(C.byref(p), C.byref(s), job, C.byref(t), C.byref(sc), C.byref(ch), C.byref(st))
    return t, sc.value, ch.value, st.value


def host_threads():
    """Host threads for parallel replays: the cores this process may use,
    at most 16 (a GPU box's share)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def run_chains(p: Params, s: Scorer, job=0, first=0, count=None, n_threads=None):
    """Chains [first, first + count) in parallel; a ctypes array of
    ChainResult (index i = chain first + i)."""
    if count is None:
        count = p.n_chains - first
    res = (ChainResult * count)()
    lib().orc_run_chains(C.byref(p), C.byref(s), job, first, count,
                         n_threads or host_threads(), C.cast(res, C.c_void_p))
    return res


def run_chain_list(p: Params, s: Scorer, chains, job=0, n_threads=None):
    """orc_run_chain for each chain of a list (any chain indices), on a pool
    of host threads (ctypes releases the interpreter lock during the call);
    ChainResult per chain, in list order."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(n_threads or host_threads()) as ex:
        return list(ex.map(lambda c: run_chain(p, s, c, c == 0, job=job), chains))


def chain_major_best(results):
    """The reference's global maximum (src/Application.hpp:456, P = 1) from
    per-chain results of chains 0..n-1: chain 0's record seeds it, a later
    chain replaces it only with a strictly greater score. Returns the chain
    index (or None when no chain has a record)."""
    best = None
    for c, r in enumerate(results):
        if r.best_step == 0xFFFFFFFF:
            continue
        if best is None or c == 0 or r.best_score > results[best].best_score:
            best = c
    return best


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox(_ptr(c), _ptr(k), _ptr(out))
    return out


def draw_move(seed, job, chain, step, n_nodes, probs: MoveProbs) -> RawMove:
    return lib().orc_draw_move(seed, job, chain, step, n_nodes, probs)


def bench_score(s: Scorer, dm_rows, beta, n_threads):
    dm_rows = np.ascontiguousarray(dm_rows, np.uint64)
    ll = np.zeros(dm_rows.shape[0], np.float32)
    rate = lib().orc_bench_score(C.byref(s), _ptr(dm_rows), dm_rows.shape[0], beta, n_threads, _ptr(ll))
    return rate, ll


def bench_chains(p: Params, s: Scorer, n_chains, n_steps, n_threads):
    cs = C.c_float()
    rate = lib().orc_bench_chains(C.byref(p), C.byref(s), n_chains, n_steps, n_threads, C.byref(cs))
    return rate, cs.value
