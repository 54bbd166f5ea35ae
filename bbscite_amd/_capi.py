"""ctypes binding of the C-ABI in include/bbscite.h (libbbscite.so).

The shared library is built in-tree by bbscite_amd.build (hipcc for gfx950)
and loaded from this package directory. There is no fallback: if the
library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbbscite.so")

MAX_GENES = 127
MAX_CELLS = 128
MAX_NODES = 128

OK = 0
ERR_ARG = -1
ERR_SIZE = -2
ERR_HIP = -3
ERR_BUDGET = -4
ERR_STATE = -5
ERR_NODEV = -6
NOALLOC_SCRATCH_BYTES = 16 << 20

MOVE_CHANGE_BETA = 0
MOVE_PRUNE_REATTACH = 1
MOVE_SWAP_NODES = 2
MOVE_SWAP_SUBTREES = 3

SCORE_NO_PRIOR = 1
SCORE_CELL_LANES = 2
SCORE_VALU_LANES = 4

PURPOSE_INIT_TREE = 2
PURPOSE_TREE_SAMPLE = 6


class Params(C.Structure):
    """bbs_params (include/bbscite.h); defaults of src/Parameters.hpp:39-45."""

    _fields_ = [
        ("alpha_mean", C.c_float),
        ("beta_mean", C.c_float),
        ("beta_sd", C.c_float),
        ("gamma", C.c_float),
        ("prob_beta_change", C.c_float),
        ("prob_prune_n_reattach", C.c_float),
        ("prob_swap_nodes", C.c_float),
        ("prob_swap_subtrees", C.c_float),
        ("beta_jump_scaling_chi", C.c_float),
        ("n_chains", C.c_uint32),
        ("n_steps", C.c_uint32),
        ("seed", C.c_uint64),
    ]


class StepTrace(C.Structure):
    _fields_ = [
        ("move_type", C.c_uint32),
        ("v", C.c_uint32),
        ("w", C.c_uint32),
        ("parent_of_v", C.c_uint32),
        ("parent_of_w", C.c_uint32),
        ("descendant_of_v", C.c_uint32),
        ("nondescendant_of_v", C.c_uint32),
        ("new_beta", C.c_float),
        ("proposed_score", C.c_float),
        ("flags", C.c_uint32),
    ]


TRACE_DTYPE = np.dtype(
    [
        ("move_type", "<u4"),
        ("v", "<u4"),
        ("w", "<u4"),
        ("parent_of_v", "<u4"),
        ("parent_of_w", "<u4"),
        ("descendant_of_v", "<u4"),
        ("nondescendant_of_v", "<u4"),
        ("new_beta", "<f4"),
        ("proposed_score", "<f4"),
        ("flags", "<u4"),
    ]
)

class Job(C.Structure):
    """bbs_job (include/bbscite.h): one dataset of a batched run."""

    _fields_ = [
        ("params", Params),
        ("n_cells", C.c_uint32),
        ("n_genes", C.c_uint32),
        ("is_mutated", C.c_void_p),
        ("is_known", C.c_void_p),
        ("job_tag", C.c_uint32),
    ]


class BatchConfig(C.Structure):
    _fields_ = [
        ("waves_per_group", C.c_uint32),
        ("n_streams", C.c_uint32),
        ("order", C.c_uint32),
        ("persistent", C.c_uint32),
        ("scoring", C.c_uint32),
        ("graph", C.c_uint32),
        ("two_wave_layouts", C.c_uint32),
    ]


class ServerConfig(C.Structure):
    """bbs_server_config (include/bbscite.h)."""

    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("round_cost", C.c_uint32),
        ("round_ms", C.c_float),
        ("max_resident_chains", C.c_uint32),
        ("waves_per_group", C.c_uint32),
        ("n_streams", C.c_uint32),
        ("two_wave_layouts", C.c_uint32),
        ("admission", C.c_uint32),
        ("aging_s", C.c_float),
    ]


ADMIT_ARRIVAL = 0
ADMIT_LEAST_WORK = 1


class JobResult(C.Structure):
    """bbs_job_result (include/bbscite.h)."""

    _fields_ = [
        ("id", C.c_uint64),
        ("n_genes", C.c_uint32),
        ("n_cells", C.c_uint32),
        ("n_chains", C.c_uint32),
        ("n_steps", C.c_uint32),
        ("best_score", C.c_float),
        ("best_beta", C.c_float),
        ("best_chain", C.c_uint32),
        ("best_step", C.c_uint32),
        ("submit_s", C.c_double),
        ("admit_s", C.c_double),
        ("done_s", C.c_double),
        ("best_am", C.c_uint64 * 2 * MAX_NODES),
        ("best_dm", C.c_uint64 * 2 * MAX_NODES),
    ]

    def rows(self, which="dm"):
        n = self.n_genes + 1
        a = np.ctypeslib.as_array(getattr(self, "best_" + which))
        return a.reshape(MAX_NODES, 2)[:n].copy()


_P = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64

# name: (restype, argtypes)
_SIGS = {
    "bbs_params_default": (None, [C.POINTER(Params)]),
    "bbs_create": (C.c_int, [C.POINTER(Params), _u32, _u32, _P, _P, C.POINTER(_P)]),
    "bbs_set_initial_states": (C.c_int, [_P, _P, _P, _P, _P]),
    "bbs_set_trace": (C.c_int, [_P, _u32]),
    "bbs_set_proposal_trace": (C.c_int, [_P, _u32]),
    "bbs_set_scoring_mode": (C.c_int, [_P, C.c_int]),
    "bbs_set_chain_split": (C.c_int, [_P, C.c_uint32]),
    "bbs_set_speculation": (C.c_int, [_P, C.c_uint32]),
    "bbs_run": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "bbs_begin": (C.c_int, [_P]),
    "bbs_advance": (C.c_int, [_P, _u32, C.POINTER(C.c_float)]),
    "bbs_get_best": (C.c_int, [_P, _P, _P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "bbs_get_chain_results": (C.c_int, [_P, _P, _P, _P, _P]),
    "bbs_get_chain_states": (C.c_int, [_P, _P, _P]),
    "bbs_get_trace": (C.c_int, [_P, _P]),
    "bbs_get_proposal_trace": (C.c_int, [_P, _P, _P]),
    "bbs_score_trees": (C.c_int, [_P, _P, _u32, _P, _u32, _P, _P]),
    "bbs_score_trees_noalloc": (C.c_int, [_P, _P, _u32, _P, _u32, _P, _P]),
    "bbs_bench_score_random": (C.c_int, [_P, _u64, _u32, _u32, _u32, _P, _P, _u32]),
    "bbs_random_trees_device": (C.c_int, [_P, _u64, _u32, _u32, _P, _P]),
    "bbs_eval_libm": (C.c_int, [_P, _P, _u32, _P, _P, _P]),
    "bbs_measure_valu_peak": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bbs_last_error": (C.c_char_p, [_P]),
    "bbs_batch_config_default": (None, [C.POINTER(BatchConfig)]),
    "bbs_batch_create": (C.c_int, [C.POINTER(Job), _u32, C.POINTER(BatchConfig), C.POINTER(_P)]),
    "bbs_batch_run": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "bbs_batch_total_steps": (_u64, [_P]),
    "bbs_batch_get_best": (C.c_int, [_P, _u32, _P, _P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "bbs_batch_get_chain_results": (C.c_int, [_P, _u32, _P, _P, _P, _P]),
    "bbs_batch_last_error": (C.c_char_p, [_P]),
    "bbs_batch_destroy": (None, [_P]),
    "bbs_batch_autotune": (
        C.c_int,
        [C.POINTER(Job), _u32, _u32, C.POINTER(BatchConfig), C.POINTER(C.c_float)],
    ),
    "bbs_newick_to_parent_vector": (C.c_int, [C.c_char_p, _P, C.POINTER(_u32)]),
    "bbs_score_attachments": (C.c_int, [_P, _u32, _P, _P, _u32, C.c_double, C.c_double, _P,
                                        C.POINTER(C.c_double)]),
    "bbs_tost": (C.c_int, [_P, _u32, C.c_double, C.POINTER(C.c_double),
                           C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bbs_server_config_default": (None, [C.POINTER(ServerConfig)]),
    "bbs_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(_P)]),
    "bbs_server_submit": (C.c_int, [_P, C.POINTER(Job), C.POINTER(C.c_uint64)]),
    "bbs_server_step": (C.c_int, [_P, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
    "bbs_server_poll": (C.c_int, [_P, C.POINTER(JobResult), _u32, C.POINTER(_u32)]),
    "bbs_server_now": (C.c_double, [_P]),
    "bbs_server_stats": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                   C.POINTER(C.c_uint64)]),
    "bbs_server_last_error": (C.c_char_p, [_P]),
    "bbs_server_destroy": (None, [_P]),
    "bbs_destroy": (None, [_P]),
    "bbs_parse_csv": (C.c_int, [C.c_char_p, _u64, _P, C.POINTER(_u32), C.POINTER(_u32)]),
    "bbs_encode_matrix": (C.c_int, [_P, _u32, _u32, _P, _P]),
    "bbs_generate_dataset": (
        C.c_int,
        [_u64, _u32, _u32, C.c_float, C.c_float, C.c_float, _P, _P, _P],
    ),
    "bbs_random_trees": (C.c_int, [_u64, _u32, _u32, _P]),
    "bbs_initial_tree": (C.c_int, [_u64, _u32, _u32, _u32, _P, _P]),
    "bbs_parent_vector_to_matrix": (C.c_int, [_P, _u32, _P, _P]),
    "bbs_pruefer_to_parent_vector": (C.c_int, [_P, _u32, _P]),
    "bbs_tree_to_graphviz": (C.c_int64, [_P, _P, _u32, C.c_char_p, _u64]),
    "bbs_tree_to_newick": (C.c_int64, [_P, _P, _u32, C.c_char_p, _u64]),
}

EXPORTED = tuple(_SIGS)

_lib = None


class BbsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bbscite error {code}: {msg}")
        self.code = code


def lib():
    """Load libbbscite.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`"
            )
        lib_ = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib_, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib_
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc, ctx=None):
    if rc != OK:
        msg = lib().bbs_last_error(ctx)
        raise BbsError(rc, msg.decode() if msg else "")


def default_params(**kw) -> Params:
    p = Params()
    lib().bbs_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# ---------------------------------------------------------------- host helpers


def parse_csv(text: str | bytes):
    """Parse the reference CSV format (src/main.cpp:52-144)."""
    if isinstance(text, str):
        text = text.encode()
    m = np.zeros(MAX_GENES * MAX_CELLS, np.int8)
    g, c = _u32(), _u32()
    _check(lib().bbs_parse_csv(text, len(text), _ptr(m), C.byref(g), C.byref(c)))
    return m[: g.value * c.value].reshape(g.value, c.value).copy()


def encode_matrix(matrix: np.ndarray):
    """genes x cells int8 -> (is_mutated, is_known) as [cells][2] uint64."""
    matrix = np.ascontiguousarray(matrix, np.int8)
    G, Cn = matrix.shape
    mut = np.zeros((Cn, 2), np.uint64)
    kno = np.zeros((Cn, 2), np.uint64)
    _check(lib().bbs_encode_matrix(_ptr(matrix), G, Cn, _ptr(mut), _ptr(kno)))
    return mut, kno


def generate_dataset(seed, n_genes, n_cells, fd, ad, missing):
    """Synthetic data: returns (matrix[G][C] int8, true parent[N], attachment[C])."""
    m = np.zeros((n_genes, n_cells), np.int8)
    par = np.zeros(n_genes + 1, np.uint32)
    att = np.zeros(n_cells, np.uint32)
    _check(
        lib().bbs_generate_dataset(
            seed, n_genes, n_cells, fd, ad, missing, _ptr(m), _ptr(par), _ptr(att)
        )
    )
    return m, par, att


def random_trees(seed, n_genes, n_trees):
    dm = np.zeros((n_trees, n_genes + 1, 2), np.uint64)
    _check(lib().bbs_random_trees(seed, n_genes, n_trees, _ptr(dm)))
    return dm


def initial_tree(seed, job, chain, n_genes):
    am = np.zeros((n_genes + 1, 2), np.uint64)
    dm = np.zeros((n_genes + 1, 2), np.uint64)
    _check(lib().bbs_initial_tree(seed, job, chain, n_genes, _ptr(am), _ptr(dm)))
    return am, dm


def parent_vector_to_matrix(parent):
    parent = np.ascontiguousarray(parent, np.uint32)
    n = len(parent)
    am = np.zeros((n, 2), np.uint64)
    dm = np.zeros((n, 2), np.uint64)
    _check(lib().bbs_parent_vector_to_matrix(_ptr(parent), n, _ptr(am), _ptr(dm)))
    return am, dm


def pruefer_to_parent_vector(code):
    code = np.ascontiguousarray(code, np.uint32)
    par = np.zeros(len(code) + 2, np.uint32)
    _check(lib().bbs_pruefer_to_parent_vector(_ptr(code), len(code), _ptr(par)))
    return par


def _text(fn, am, dm):
    am = np.ascontiguousarray(am, np.uint64)
    dm = np.ascontiguousarray(dm, np.uint64)
    n = dm.shape[0]
    cap = 64 * n + 256
    buf = C.create_string_buffer(cap)
    ln = fn(_ptr(am), _ptr(dm), n, buf, cap)
    if ln < 0:
        _check(int(ln))
    return buf.value.decode()


def tree_to_graphviz(am, dm):
    return _text(lib().bbs_tree_to_graphviz, am, dm)


def tree_to_newick(am, dm):
    return _text(lib().bbs_tree_to_newick, am, dm)


def newick_to_parent_vector(text):
    """Newick text -> parent vector (root last, parent[root] = root)."""
    if isinstance(text, str):
        text = text.encode()
    par = np.zeros(MAX_NODES, np.uint32)
    n = _u32()
    _check(lib().bbs_newick_to_parent_vector(text, _ptr(par), C.byref(n)))
    return par[: n.value].copy()


def score_attachments(dm, is_mutated, is_known, alpha, beta):
    """(log-likelihood in double, attachment node per cell) of a tree with
    every cell at its most likely node."""
    dm = np.ascontiguousarray(dm, np.uint64)
    mut = np.ascontiguousarray(is_mutated, np.uint64)
    kno = np.ascontiguousarray(is_known, np.uint64)
    att = np.zeros(mut.shape[0], np.uint32)
    ll = C.c_double()
    _check(lib().bbs_score_attachments(_ptr(dm), dm.shape[0], _ptr(mut), _ptr(kno), mut.shape[0],
                                       alpha, beta, _ptr(att), C.byref(ll)))
    return ll.value, att


def tost(diff, delta):
    """Two one-sided t-tests against +-delta: (mean, p_lower, p_upper)."""
    d = np.ascontiguousarray(diff, np.float64)
    m, pl, pu = C.c_double(), C.c_double(), C.c_double()
    _check(lib().bbs_tost(_ptr(d), len(d), delta, C.byref(m), C.byref(pl), C.byref(pu)))
    return m.value, pl.value, pu.value


# ---------------------------------------------------------------- device


class Context:
    """Owns one bbs_ctx (the Application counterpart, src/Application.hpp:88-256)."""

    def __init__(self, params: Params, is_mutated, is_known, n_genes):
        self.params = params
        self.is_mutated = np.ascontiguousarray(is_mutated, np.uint64)
        self.is_known = np.ascontiguousarray(is_known, np.uint64)
        self.n_cells = self.is_mutated.shape[0]
        self.n_genes = n_genes
        self.n_nodes = n_genes + 1
        h = C.c_void_p()
        rc = lib().bbs_create(
            C.byref(params),
            self.n_cells,
            n_genes,
            _ptr(self.is_mutated),
            _ptr(self.is_known),
            C.byref(h),
        )
        _check(rc)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().bbs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        _check(rc, self.h)

    def set_initial_states(self, am, dm, beta, score=None):
        am = np.ascontiguousarray(am, np.uint64)
        dm = np.ascontiguousarray(dm, np.uint64)
        beta = np.ascontiguousarray(beta, np.float32)
        sc = None if score is None else np.ascontiguousarray(score, np.float32)
        self._c(lib().bbs_set_initial_states(self.h, _ptr(am), _ptr(dm), _ptr(beta), _ptr(sc)))

    def set_trace(self, n):
        self._c(lib().bbs_set_trace(self.h, n))

    def set_proposal_trace(self, n):
        """Record every proposed tree's rows for chains [0, n)."""
        self._c(lib().bbs_set_proposal_trace(self.h, n))

    def set_scoring_mode(self, mode):
        """"auto", "full", "incremental" or "half" (the table of cells 0..63,
        over 96 cells); results are identical."""
        code = {"auto": 0, "full": 1, "incremental": 2, "half": 3}[mode]
        self._c(lib().bbs_set_scoring_mode(self.h, code))

    def set_chain_split(self, waves_per_chain):
        """Waves per chain: 0 (auto), 1, 2 or 4 (results are identical)."""
        self._c(lib().bbs_set_chain_split(self.h, waves_per_chain))

    def set_speculation(self, waves_per_chain):
        """Speculative waves per chain: 0 (auto), 1 (off), 2, 4 or 8
        (results are identical)."""
        self._c(lib().bbs_set_speculation(self.h, waves_per_chain))

    def run(self):
        ms = C.c_float()
        self._c(lib().bbs_run(self.h, C.byref(ms)))
        return ms.value

    def begin(self):
        self._c(lib().bbs_begin(self.h))

    def advance(self, n):
        ms = C.c_float()
        self._c(lib().bbs_advance(self.h, n, C.byref(ms)))
        return ms.value

    def best(self):
        am = np.zeros((self.n_nodes, 2), np.uint64)
        dm = np.zeros((self.n_nodes, 2), np.uint64)
        b, s = C.c_float(), C.c_float()
        self._c(lib().bbs_get_best(self.h, _ptr(am), _ptr(dm), C.byref(b), C.byref(s)))
        return am, dm, b.value, s.value

    def chain_results(self):
        R = self.params.n_chains
        bs = np.zeros(R, np.float32)
        st = np.zeros(R, np.uint32)
        fs = np.zeros(R, np.float32)
        fb = np.zeros(R, np.float32)
        self._c(lib().bbs_get_chain_results(self.h, _ptr(bs), _ptr(st), _ptr(fs), _ptr(fb)))
        return bs, st, fs, fb

    def chain_states(self):
        R = self.params.n_chains
        am = np.zeros((R, self.n_nodes, 2), np.uint64)
        dm = np.zeros((R, self.n_nodes, 2), np.uint64)
        self._c(lib().bbs_get_chain_states(self.h, _ptr(am), _ptr(dm)))
        return am, dm

    def trace(self, n_traced):
        out = np.zeros((n_traced, self.params.n_steps), TRACE_DTYPE)
        self._c(lib().bbs_get_trace(self.h, _ptr(out)))
        return out

    def proposal_trace(self, n_traced):
        """(am, dm): [n_traced][n_steps][n_nodes][2] uint64 rows of every
        step's proposed tree."""
        shape = (n_traced, self.params.n_steps, self.n_genes + 1, 2)
        am = np.zeros(shape, np.uint64)
        dm = np.zeros(shape, np.uint64)
        self._c(lib().bbs_get_proposal_trace(self.h, _ptr(am), _ptr(dm)))
        return am, dm

    def score_trees(self, dm_rows, betas=None, with_prior=True, counts=False,
                    cell_lanes=False, valu_lanes=False, no_alloc=False):
        dm_rows = np.ascontiguousarray(dm_rows, np.uint64)
        T = dm_rows.shape[0]
        ll = np.zeros(T, np.float32)
        b = None if betas is None else np.ascontiguousarray(betas, np.float32)
        cnt = np.zeros((T, self.n_cells, 4), np.uint8) if counts else None
        flags = ((0 if with_prior else SCORE_NO_PRIOR) | (SCORE_CELL_LANES if cell_lanes else 0)
                 | (SCORE_VALU_LANES if valu_lanes else 0))
        fn = lib().bbs_score_trees_noalloc if no_alloc else lib().bbs_score_trees
        self._c(fn(self.h, _ptr(dm_rows), T, _ptr(b), flags, _ptr(ll), _ptr(cnt)))
        return (ll, cnt) if counts else ll

    def bench_score_random(self, seed, n_trees, reps, with_prior=False, n_head=0,
                           cell_lanes=False, valu_lanes=False):
        ms = np.zeros(reps, np.float32)
        head = np.zeros(n_head, np.float32)
        flags = ((0 if with_prior else SCORE_NO_PRIOR) | (SCORE_CELL_LANES if cell_lanes else 0)
                 | (SCORE_VALU_LANES if valu_lanes else 0))
        self._c(
            lib().bbs_bench_score_random(
                self.h, seed, n_trees, reps, flags, _ptr(ms), _ptr(head) if n_head else None, n_head
            )
        )
        return ms, head

    def measure_valu_peak(self):
        ops, insts = C.c_double(), C.c_double()
        self._c(lib().bbs_measure_valu_peak(self.h, C.byref(ops), C.byref(insts)))
        return ops.value, insts.value

    def eval_libm(self, x):
        x = np.ascontiguousarray(x, np.float32)
        e = np.zeros_like(x)
        l = np.zeros_like(x)
        l1 = np.zeros_like(x)
        self._c(lib().bbs_eval_libm(self.h, _ptr(x), len(x), _ptr(e), _ptr(l), _ptr(l1)))
        return e, l, l1

    def random_trees_device(self, seed, purpose, n_trees, with_am=False):
        dm = np.zeros((n_trees, self.n_nodes, 2), np.uint64)
        am = np.zeros_like(dm) if with_am else None
        self._c(lib().bbs_random_trees_device(self.h, seed, purpose, n_trees, _ptr(am), _ptr(dm)))
        return (am, dm) if with_am else dm


# ---------------------------------------------------------------- batches


def _jobs_array(jobs):
    """jobs: list of (params, is_mutated, is_known, n_genes[, job_tag])."""
    arr = (Job * len(jobs))()
    keep = []
    for i, j in enumerate(jobs):
        params, mut, kno, G = j[:4]
        tag = j[4] if len(j) > 4 else 0
        mut = np.ascontiguousarray(mut, np.uint64)
        kno = np.ascontiguousarray(kno, np.uint64)
        keep += [mut, kno]
        arr[i].params = params
        arr[i].n_cells = mut.shape[0]
        arr[i].n_genes = G
        arr[i].is_mutated = mut.ctypes.data
        arr[i].is_known = kno.ctypes.data
        arr[i].job_tag = tag
    return arr, keep


def _bcheck(rc, b=None):
    if rc != OK:
        msg = lib().bbs_batch_last_error(b)
        raise BbsError(rc, msg.decode() if msg else "")


def layout_bit(ns, cs, w):
    """BBS_LAYOUT_BIT: the layout's bit in two_wave_layouts."""
    return 1 << ((ns - 1) * 12 + (cs - 1) * 4 + (w - 1))


def batch_config(waves_per_group=None, n_streams=None, order=None, scoring=None, graph=None,
                 two_wave_layouts=None):
    """scoring: None/"auto" (incremental where the tables fit) or "full";
    graph: capture the run in a hipGraph on the first run, replay later;
    two_wave_layouts: layout bits (layout_bit) of the two-waves-per-SIMD
    register budget."""
    c = BatchConfig()
    lib().bbs_batch_config_default(C.byref(c))
    if two_wave_layouts is not None:
        c.two_wave_layouts = two_wave_layouts
    if graph is not None:
        c.graph = 1 if graph else 0
    if scoring is not None:
        c.scoring = {"auto": 0, "full": 1}[scoring]
    if waves_per_group is not None:
        c.waves_per_group = waves_per_group
    if n_streams is not None:
        c.n_streams = n_streams
    if order is not None:
        c.order = order
    return c


def batch_autotune(jobs, probe_steps=2000):
    arr, keep = _jobs_array(jobs)
    cfg = BatchConfig()
    ms = C.c_float()
    _bcheck(lib().bbs_batch_autotune(arr, len(jobs), probe_steps, C.byref(cfg), C.byref(ms)))
    return cfg, ms.value


class Batch:
    """Owns one bbs_batch: many datasets, one launch per register layout
    (optionally captured in a hipGraph and replayed)."""

    def __init__(self, jobs, cfg: BatchConfig | None = None):
        arr, keep = _jobs_array(jobs)
        self.jobs = jobs
        h = C.c_void_p()
        _bcheck(lib().bbs_batch_create(arr, len(jobs), C.byref(cfg) if cfg else None, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().bbs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self):
        ms = C.c_float()
        _bcheck(lib().bbs_batch_run(self.h, C.byref(ms)), self.h)
        return ms.value

    @property
    def total_steps(self):
        return lib().bbs_batch_total_steps(self.h)

    def chain_results(self, job):
        R = self.jobs[job][0].n_chains
        bs = np.zeros(R, np.float32)
        st = np.zeros(R, np.uint32)
        fs = np.zeros(R, np.float32)
        fb = np.zeros(R, np.float32)
        _bcheck(lib().bbs_batch_get_chain_results(self.h, job, _ptr(bs), _ptr(st), _ptr(fs), _ptr(fb)), self.h)
        return bs, st, fs, fb

    def best(self, job):
        n = self.jobs[job][3] + 1
        am = np.zeros((n, 2), np.uint64)
        dm = np.zeros((n, 2), np.uint64)
        b, s = C.c_float(), C.c_float()
        _bcheck(lib().bbs_batch_get_best(self.h, job, _ptr(am), _ptr(dm), C.byref(b), C.byref(s)), self.h)
        return am, dm, b.value, s.value


# ---------------------------------------------------------------- serving


def server_config(**kw) -> ServerConfig:
    c = ServerConfig()
    lib().bbs_server_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Server:
    """A continuous-batching server (bbs_server_*): submit jobs at any time,
    step() collects ended slices, admits and relaunches (one lane per register
    layout), poll() collects finished jobs."""

    def __init__(self, cfg: ServerConfig = None):
        self.h = C.c_void_p()
        self._check(lib().bbs_server_create(C.byref(cfg) if cfg is not None else None,
                                            C.byref(self.h)), create=True)

    def _check(self, rc, create=False):
        if rc != OK:
            msg = lib().bbs_server_last_error(None if create else self.h)
            raise BbsError(rc, msg.decode() if msg else "")

    def submit(self, params, is_mutated, is_known, n_genes, job_tag=0):
        arr, keep = _jobs_array([(params, is_mutated, is_known, n_genes, job_tag)])
        jid = C.c_uint64()
        self._check(lib().bbs_server_submit(self.h, arr, C.byref(jid)))
        del keep  # the server copied the data
        return jid.value

    def step(self):
        r, p, d = _u32(), _u32(), _u32()
        self._check(lib().bbs_server_step(self.h, C.byref(r), C.byref(p), C.byref(d)))
        return r.value, p.value, d.value

    def poll(self, max_results=256):
        out = (JobResult * max_results)()
        n = _u32()
        self._check(lib().bbs_server_poll(self.h, out, max_results, C.byref(n)))
        return [out[i] for i in range(n.value)]

    def now(self):
        return lib().bbs_server_now(self.h)

    def stats(self):
        st, bs, rd = C.c_uint64(), C.c_double(), C.c_uint64()
        self._check(lib().bbs_server_stats(self.h, C.byref(st), C.byref(bs), C.byref(rd)))
        return st.value, bs.value, rd.value

    def close(self):
        if self.h:
            lib().bbs_server_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
