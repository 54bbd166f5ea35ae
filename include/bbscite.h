/*
 * bbscite: C-ABI of the MI355X implementation of bbSCITE's MCMC
 * mutation-tree search.
 *
 * This is the drop-in boundary for the chain loop of the reference. Each
 * entry point names the reference interface it replaces. All pointers are
 * host memory owned by the caller; the library copies in and out and owns
 * all device memory. No exception crosses this boundary: every function
 * returns a status (0 = ok, < 0 = error) and bbs_last_error() describes the
 * last failure.
 *
 * Bit rows: a tree row or a data row is two uint64_t words, bit i of the row
 * being bit (i % 64) of word (i / 64). A tree of n_nodes = n_genes + 1 nodes
 * is n_nodes rows; node n_nodes - 1 is the root.
 *   am[x] bit y : x is an ancestor-or-self of y   ("ancestor" matrix, the
 *                 descendants of x; src/MutationTree.hpp:645-650)
 *   dm[x] bit y : y is an ancestor-or-self of x   ("descendant" matrix, the
 *                 ancestors of x; src/MutationTree.hpp:718-723)
 *   is_mutated[c] / is_known[c] bit g : gene g of cell c
 *                 (src/main.cpp:103-144)
 */
#ifndef BBSCITE_H
#define BBSCITE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BBS_MAX_GENES 127u
#define BBS_MAX_CELLS 128u
#define BBS_MAX_NODES 128u

/* Status codes. */
#define BBS_OK 0
#define BBS_ERR_ARG (-1)    /* invalid argument                         */
#define BBS_ERR_SIZE (-2)   /* genes/cells beyond the supported limits  */
#define BBS_ERR_HIP (-3)    /* HIP runtime error                        */
#define BBS_ERR_BUDGET (-4) /* request exceeds the device memory budget */
#define BBS_ERR_STATE (-5)  /* call out of order (e.g. results before run) */
#define BBS_ERR_NODEV (-6)  /* no usable gfx950 device                  */

/* Move types (src/MoveType.hpp:23-42). */
#define BBS_MOVE_CHANGE_BETA 0u
#define BBS_MOVE_PRUNE_REATTACH 1u
#define BBS_MOVE_SWAP_NODES 2u
#define BBS_MOVE_SWAP_SUBTREES 3u

/*
 * Run parameters; the counterpart of class Parameters
 * (src/Parameters.hpp:39-45 defaults, :61-281 flags). Move probabilities are
 * given unnormalised, as on the command line, and normalised by their sum in
 * float as Parameters::get_prob_* do (src/Parameters.hpp:451-541,609-612).
 */
typedef struct bbs_params {
  float alpha_mean;            /* -fd : false-positive rate, 6.04e-5        */
  float beta_mean;             /* -ad : false-negative rate, 0.4309          */
  float beta_sd;               /* -sd : prior sd of beta, 0.1                */
  float gamma;                 /* -g  : likelihood exponent, 1               */
  float prob_beta_change;      /* -e  : 0                                    */
  float prob_prune_n_reattach; /* -move_probs p q r : 0.55 0.4 0.05          */
  float prob_swap_nodes;
  float prob_swap_subtrees;
  float beta_jump_scaling_chi; /* -x  : 10; beta jump sd = sd / chi          */
  uint32_t n_chains;           /* -r                                         */
  uint32_t n_steps;            /* -l                                         */
  uint64_t seed;               /* -seed (the reference ignores it)           */
} bbs_params;

/* Fill *p with the reference defaults (src/Parameters.hpp:39-45),
 * n_chains = n_steps = 0 and seed = 0. */
void bbs_params_default(bbs_params *p);

typedef struct bbs_ctx bbs_ctx;

/* One step of one chain, for replay against the CPU oracle. The fields are
 * ModificationParameters (src/MutationTree.hpp:84-89) plus the proposal's
 * score and the Metropolis outcome (src/Application.hpp:441-456). */
typedef struct bbs_step_trace {
  uint32_t move_type;
  uint32_t v, w, parent_of_v, parent_of_w, descendant_of_v, nondescendant_of_v;
  float new_beta;
  float proposed_score;
  uint32_t flags; /* bit 0: accepted; bit 1: new chain-local maximum */
} bbs_step_trace;

/*
 * Replaces Application's constructor (src/Application.hpp:88-186) minus the
 * host-side move pre-generation: binds the encoded data and parameters to
 * the first gfx950 device. Initial chain states are drawn from the seed
 * (uniform Pruefer code per chain, src/Application.hpp:149-164) unless
 * bbs_set_initial_states is called. The chain count is used as given (the
 * reference's emulation build, pipeline capacity 1, src/main.cpp:28-29).
 * is_mutated / is_known: n_cells rows each.
 */
int bbs_create(const bbs_params *params, uint32_t n_cells, uint32_t n_genes,
               const uint64_t *is_mutated, const uint64_t *is_known,
               bbs_ctx **out);

/* Replaces the initial-state generation of src/Application.hpp:149-164 with
 * caller-provided states: am/dm [n_chains][n_nodes] rows, beta[n_chains].
 * score may be NULL, in which case the device scores the initial trees. */
int bbs_set_initial_states(bbs_ctx *ctx, const uint64_t *am, const uint64_t *dm,
                           const float *beta, const float *score);

/* Record a bbs_step_trace for every step of chains [0, n_traced_chains). */
int bbs_set_trace(bbs_ctx *ctx, uint32_t n_traced_chains);

/* Record the proposed tree of every step of chains [0, n_traced_chains),
 * accepted or not: its ancestor and descendant rows as the device built
 * them, so that each move's update can be checked matrix by matrix (the
 * reference's move test checks the full matrices after every update,
 * test/MutationTree.cpp:697-735; Application itself has no such record).
 * Device memory: 32 * n_nodes bytes per traced step. */
int bbs_set_proposal_trace(bbs_ctx *ctx, uint32_t n_traced_chains);

/* How the chain kernel scores proposals; results are bit-identical in every
 * mode (no reference counterpart: the reference rescores every pair,
 * src/TreeScorer.hpp:131-210). FULL rescores all (node, cell) pairs each
 * step; INCREMENTAL keeps the current tree's pair scores in LDS and
 * rescores only the nodes whose ancestor set a move changes; AUTO (the
 * default) uses INCREMENTAL when its table fits without costing concurrent
 * chains. BBS_ERR_SIZE if INCREMENTAL is impossible for this size. */
#define BBS_SCORING_AUTO 0
#define BBS_SCORING_FULL 1
#define BBS_SCORING_INCREMENTAL 2
/* HALF_TABLE (datasets of over 96 cells): the table holds cells 0..63 and
 * the other cells are rescored every step, in about half the LDS. AUTO
 * takes it where the whole table would cost concurrent chains and the half
 * table does not. BBS_ERR_SIZE for 96 cells or fewer. */
#define BBS_SCORING_HALF_TABLE 3
int bbs_set_scoring_mode(bbs_ctx *ctx, int mode);

/* Waves per chain of the plain launch (no reference counterpart: a launch
 * shape; the reference runs one pipelined chain step at a time,
 * src/Application.hpp:398-496). 1: one wavefront per chain. 2 or 4: each
 * chain runs on a workgroup of that many wavefronts that split its cells
 * (ceil(n_cells / k) each) and compute the same move, acceptance and ordered
 * cell sum, with one workgroup barrier per step; with over 64 cells (two
 * cell slots in one wave) it shortens a step where the launch has fewer
 * chains than the device has SIMDs. 0 (the default): 2 when the dataset has
 * over 64 cells, 2 n_chains waves fit the device's SIMDs, the incremental
 * scorer is in use and no speculation applies (bbs_set_speculation), else 1.
 * Results are
 * bit-identical for every value; with BBS_SCORING_FULL a chain always has
 * one wavefront (a split chain keeps the incremental scorer's table).
 * BBS_ERR_ARG for other values, BBS_ERR_SIZE
 * if the split does not fit this dataset (over the LDS of a workgroup or a
 * wave without cells). */
int bbs_set_chain_split(bbs_ctx *ctx, uint32_t waves_per_chain);

/* Speculative waves per chain of the plain launch (no reference counterpart:
 * the reference evaluates one step of a chain at a time,
 * src/Application.hpp:412-487). k = 2, 4 or 8: each chain runs on k
 * wavefronts; in each round wave i proposes step s + i against the same
 * current tree and pair-score table, which is the chain's own proposal of
 * that step whenever steps s .. s + i - 1 are rejected. The steps up to the
 * first accepted one are kept and the rest proposed again from the new tree,
 * so a round advances the chain by up to k steps in about the time of one.
 * A round lasts as long as its slowest proposal (a beta move rescores every
 * pair), so the gain is below k. 1: off. 0 (the default): the largest of 8,
 * 4 and 2 whose waves run at most two per SIMD with every chain's LDS at
 * once, while the incremental scorer is in use, else 1; when it applies it
 * takes precedence over bbs_set_chain_split.
 * Trajectories, traces and
 * bests are bit-identical for every value; with BBS_SCORING_FULL it is off.
 * BBS_ERR_ARG for other values, BBS_ERR_SIZE if the k waves' LDS does not fit
 * one workgroup. */
int bbs_set_speculation(bbs_ctx *ctx, uint32_t waves_per_chain);

/* Replaces Application::run_simulation (src/Application.hpp:197-219): runs
 * all chains for n_steps steps and blocks; *makespan_ms (nullable) receives
 * the device time in milliseconds, measured with HIP events. */
int bbs_run(bbs_ctx *ctx, float *makespan_ms);

/* Replaces get_best_am/get_best_dm/get_best_beta/get_best_score
 * (src/Application.hpp:224-256): the first strictly greatest proposed tree
 * in chain-major step order (src/Application.hpp:456). am/dm: n_nodes rows
 * each; any pointer may be NULL. */
int bbs_get_best(bbs_ctx *ctx, uint64_t *am, uint64_t *dm, float *beta,
                 float *score);

/* Split form of bbs_run for callers that run the chains in windows:
 * bbs_begin draws (or keeps) the initial states and clears the maxima;
 * bbs_advance runs n more steps of every chain (steps continue where the
 * previous window ended, so any split gives the same result as one
 * bbs_run). No reference counterpart: the reference runs all steps in one
 * kernel invocation. */
int bbs_begin(bbs_ctx *ctx);
int bbs_advance(bbs_ctx *ctx, uint32_t n_steps, float *makespan_ms);

/* Per-chain outcome after bbs_run (no reference counterpart; the reference
 * keeps only the global best). Each array has n_chains entries and may be
 * NULL. best_step is UINT32_MAX for a chain without a finite maximum. */
int bbs_get_chain_results(bbs_ctx *ctx, float *best_score, uint32_t *best_step,
                          float *final_score, float *final_beta);

/* Final current state of every chain: am/dm [n_chains][n_nodes] rows. */
int bbs_get_chain_states(bbs_ctx *ctx, uint64_t *am, uint64_t *dm);

/* Copy the recorded trace: [n_traced_chains][n_steps] records. */
int bbs_get_trace(bbs_ctx *ctx, bbs_step_trace *out);

/* Copy the recorded proposals: am/dm [n_traced_chains][n_steps][n_nodes]
 * rows (each row two 64-bit words, as bbs_get_chain_states); either may be
 * NULL. */
int bbs_get_proposal_trace(bbs_ctx *ctx, uint64_t *am, uint64_t *dm);

#define BBS_SCORE_NO_PRIOR 1u
/* Use the lane-per-cell scorer (the chain kernel's) instead of the bulk
 * scorer; results are identical. */
#define BBS_SCORE_CELL_LANES 2u
/* Bulk scorer with the pair counts on the vector unit (one tree per lane)
 * instead of the matrix cores (one tree per wave); results are identical. */
#define BBS_SCORE_VALU_LANES 4u

/*
 * Replaces TreeScorer::logscore_tree (src/TreeScorer.hpp:131-210) over many
 * trees: dm_rows [n_trees][n_nodes] rows (ancestor sets, the rows the scorer
 * reads, src/TreeScorer.hpp:155), betas [n_trees] or NULL for beta_mean.
 * ll[n_trees] receives the log-likelihood (plus the beta prior unless
 * flags has BBS_SCORE_NO_PRIOR). counts (nullable) receives, per tree and
 * cell, the four counts n00,n01,n10,n11 (observed x true) at the cell's
 * best attachment node, the first maximal node: [n_trees][n_cells][4].
 */
int bbs_score_trees(bbs_ctx *ctx, const uint64_t *dm_rows, uint32_t n_trees,
                    const float *betas, uint32_t flags, float *ll,
                    uint8_t *counts);

/*
 * No counterpart in the reference: an allocation-free bbs_score_trees for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Arguments,
 * results and errors are bbs_score_trees's; the per-call device staging
 * (the trees' rows, betas, scores and counts) lives in one fixed device
 * array of BBS_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends,
 * instead of buffers allocated per call, and the call makes no hipMalloc or
 * hipFree. The staging of one call, with each part rounded up to 256 bytes,
 * is 16 n_trees n_nodes (rows) + 4 n_trees (betas, when given) + 4 n_trees
 * (scores) + 4 n_trees n_cells (counts, when requested) bytes; a call that
 * needs more returns BBS_ERR_SIZE. Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (from several threads or contexts)
 * race.
 */
#define BBS_NOALLOC_SCRATCH_BYTES (16u << 20)
int bbs_score_trees_noalloc(bbs_ctx *ctx, const uint64_t *dm_rows, uint32_t n_trees,
                            const float *betas, uint32_t flags, float *ll,
                            uint8_t *counts);

/*
 * Bulk scoring with device-resident inputs (the config of
 * cpu_scoring_benchmark/main.cpp:168-212: random trees scored against one
 * matrix): n_trees uniform random trees are drawn on the device from seed
 * (bbs_random_trees gives the same trees on the host), then scored reps
 * times. kernel_ms[reps] receives each scoring launch's device time (HIP
 * events on the launch stream); ll_head receives the first n_head scores.
 */
int bbs_bench_score_random(bbs_ctx *ctx, uint64_t seed, uint32_t n_trees,
                           uint32_t reps, uint32_t flags, float *kernel_ms,
                           float *ll_head, uint32_t n_head);

/* Random trees drawn on the device, copied to the host: am (nullable), dm
 * [n_trees][n_nodes] rows. purpose 2 gives the initial chain states,
 * purpose 6 the bulk-scoring trees. */
int bbs_random_trees_device(bbs_ctx *ctx, uint64_t seed, uint32_t purpose,
                            uint32_t n_trees, uint64_t *am, uint64_t *dm);

/* Calibration of the int32 VALU roofline on this device: a kernel of
 * independent v_and_b32 / v_bcnt_u32_b32 chains on every lane of every CU;
 * returns the sustained int32 lane operations per second and wave64 VALU
 * instructions per second. No reference counterpart. */
int bbs_measure_valu_peak(bbs_ctx *ctx, double *int32_ops_per_s,
                          double *valu_insts_per_s);

/* Device values of expf(x), logf(x) and (float)log(1.0 - x) as the chain
 * kernel computes them (the libm calls of src/Application.hpp:454 and
 * src/TreeScorer.hpp:124,136-137), for comparison with the host libm. */
int bbs_eval_libm(bbs_ctx *ctx, const float *x, uint32_t n, float *exp_out,
                  float *log_out, float *log1m_out);

/* ---- batched form ---------------------------------------------------------
 * Many independent datasets ("jobs") run together: the (job, chain) items
 * of each register layout go out as one launch, largest work first, and the
 * layouts run side by side on up to BBS_MAX_STREAMS (16) streams. Job j
 * draws its random
 * stream from
 * (params.seed, job_tag, chain, step), so a job's results are identical to
 * a plain bbs_run of the same data and parameters when job_tag is 0,
 * whatever else is in the batch. No reference counterpart: the reference
 * runs one dataset per FPGA invocation (src/main.cpp:159-161). */
typedef struct bbs_job {
  bbs_params params;             /* n_chains <= 65535                    */
  uint32_t n_cells, n_genes;
  const uint64_t *is_mutated;    /* n_cells rows, read at create time    */
  const uint64_t *is_known;
  uint32_t job_tag;              /* Philox job word; 0 = as bbs_run      */
} bbs_job;

/* Launch shape (chosen by bbs_batch_autotune). */
#define BBS_MAX_STREAMS 16
typedef struct bbs_batch_config {
  uint32_t waves_per_group; /* chains per workgroup: 1, 2 or 4          */
  uint32_t n_streams;       /* streams the layouts are spread over,
                               1..BBS_MAX_STREAMS (they overlap on as
                               many hardware queues as GPU_MAX_HW_QUEUES
                               gives the process)                       */
  uint32_t order;           /* 0: largest work first; 1: job order      */
  uint32_t persistent;      /* reserved, 0: the persistent kernel form
                               was retired; bbs_batch_run rejects 1     */
  uint32_t scoring;         /* BBS_SCORING_AUTO (incremental where the
                               pair-score tables fit) or _FULL          */
  uint32_t graph;           /* 1: the first bbs_batch_run captures its
                               launches (initial trees, resets, every
                               layout on every stream) in a hipGraph;
                               later runs replay it                     */
  uint32_t two_wave_layouts; /* register budget of the incremental scorer
                               per layout: bit BBS_LAYOUT_BIT(NS, CS, W)
                               set runs that layout (NS node slots, CS
                               cell slots, W gene words) with a build
                               held to 256 registers, so two waves share
                               a SIMD; clear, with one wave per SIMD and
                               every register. Results are identical.   */
} bbs_batch_config;

/* Register layouts (bbs_batch_config.two_wave_layouts): NS = 1 node slot
 * for up to 64 nodes, else 2; CS = 1 cell slot for up to 64 cells, 3 (a
 * paired half slot) for up to 96, else 2; W = ceil(genes / 32) 32-bit gene
 * words (at least 2 with two node slots). */
#define BBS_LAYOUT_BIT(NS, CS, W) (1u << (((NS)-1u) * 12u + ((CS)-1u) * 4u + ((W)-1u)))
/* The default: the one-node-slot layouts, which fit 256 registers with at
 * most a few spills. */
#define BBS_TWO_WAVE_DEFAULT 0x00000fffu

typedef struct bbs_batch bbs_batch;

void bbs_batch_config_default(bbs_batch_config *cfg);
/* cfg may be NULL (defaults). */
int bbs_batch_create(const bbs_job *jobs, uint32_t n_jobs,
                     const bbs_batch_config *cfg, bbs_batch **out);
/* Runs every job from its initial states to completion; rerunning
 * reproduces the same results. */
int bbs_batch_run(bbs_batch *b, float *makespan_ms);
uint64_t bbs_batch_total_steps(const bbs_batch *b);
int bbs_batch_get_best(bbs_batch *b, uint32_t job, uint64_t *am, uint64_t *dm,
                       float *beta, float *score);
int bbs_batch_get_chain_results(bbs_batch *b, uint32_t job, float *best_score,
                                uint32_t *best_step, float *final_score,
                                float *final_beta);
const char *bbs_batch_last_error(const bbs_batch *b);
void bbs_batch_destroy(bbs_batch *b);
/* Tunes the launch on the jobs with n_steps capped at probe_steps: first
 * each layout's register budget where the two builds differ in occupancy
 * (its jobs alone, one wave per SIMD against two), then waves per workgroup
 * {1, 2} x streams {2, 16} x item order with those budgets; returns the
 * fastest configuration. */
int bbs_batch_autotune(const bbs_job *jobs, uint32_t n_jobs, uint32_t probe_steps,
                       bbs_batch_config *best, float *best_ms);

/* ---- serving: continuous batching of tree-search jobs ------------------
 * No reference counterpart: the reference runs one dataset per invocation
 * (src/main.cpp:159-169). A server keeps every admitted job's chains and
 * best trees on the device and advances them in slices. The resident jobs
 * of one register layout form a lane: one kernel launch per slice on the
 * lane's own stream, relaunched as soon as that slice ends, so layouts never
 * wait for each other. Jobs that arrive meanwhile are admitted as device
 * memory and concurrent-wave capacity free up. Every random draw is keyed by
 * (seed, job_tag, chain, step), so a job's result is the same alone
 * (bbs_create + bbs_run with job_tag 0) and in any mix. */
typedef struct bbs_server_config {
  uint64_t hbm_budget_bytes;    /* device memory for resident jobs' state  */
  uint32_t round_cost;          /* first slice of a layout: a job advances
                                   round_cost / its step cost steps (the
                                   cost model counts about one unit per
                                   VALU instruction of a full rescoring)   */
  float round_ms;               /* later slices: steps sized so that each
                                   job's slice takes about round_ms, from
                                   the measured time of its layout's last
                                   slice                                   */
  uint32_t max_resident_chains; /* 0: 16 times the chain waves the device
                                   runs at once (two per SIMD); waves
                                   beyond those are queued by the hardware,
                                   and the LDS budget of admission grows by
                                   the same factor                         */
  uint32_t waves_per_group;     /* chain waves per workgroup, 1..4          */
  uint32_t n_streams;           /* 1..BBS_MAX_STREAMS: lanes are dealt out
                                   over these streams; they run at once
                                   only on as many hardware queues as HIP
                                   gives the process (GPU_MAX_HW_QUEUES)   */
  uint32_t two_wave_layouts;    /* as bbs_batch_config.two_wave_layouts     */
  uint32_t admission;           /* BBS_ADMIT_*: which pending job is
                                   admitted next                           */
  float aging_s;                /* BBS_ADMIT_LEAST_WORK: a job's predicted
                                   work counts as multiplied by
                                   exp(-(seconds pending) / aging_s), so
                                   aging_s seconds of waiting weigh as a
                                   factor e of work; small values approach
                                   arrival order; 0: no aging (a steady
                                   stream of smaller jobs can then hold a
                                   large one back)                         */
} bbs_server_config;

/* Admission order of pending jobs (no reference counterpart: the reference
 * runs one job per process). The policy picks the next candidate; it is
 * admitted if it fits the budgets, else admission waits for resources. */
#define BBS_ADMIT_ARRIVAL 0u    /* arrival order                          */
#define BBS_ADMIT_LEAST_WORK 1u /* least predicted work first: chains x
                                   steps x the layout's step cost (ties in
                                   arrival order)                         */

typedef struct bbs_job_result {
  uint64_t id;
  uint32_t n_genes, n_cells, n_chains, n_steps;
  float best_score, best_beta;
  uint32_t best_chain, best_step; /* chain-major first maximum          */
  double submit_s, admit_s, done_s; /* seconds on the server clock      */
  uint64_t best_am[BBS_MAX_NODES][2], best_dm[BBS_MAX_NODES][2];
} bbs_job_result;

typedef struct bbs_server bbs_server;
void bbs_server_config_default(bbs_server_config *cfg);
int bbs_server_create(const bbs_server_config *cfg, bbs_server **out);
/* Copies the job (parameters and data); *id identifies its result. */
int bbs_server_submit(bbs_server *s, const bbs_job *job, uint64_t *id);
/* Collects the lanes whose slice has ended (retiring finished jobs), admits
 * pending jobs in the configured order while they fit, and launches the next slice
 * of every idle lane. If none of that was possible while slices run, waits
 * for the oldest one to end first. Counts may be NULL. */
int bbs_server_step(bbs_server *s, uint32_t *n_resident, uint32_t *n_pending,
                    uint32_t *n_done_ready);
/* Moves up to max finished results (in completion order) to out. */
int bbs_server_poll(bbs_server *s, bbs_job_result *out, uint32_t max, uint32_t *n_out);
/* Seconds since the server was created (the clock of the results). */
double bbs_server_now(const bbs_server *s);
/* Chain steps run so far, the time during which any slice ran, and the
 * number of slices launched. */
int bbs_server_stats(const bbs_server *s, uint64_t *chain_steps, double *busy_s,
                     uint64_t *slices);
const char *bbs_server_last_error(const bbs_server *s);
void bbs_server_destroy(bbs_server *s);

/* Message of the last failure on ctx, or of the last failed bbs_create on
 * this thread when ctx is NULL. */
const char *bbs_last_error(const bbs_ctx *ctx);

void bbs_destroy(bbs_ctx *ctx);

/* ---- host-only helpers (no device needed) ------------------------------ */

/* Parse the reference CSV format (src/main.cpp:52-101): entries 0..3
 * separated by spaces, genes as rows. matrix receives n_genes x n_cells
 * int8 entries (capacity BBS_MAX_GENES x BBS_MAX_CELLS is always enough
 * for accepted inputs). */
int bbs_parse_csv(const char *text, uint64_t length, int8_t *matrix,
                  uint32_t *n_genes, uint32_t *n_cells);

/* Encode a genes x cells matrix into per-cell rows (src/main.cpp:103-144):
 * 0 -> known, not mutated; 1, 2 -> known, mutated; 3 -> unknown. */
int bbs_encode_matrix(const int8_t *matrix, uint32_t n_genes, uint32_t n_cells,
                      uint64_t *is_mutated, uint64_t *is_known);

/* Synthetic data: uniform Pruefer tree over n_genes + 1 nodes, uniform cell
 * attachment, truth = ancestor-or-self, then 0->1 with prob
 * prob_false_positive, 1->0 with prob prob_false_negative, then missing (3)
 * with prob prob_missing. matrix [n_genes][n_cells]; true_parent
 * [n_genes + 1] and attachment [n_cells] are optional. */
int bbs_generate_dataset(uint64_t seed, uint32_t n_genes, uint32_t n_cells,
                         float prob_false_positive, float prob_false_negative,
                         float prob_missing, int8_t *matrix,
                         uint32_t *true_parent, uint32_t *attachment);

/* Uniform random trees (Pruefer codes) as dm rows [n_trees][n_genes + 1]. */
int bbs_random_trees(uint64_t seed, uint32_t n_genes, uint32_t n_trees,
                     uint64_t *dm_rows);

/* The initial tree of (job, chain) as the device draws it (am/dm rows). */
int bbs_initial_tree(uint64_t seed, uint32_t job, uint32_t chain,
                     uint32_t n_genes, uint64_t *am, uint64_t *dm);

/* Parent vector -> am/dm rows (src/MutationTree.hpp:379-409); Pruefer code
 * -> parent vector (src/MutationTree.hpp:427-476). */
int bbs_parent_vector_to_matrix(const uint32_t *parent, uint32_t n_nodes,
                                uint64_t *am, uint64_t *dm);
int bbs_pruefer_to_parent_vector(const uint32_t *code, uint32_t code_len,
                                 uint32_t *parent);

/* Graphviz and Newick text of a tree given by its am/dm rows
 * (src/MutationTree.hpp:799-837,880-894). Returns the text length, or a
 * negative status; writes at most cap bytes including the terminating 0. */
int64_t bbs_tree_to_graphviz(const uint64_t *am, const uint64_t *dm,
                             uint32_t n_nodes, char *buf, uint64_t cap);
int64_t bbs_tree_to_newick(const uint64_t *am, const uint64_t *dm,
                           uint32_t n_nodes, char *buf, uint64_t cap);

/* ---- quality harness (verification/__main__.py:54-163 restated) ---------
 * Newick text as written by bbs_tree_to_newick (or any labelling that
 * src/MutationTree.hpp:880-894 produces) -> parent vector; labels that are
 * not 0..n-1 are renumbered in sorted order (verification/lib.py:148-164);
 * the root must carry the largest label. parent[root] = root. */
int bbs_newick_to_parent_vector(const char *text, uint32_t *parent, uint32_t *n_nodes);
/* Log-likelihood of a tree in double precision with each cell attached to
 * its most likely node (verification/__main__.py:54-87,
 * verification/lib.py:170-196 with the entry meanings of src/main.cpp:
 * 0 = not mutated, 1 and 2 = mutated, 3 = missing; the first maximal node
 * in index order wins). attachment may be NULL. */
int bbs_score_attachments(const uint64_t *dm, uint32_t n_nodes, const uint64_t *is_mutated,
                          const uint64_t *is_known, uint32_t n_cells, double alpha,
                          double beta, uint32_t *attachment, double *log_likelihood);
/* Two one-sided t-tests (verification/__main__.py:90-120) of n paired
 * differences against the margin +-delta: p_lower for H0 mean <= -delta,
 * p_upper for H0 mean >= delta; equivalence at level a when both are
 * below a. All-equal differences give p = 0 (or 1) exactly. */
int bbs_tost(const double *diff, uint32_t n, double delta, double *mean, double *p_lower,
             double *p_upper);

#ifdef __cplusplus
}
#endif

#endif /* BBSCITE_H */
