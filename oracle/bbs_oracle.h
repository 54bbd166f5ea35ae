/*
 * CPU oracle: a plain-C restatement of the reference's chain-loop hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This code is the checker for the HIP path; it is
 * used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg, never by the product. Every function cites the reference file:line it
 * follows. Transcendentals are glibc's (as in the reference, which runs on
 * the host); random draws come from the shared Philox stream
 * (bbscite_amd/csrc/bbs_rng.h), which is an input to both sides.
 *
 * Pinning (what checks each part; see tests/test_oracle_*.py):
 *   parent/matrix/Pruefer/queries : reference known answers
 *       test/MutationTree.cpp:49-408 (fixtures in tests/golden/)
 *   move updates                  : reference known answers
 *       test/MutationTree.cpp:410-611 and an independent parent-vector
 *       oracle (restated from test/MutationTree.cpp:697-735), seeded,
 *       advancing the tree
 *   move realisation              : brute force (sorted (non)descendant
 *       lists indexed with floor(u*n)), invariants
 *   scorer counts and likelihood  : closed form test/TreeScorer.cpp:90-92;
 *       brute-force SCITE likelihood by walking parents; exhaustive over
 *       all labelled trees with 5 nodes
 *   beta prior                    : the formula of src/TreeScorer.hpp:122-126
 *       evaluated independently in double
 *   chain loop                    : restated test/Application.cpp:38-142
 *       (the planted 15x4 tree is found); on a fixture with duplicated
 *       gene rows (many equal proposals) the global and per-chain maxima
 *       equal an independent scan of the traces (first strictly greater
 *       proposal in chain-major order, chain 0's step 0 seeding it); on
 *       subtree-swap-only chains every acceptance decision equals one
 *       recomputed from parent vectors (which node is the ancestor, the
 *       nd(v)/nd(w) ratio, glibc expf) and every proposal's score equals
 *       the score of the brute-force-moved tree
 *       (tests/test_oracle_chain_rules.py)
 */
#ifndef BBS_ORACLE_H
#define BBS_ORACLE_H

#include <stdint.h>

#include "../bbscite_amd/csrc/bbs_rng.h"
#include "../include/bbscite.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_NODES 128

typedef struct {
  uint64_t w[2];
} orc_row;

/* MutationTree state (src/MutationTree.hpp:53-107,896-898). */
typedef struct {
  uint32_t n_nodes;
  orc_row am[ORC_MAX_NODES]; /* ancestor[x]: descendants of x   */
  orc_row dm[ORC_MAX_NODES]; /* descendant[x]: ancestors of x   */
  float beta;
} orc_tree;

/* ModificationParameters (src/MutationTree.hpp:84-89). */
typedef struct {
  uint32_t move_type;
  uint32_t v, w, parent_of_v, parent_of_w, descendant_of_v, nondescendant_of_v;
  float new_beta;
} orc_mod;

/* TreeScorer state (src/TreeScorer.hpp:92-117,212-217). */
typedef struct {
  uint32_t n_cells, n_genes;
  orc_row is_mutated[128], is_known[128];
  float L00, L10; /* log(1-alpha) via double, logf(alpha) */
  float bpriora, bpriorb;
  float lg_ab, lg_a, lg_b; /* log(tgamma(.)) terms of the beta prior */
  int prior_fallback;      /* 1 if tgammaf overflowed (documented deviation) */
} orc_scorer;

/* ---- MutationTree -------------------------------------------------- */
int orc_pruefer_to_parent(const uint32_t *code, uint32_t len, uint32_t *parent);
void orc_parent_to_tree(const uint32_t *parent, uint32_t n_nodes, float beta,
                        orc_tree *t);
int orc_is_parent(const orc_tree *t, uint32_t parent, uint32_t child);
uint32_t orc_get_parent(const orc_tree *t, uint32_t node);
int orc_is_ancestor(const orc_tree *t, uint32_t a, uint32_t b);
uint32_t orc_n_descendants(const orc_tree *t, uint32_t node);
uint32_t orc_n_ancestors(const orc_tree *t, uint32_t node);
uint32_t orc_select(const orc_tree *t, uint32_t node, uint32_t index,
                    int want_descendant);
orc_mod orc_realize(const orc_tree *t, const bbs_raw_move *raw);
void orc_update(const orc_tree *old_tree, const orc_mod *mod, orc_tree *out);
int64_t orc_to_graphviz(const orc_tree *t, char *buf, uint64_t cap);
int64_t orc_to_newick(const orc_tree *t, char *buf, uint64_t cap);

/* ---- TreeScorer ---------------------------------------------------- */
void orc_scorer_init(orc_scorer *s, float alpha_mean, float beta_mean,
                     float beta_sd, uint32_t n_cells, uint32_t n_genes,
                     const uint64_t *is_mutated, const uint64_t *is_known);
float orc_logscore_beta(const orc_scorer *s, float beta);
/* counts: nullable [n_cells][4] (n00,n01,n10,n11 at the first maximal node);
 * best_node: nullable [n_cells]. with_prior: add logscore_beta. */
float orc_logscore_tree(const orc_scorer *s, const orc_tree *t, int with_prior,
                        uint8_t *counts, uint32_t *best_node);

/* ---- chain loop (src/Application.hpp:146-178, 398-496) ------------- */
bbs_move_probs orc_move_probs(const bbs_params *p);
void orc_initial_tree(const bbs_params *p, uint32_t job, uint32_t chain,
                      uint32_t n_genes, orc_tree *t);

typedef struct {
  float best_score;
  uint32_t best_step; /* UINT32_MAX: no record */
  orc_tree best_tree;
  float final_score;
  orc_tree final_tree;
} orc_chain_result;

/* Run one chain. is_first_chain: the chain whose step 0 seeds the global
 * maximum unconditionally (src/Application.hpp:456, i == 0). init: initial
 * tree (NULL: drawn from the seed); init_score: its score (NaN: computed).
 * trace: nullable [n_steps]. */
void orc_run_chain(const bbs_params *p, const orc_scorer *s, uint32_t job,
                   uint32_t chain, int is_first_chain, const orc_tree *init,
                   float init_score, orc_chain_result *res,
                   bbs_step_trace *trace);

/* orc_run_chain that also records every step's proposed tree (accepted or
 * not): prop_am / prop_dm [n_steps][n_nodes][2] words, both or neither
 * NULL. */
void orc_run_chain_proposals(const bbs_params *p, const orc_scorer *s,
                             uint32_t job, uint32_t chain, int is_first_chain,
                             const orc_tree *init, float init_score,
                             orc_chain_result *res, bbs_step_trace *trace,
                             uint64_t *prop_am, uint64_t *prop_dm);

/* Whole run, chain-major (P = 1): the global best as the reference keeps it. */
void orc_run(const bbs_params *p, const orc_scorer *s, uint32_t job,
             orc_tree *best_tree, float *best_score, uint32_t *best_chain,
             uint32_t *best_step);

/* Chains [first, first + count) of one job, each as orc_run_chain (chain 0
 * is the first chain), spread over n_threads host threads (0: all);
 * results[i] is chain first + i. Replays whole batched or served runs. */
void orc_run_chains(const bbs_params *p, const orc_scorer *s, uint32_t job,
                    uint32_t first, uint32_t count, int n_threads,
                    orc_chain_result *results);

/* The shared random stream (bbscite_amd/csrc/bbs_rng.h), exported so that
 * tests can pin it: Philox4x32-10 blocks and raw moves. */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
bbs_raw_move orc_draw_move(uint64_t seed, uint32_t job, uint32_t chain,
                           uint32_t step, uint32_t n_nodes, bbs_move_probs p);
double orc_normal(uint32_t a, uint32_t b);

/* ---- CPU baselines (timed by bench.py) ------------------------------ */
/* Chains run one per thread; returns steps per second. */
double orc_bench_chains(const bbs_params *p, const orc_scorer *s,
                        uint32_t n_chains, uint32_t n_steps, int n_threads,
                        float *checksum);
/* TreeScorer-semantics scorer with CPUTreeScorer's loop nest
 * (cpu_scoring_benchmark/CPUTreeScorer.hpp:165-213: node outer, cell inner,
 * 4 classes x 2 uint64 words with popcount); dm_rows [n_trees][n_nodes][2].
 * Returns trees per second; ll receives the scores without prior. */
double orc_bench_score(const orc_scorer *s, const uint64_t *dm_rows,
                       uint32_t n_trees, float beta, int n_threads, float *ll);

#ifdef __cplusplus
}
#endif

#endif /* BBS_ORACLE_H */
