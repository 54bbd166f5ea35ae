/*
 * CPU oracle: plain-C restatement of the reference's chain-loop hot path.
 * TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py
 * cpu_baseline); never part of the product. See bbs_oracle.h for what pins
 * each function.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off; glibc libm).
 */
#include "bbs_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline int rget(const orc_row *r, uint32_t i) {
  return (int)((r->w[i >> 6] >> (i & 63u)) & 1u);
}

static inline void rset(orc_row *r, uint32_t i, int b) {
  uint64_t m = 1ull << (i & 63u);
  if (b)
    r->w[i >> 6] |= m;
  else
    r->w[i >> 6] &= ~m;
}

/* src/MutationTree.hpp:427-476 */
int orc_pruefer_to_parent(const uint32_t *code, uint32_t len, uint32_t *parent) {
  uint32_t n = len + 2;
  uint32_t degree[ORC_MAX_NODES + 2];
  if (n > ORC_MAX_NODES) return -1;
  for (uint32_t i = 0; i < n; i++) degree[i] = 1;
  for (uint32_t i = 0; i < len; i++) {
    if (code[i] >= n) return -1;
    degree[code[i]]++;
  }
  for (uint32_t i = 0; i < n; i++) parent[i] = n - 1;
  for (uint32_t i = 0; i < len; i++) {
    for (uint32_t j = 0; j < n; j++) {
      if (degree[j] == 1) {
        parent[j] = code[i];
        degree[code[i]]--;
        degree[j]--;
        break;
      }
    }
  }
  for (uint32_t i = 0; i < n; i++) {
    if (degree[i] == 1) {
      parent[i] = n - 1;
      break;
    }
  }
  return 0;
}

/* src/MutationTree.hpp:379-409 */
void orc_parent_to_tree(const uint32_t *parent, uint32_t n, float beta,
                        orc_tree *t) {
  memset(t, 0, sizeof(*t));
  t->n_nodes = n;
  t->beta = beta;
  uint32_t root = n - 1;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t anc = i;
    uint32_t guard = 0;
    while (anc != root && guard++ <= n) {
      rset(&t->am[anc], i, 1);
      rset(&t->dm[i], anc, 1);
      anc = parent[anc];
    }
    rset(&t->am[i], i, 1);
    rset(&t->dm[i], i, 1);
    rset(&t->am[root], i, 1);
    rset(&t->dm[i], root, 1);
  }
}

/* src/MutationTree.hpp:549-575 */
int orc_is_parent(const orc_tree *t, uint32_t parent, uint32_t child) {
  const orc_row *pa = &t->dm[parent], *ca = &t->dm[child];
  if (!rget(ca, parent)) return 0;
  if (parent == child) return child == t->n_nodes - 1;
  for (uint32_t i = 0; i < ORC_MAX_NODES; i++) {
    if (i >= t->n_nodes || i == child) continue;
    if (rget(pa, i) != rget(ca, i)) return 0;
  }
  return 1;
}

/* src/MutationTree.hpp:585-596 */
uint32_t orc_get_parent(const orc_tree *t, uint32_t node) {
  uint32_t parent = 0;
  for (uint32_t j = 0; j < t->n_nodes; j++)
    if (orc_is_parent(t, j, node)) parent = j;
  return parent;
}

/* src/MutationTree.hpp:609-614 */
int orc_is_ancestor(const orc_tree *t, uint32_t a, uint32_t b) {
  return rget(&t->am[a], b);
}

/* src/MutationTree.hpp:661-674 */
uint32_t orc_n_descendants(const orc_tree *t, uint32_t node) {
  uint32_t n = 0;
  for (uint32_t i = 0; i < t->n_nodes; i++) n += (uint32_t)rget(&t->am[node], i);
  return n;
}

/* src/MutationTree.hpp:733-746 */
uint32_t orc_n_ancestors(const orc_tree *t, uint32_t node) {
  uint32_t n = 0;
  for (uint32_t i = 0; i < t->n_nodes; i++) n += (uint32_t)rget(&t->dm[node], i);
  return n;
}

/* src/MutationTree.hpp:840-878 */
uint32_t orc_select(const orc_tree *t, uint32_t node, uint32_t index,
                    int want_descendant) {
  orc_row d = t->am[node];
  if (!want_descendant) {
    d.w[0] = ~d.w[0];
    d.w[1] = ~d.w[1];
  }
  uint32_t result = 0;
  for (uint32_t i = 0; i < t->n_nodes; i++) {
    if (rget(&d, i)) {
      if (index == 0) {
        result = i;
        break;
      }
      index--;
    }
  }
  return result;
}

/* src/MutationTree.hpp:323-363 */
orc_mod orc_realize(const orc_tree *t, const bbs_raw_move *raw) {
  orc_mod m;
  uint32_t v = raw->raw_v, w = raw->raw_w;
  if (orc_is_ancestor(t, v, w)) {
    uint32_t tmp = v;
    v = w;
    w = tmp;
  }
  m.move_type = raw->move_type;
  m.v = v;
  m.w = w;
  m.parent_of_v = orc_get_parent(t, v);
  m.parent_of_w = orc_get_parent(t, w);
  uint32_t nd = orc_n_descendants(t, v);
  uint32_t nn = t->n_nodes - nd;
  uint32_t i_d = (uint32_t)floorf(raw->raw_descendant_of_v * (float)nd);
  uint32_t i_n = (uint32_t)floorf(raw->raw_nondescendant_of_v * (float)nn);
  m.descendant_of_v = orc_select(t, v, i_d, 1);
  m.nondescendant_of_v = orc_select(t, v, i_n, 0);
  float nb = t->beta + raw->beta_jump;
  if (nb < 0) nb = fabsf(nb);
  if (nb > 1) nb = nb - 2 * (nb - 1);
  m.new_beta = nb;
  return m;
}

/* src/MutationTree.hpp:123-307, bit by bit as the reference computes it. */
void orc_update(const orc_tree *old, const orc_mod *mod, orc_tree *out) {
  uint32_t n = old->n_nodes;
  memset(out, 0, sizeof(*out));
  out->n_nodes = n;
  out->beta = old->beta;
  if (mod->move_type == BBS_MOVE_CHANGE_BETA) out->beta = mod->new_beta;
  uint32_t v = mod->v, w = mod->w;
  orc_row v_desc = old->am[v], w_desc = old->am[w];
  orc_row v_anc = old->dm[v], w_anc = old->dm[w];
  uint32_t v_target, w_target;
  switch (mod->move_type) {
  case BBS_MOVE_SWAP_SUBTREES:
    v_target = mod->parent_of_w;
    w_target = rget(&w_desc, v) ? mod->descendant_of_v : mod->parent_of_v;
    break;
  case BBS_MOVE_PRUNE_REATTACH:
    v_target = mod->nondescendant_of_v;
    w_target = 0;
    break;
  default:
    v_target = w_target = 0;
  }
  orc_row vt_anc = old->dm[v_target], wt_anc = old->dm[w_target];

  for (uint32_t x = 0; x < n; x++) {
    orc_row x_desc = old->am[x], x_anc = old->dm[x];
    out->am[x] = x_desc;
    out->dm[x] = x_anc;
    for (uint32_t y = 0; y < ORC_MAX_NODES; y++) {
      int x_to_y, y_to_x;
      switch (mod->move_type) {
      case BBS_MOVE_SWAP_NODES:
        if (x == v) {
          if (y == v) {
            x_to_y = y_to_x = 1;
          } else if (y == w) {
            x_to_y = rget(&w_desc, v);
            y_to_x = rget(&v_desc, w);
          } else {
            x_to_y = rget(&w_desc, y);
            y_to_x = rget(&w_anc, y);
          }
        } else if (x == w) {
          if (y == v) {
            x_to_y = rget(&v_desc, w);
            y_to_x = rget(&w_desc, v);
          } else if (y == w) {
            x_to_y = y_to_x = 1;
          } else {
            x_to_y = rget(&v_desc, y);
            y_to_x = rget(&v_anc, y);
          }
        } else {
          if (y == v) {
            x_to_y = rget(&x_desc, w);
            y_to_x = rget(&x_anc, w);
          } else if (y == w) {
            x_to_y = rget(&x_desc, v);
            y_to_x = rget(&x_anc, v);
          } else {
            x_to_y = rget(&x_desc, y);
            y_to_x = rget(&x_anc, y);
          }
        }
        break;
      case BBS_MOVE_PRUNE_REATTACH:
        if (rget(&v_desc, y))
          x_to_y = rget(&x_desc, v_target) ||
                   (rget(&v_desc, x) && rget(&x_desc, y));
        else
          x_to_y = rget(&x_desc, y);
        if (rget(&v_desc, x))
          y_to_x = rget(&vt_anc, y) || (rget(&x_anc, y) && rget(&v_desc, y));
        else
          y_to_x = rget(&x_anc, y);
        break;
      case BBS_MOVE_SWAP_SUBTREES:
        if (rget(&w_desc, v)) {
          int cx = rget(&v_desc, x) ? 2 : (rget(&w_desc, x) ? 1 : 0);
          int cy = rget(&v_desc, y) ? 2 : (rget(&w_desc, y) ? 1 : 0);
          if (cx == cy || (cx == 0 && (cy == 1 || cy == 2)))
            x_to_y = rget(&x_desc, y);
          else if (cx == 2 && cy == 1)
            x_to_y = rget(&x_desc, w_target);
          else
            x_to_y = 0;
          if (cx == cy || (cy == 0 && (cx == 1 || cx == 2)))
            y_to_x = rget(&x_anc, y);
          else if (cy == 2 && cx == 1)
            y_to_x = rget(&wt_anc, y);
          else
            y_to_x = 0;
        } else {
          if (rget(&v_desc, y) && !rget(&w_desc, y))
            x_to_y = rget(&x_desc, v_target) ||
                     (rget(&v_desc, x) && rget(&x_desc, y));
          else if (!rget(&v_desc, y) && rget(&w_desc, y))
            x_to_y = rget(&x_desc, w_target) ||
                     (rget(&w_desc, x) && rget(&x_desc, y));
          else
            x_to_y = rget(&x_desc, y);
          if (rget(&v_desc, x) && !rget(&w_desc, x))
            y_to_x = rget(&vt_anc, y) || (rget(&v_desc, y) && rget(&x_anc, y));
          else if (!rget(&v_desc, x) && rget(&w_desc, x))
            y_to_x = rget(&wt_anc, y) || (rget(&w_desc, y) && rget(&x_anc, y));
          else
            y_to_x = rget(&x_anc, y);
        }
        break;
      default: /* ChangeBeta */
        x_to_y = rget(&x_desc, y);
        y_to_x = rget(&x_anc, y);
      }
      rset(&out->am[x], y, x_to_y);
      rset(&out->dm[x], y, y_to_x);
    }
  }
}

/* ---- text outputs --------------------------------------------------- */

typedef struct {
  char *buf;
  uint64_t cap, len;
} orc_sb;

static void sb_put(orc_sb *sb, const char *s) {
  uint64_t l = strlen(s);
  for (uint64_t i = 0; i < l; i++) {
    if (sb->len + 1 < sb->cap) sb->buf[sb->len] = s[i];
    sb->len++;
  }
  if (sb->cap) sb->buf[sb->len + 1 < sb->cap ? sb->len : sb->cap - 1] = 0;
}

/* src/MutationTree.hpp:799-809 */
int64_t orc_to_graphviz(const orc_tree *t, char *buf, uint64_t cap) {
  orc_sb sb = {buf, cap, 0};
  char line[64];
  sb_put(&sb, "digraph G {\n");
  sb_put(&sb, "node [color=deeppink4, style=filled, fontcolor=white];\n");
  for (uint32_t i = 0; i + 1 < t->n_nodes; i++) {
    snprintf(line, sizeof line, "%u -> %u;\n", orc_get_parent(t, i), i);
    sb_put(&sb, line);
  }
  sb_put(&sb, "}\n");
  return (int64_t)sb.len;
}

static void newick_rec(const orc_tree *t, const uint32_t *parent, uint32_t node,
                       orc_sb *sb, int depth) {
  char num[16];
  int first = 1, any = 0;
  if (depth > ORC_MAX_NODES) return;
  for (uint32_t c = 0; c + 1 < t->n_nodes; c++) {
    if (parent[c] != node) continue;
    sb_put(sb, first ? "(" : ",");
    first = 0;
    any = 1;
    newick_rec(t, parent, c, sb, depth + 1);
  }
  if (any) sb_put(sb, ")");
  snprintf(num, sizeof num, "%u", node);
  sb_put(sb, num);
}

/* src/MutationTree.hpp:818-837,880-894 */
int64_t orc_to_newick(const orc_tree *t, char *buf, uint64_t cap) {
  orc_sb sb = {buf, cap, 0};
  uint32_t parent[ORC_MAX_NODES];
  for (uint32_t i = 0; i + 1 < t->n_nodes; i++) parent[i] = orc_get_parent(t, i);
  parent[t->n_nodes - 1] = UINT32_MAX;
  newick_rec(t, parent, t->n_nodes - 1, &sb, 0);
  sb_put(&sb, "\n");
  return (int64_t)sb.len;
}

/* ---- TreeScorer ------------------------------------------------------ */

static float orc_lgamma_term(float x, int *fallback) {
  float g = tgammaf(x);
  if (isinf(g) && g > 0 && x > 0) {
    /* Deviation: tgammaf overflows (x > ~35.04) and the reference's float
     * prior becomes inf - inf = NaN; use log-gamma in double instead. */
    *fallback = 1;
    return (float)lgamma((double)x);
  }
  return logf(g);
}

/* src/TreeScorer.hpp:92-117 */
void orc_scorer_init(orc_scorer *s, float alpha_mean, float beta_mean,
                     float beta_sd, uint32_t n_cells, uint32_t n_genes,
                     const uint64_t *is_mutated, const uint64_t *is_known) {
  memset(s, 0, sizeof(*s));
  s->n_cells = n_cells;
  s->n_genes = n_genes;
  for (uint32_t c = 0; c < n_cells; c++) {
    s->is_mutated[c].w[0] = is_mutated[2 * c];
    s->is_mutated[c].w[1] = is_mutated[2 * c + 1];
    s->is_known[c].w[0] = is_known[2 * c];
    s->is_known[c].w[1] = is_known[2 * c + 1];
  }
  s->L00 = (float)log(1.0 - (double)alpha_mean);
  s->L10 = logf(alpha_mean);
  double a = ((double)(1.0f - beta_mean) * pow((double)beta_mean, 2.0) /
              pow((double)beta_sd, 2.0)) -
             (double)beta_mean;
  s->bpriora = (float)a;
  s->bpriorb = s->bpriora * ((1.0f / beta_mean) - 1.0f);
  s->prior_fallback = 0;
  s->lg_ab = orc_lgamma_term(s->bpriora + s->bpriorb, &s->prior_fallback);
  s->lg_a = orc_lgamma_term(s->bpriora, &s->prior_fallback);
  s->lg_b = orc_lgamma_term(s->bpriorb, &s->prior_fallback);
}

/* src/TreeScorer.hpp:122-126 */
float orc_logscore_beta(const orc_scorer *s, float beta) {
  return s->lg_ab + (s->bpriora - 1.0f) * logf(beta) +
         (s->bpriorb - 1.0f) * logf(1.0f - beta) - s->lg_a - s->lg_b;
}

/* src/TreeScorer.hpp:131-210 */
float orc_logscore_tree(const orc_scorer *s, const orc_tree *t, int with_prior,
                        uint8_t *counts, uint32_t *best_node) {
  float L[2][2];
  L[0][0] = s->L00;
  L[1][0] = s->L10;
  L[0][1] = logf(t->beta);
  L[1][1] = (float)log(1.0 - (double)t->beta);
  float individual[128];
  uint8_t cnt[128][4];
  uint32_t arg[128];
  for (uint32_t node = 0; node < s->n_genes + 1; node++) {
    orc_row A = t->dm[node];
    for (uint32_t c = 0; c < s->n_cells; c++) {
      orc_row M = s->is_mutated[c], K = s->is_known[c];
      float score = 0.0f;
      uint8_t n4[4];
      for (int post = 0; post < 2; post++) {
        for (int prior = 0; prior < 2; prior++) {
          uint32_t n = 0;
          for (int wi = 0; wi < 2; wi++) {
            uint64_t occ = K.w[wi] & (post ? M.w[wi] : ~M.w[wi]) &
                           (prior ? A.w[wi] : ~A.w[wi]);
            n += (uint32_t)__builtin_popcountll(occ);
          }
          n4[post * 2 + prior] = (uint8_t)n;
          score += (float)n * L[post][prior];
        }
      }
      int take;
      if (node == 0)
        take = 1;
      else
        take = individual[c] < score; /* std::max(old, new) keeps old on ties */
      if (take) {
        individual[c] = score;
        memcpy(cnt[c], n4, 4);
        arg[c] = node;
      }
    }
  }
  float tree_score = 0.0f;
  for (uint32_t c = 0; c < s->n_cells; c++) tree_score += individual[c];
  if (counts) memcpy(counts, cnt, 4u * s->n_cells);
  if (best_node) memcpy(best_node, arg, 4u * s->n_cells);
  if (with_prior) return tree_score + orc_logscore_beta(s, t->beta);
  return tree_score;
}

/* ---- chain loop ------------------------------------------------------ */

/* src/Parameters.hpp:451-541,582-584,609-612; src/RawMoveDistribution.hpp:121-126 */
bbs_move_probs orc_move_probs(const bbs_params *p) {
  float sum = p->prob_beta_change + p->prob_prune_n_reattach +
              p->prob_swap_nodes + p->prob_swap_subtrees;
  float pb = p->prob_beta_change / sum;
  float pp = p->prob_prune_n_reattach / sum;
  float ps = p->prob_swap_nodes / sum;
  bbs_move_probs m;
  m.t_beta = pb;
  m.t_prune = pb + pp;
  m.t_swap = pb + pp + ps;
  m.beta_jump_sd = p->beta_sd / p->beta_jump_scaling_chi;
  return m;
}

/* src/Application.hpp:149-164; src/MutationTree.hpp:492-510 */
void orc_initial_tree(const bbs_params *p, uint32_t job, uint32_t chain,
                      uint32_t n_genes, orc_tree *t) {
  uint32_t n = n_genes + 1;
  uint32_t code[ORC_MAX_NODES], parent[ORC_MAX_NODES];
  for (uint32_t i = 0; i + 2 < n; i++) {
    bbs_u32x4 b =
        bbs_philox_block(p->seed, job, chain, i >> 2, BBS_PURPOSE_INIT_TREE);
    code[i] = bbs_uint_below(b.x[i & 3u], n);
  }
  orc_pruefer_to_parent(code, n - 2, parent);
  orc_parent_to_tree(parent, n, p->beta_mean, t);
}

/* src/Application.hpp:425-487 for one chain */
void orc_run_chain(const bbs_params *p, const orc_scorer *s, uint32_t job,
                   uint32_t chain, int is_first_chain, const orc_tree *init,
                   float init_score, orc_chain_result *res,
                   bbs_step_trace *trace) {
  orc_run_chain_proposals(p, s, job, chain, is_first_chain, init, init_score, res,
                          trace, NULL, NULL);
}

void orc_run_chain_proposals(const bbs_params *p, const orc_scorer *s,
                             uint32_t job, uint32_t chain, int is_first_chain,
                             const orc_tree *init, float init_score,
                             orc_chain_result *res, bbs_step_trace *trace,
                             uint64_t *prop_am, uint64_t *prop_dm) {
  static __thread orc_tree cur, prop;
  uint32_t n = s->n_genes + 1;
  if (init)
    cur = *init;
  else
    orc_initial_tree(p, job, chain, s->n_genes, &cur);
  float cur_score =
      isnan(init_score) ? orc_logscore_tree(s, &cur, 1, NULL, NULL) : init_score;
  bbs_move_probs probs = orc_move_probs(p);
  res->best_score = -INFINITY;
  res->best_step = UINT32_MAX;
  for (uint32_t step = 0; step < p->n_steps; step++) {
    bbs_raw_move raw = bbs_draw_move(p->seed, job, chain, step, n, probs);
    orc_mod mod = orc_realize(&cur, &raw);
    orc_update(&cur, &mod, &prop);
    for (uint32_t x = 0; prop_am && x < n; x++) {
      uint64_t *a = prop_am + 2u * ((size_t)step * n + x);
      uint64_t *d = prop_dm + 2u * ((size_t)step * n + x);
      a[0] = prop.am[x].w[0];
      a[1] = prop.am[x].w[1];
      d[0] = prop.dm[x].w[0];
      d[1] = prop.dm[x].w[1];
    }
    float prop_score = orc_logscore_tree(s, &prop, 1, NULL, NULL);
    float corr;
    if (raw.move_type == BBS_MOVE_SWAP_SUBTREES &&
        orc_is_ancestor(&cur, mod.w, mod.v))
      corr = (float)orc_n_descendants(&cur, mod.v) /
             (float)orc_n_descendants(&cur, mod.w);
    else
      corr = 1.0f;
    float acc = corr * expf((prop_score - cur_score) * p->gamma);
    int accept = acc > raw.acceptance_level;
    int new_max = (is_first_chain && step == 0) || prop_score > res->best_score;
    if (new_max) {
      res->best_score = prop_score;
      res->best_step = step;
      res->best_tree = prop;
    }
    if (trace) {
      bbs_step_trace *tr = &trace[step];
      tr->move_type = mod.move_type;
      tr->v = mod.v;
      tr->w = mod.w;
      tr->parent_of_v = mod.parent_of_v;
      tr->parent_of_w = mod.parent_of_w;
      tr->descendant_of_v = mod.descendant_of_v;
      tr->nondescendant_of_v = mod.nondescendant_of_v;
      tr->new_beta = mod.new_beta;
      tr->proposed_score = prop_score;
      tr->flags = (accept ? 1u : 0u) | (new_max ? 2u : 0u);
    }
    if (accept) {
      cur = prop;
      cur_score = prop_score;
    }
  }
  res->final_score = cur_score;
  res->final_tree = cur;
}

/* src/Application.hpp:398-496 with P = 1 (chain-major order) */
void orc_run(const bbs_params *p, const orc_scorer *s, uint32_t job,
             orc_tree *best_tree, float *best_score, uint32_t *best_chain,
             uint32_t *best_step) {
  orc_chain_result *res = (orc_chain_result *)malloc(sizeof(orc_chain_result));
  *best_score = -INFINITY;
  *best_chain = UINT32_MAX;
  *best_step = UINT32_MAX;
  for (uint32_t c = 0; c < p->n_chains; c++) {
    orc_run_chain(p, s, job, c, c == 0, NULL, NAN, res, NULL);
    if (res->best_step == UINT32_MAX) continue;
    /* Chain 0's step 0 seeds the maximum (i == 0); later chains replace it
     * only with a strictly greater proposal, the first one in step order. */
    if (c == 0 || res->best_score > *best_score) {
      *best_score = res->best_score;
      *best_tree = res->best_tree;
      *best_chain = c;
      *best_step = res->best_step;
    }
  }
  free(res);
}

void orc_run_chains(const bbs_params *p, const orc_scorer *s, uint32_t job,
                    uint32_t first, uint32_t count, int n_threads,
                    orc_chain_result *results) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (uint32_t i = 0; i < count; i++)
    orc_run_chain(p, s, job, first + i, first + i == 0u, NULL, NAN, &results[i], NULL);
}

void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  bbs_u32x4 r = bbs_philox4x32_10(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  memcpy(out, r.x, sizeof(r.x));
}

bbs_raw_move orc_draw_move(uint64_t seed, uint32_t job, uint32_t chain,
                           uint32_t step, uint32_t n_nodes, bbs_move_probs p) {
  return bbs_draw_move(seed, job, chain, step, n_nodes, p);
}

double orc_normal(uint32_t a, uint32_t b) { return bbs_normal_from_u32(a, b); }

/* ---- CPU baselines ---------------------------------------------------- */

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

double orc_bench_chains(const bbs_params *p, const orc_scorer *s,
                        uint32_t n_chains, uint32_t n_steps, int n_threads,
                        float *checksum) {
  bbs_params q = *p;
  q.n_steps = n_steps;
  float sum = 0.0f;
  double t0 = now_s();
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : sum)
#endif
  for (uint32_t c = 0; c < n_chains; c++) {
    orc_chain_result *res = (orc_chain_result *)malloc(sizeof(*res));
    orc_run_chain(&q, s, 0, c, c == 0, NULL, NAN, res, NULL);
    sum += res->final_score;
    free(res);
  }
  double t1 = now_s();
  if (checksum) *checksum = sum;
  return (double)n_chains * (double)n_steps / (t1 - t0);
}

/* cpu_scoring_benchmark/CPUTreeScorer.hpp:165-213 loop nest with
 * TreeScorer semantics (per-cell maximum over nodes, ordered cell sum). */
static float cpu_tree_scorer(const orc_scorer *s, const uint64_t *dm,
                             const float L[2][2]) {
  float individual[128];
  uint32_t n_nodes = s->n_genes + 1;
  for (uint32_t node = 0; node < n_nodes; node++) {
    uint64_t a0 = dm[2 * node], a1 = dm[2 * node + 1];
    for (uint32_t c = 0; c < s->n_cells; c++) {
      float score = 0.0f;
      for (int post = 0; post < 2; post++) {
        for (int prior = 0; prior < 2; prior++) {
          uint64_t o0 = s->is_known[c].w[0], o1 = s->is_known[c].w[1];
          o0 &= post ? s->is_mutated[c].w[0] : ~s->is_mutated[c].w[0];
          o1 &= post ? s->is_mutated[c].w[1] : ~s->is_mutated[c].w[1];
          o0 &= prior ? a0 : ~a0;
          o1 &= prior ? a1 : ~a1;
          uint32_t n =
              (uint32_t)(__builtin_popcountll(o0) + __builtin_popcountll(o1));
          score += (float)n * L[post][prior];
        }
      }
      if (node == 0 || individual[c] < score) individual[c] = score;
    }
  }
  float tree = 0.0f;
  for (uint32_t c = 0; c < s->n_cells; c++) tree += individual[c];
  return tree;
}

double orc_bench_score(const orc_scorer *s, const uint64_t *dm_rows,
                       uint32_t n_trees, float beta, int n_threads, float *ll) {
  float L[2][2];
  L[0][0] = s->L00;
  L[1][0] = s->L10;
  L[0][1] = logf(beta);
  L[1][1] = (float)log(1.0 - (double)beta);
  uint32_t n_nodes = s->n_genes + 1;
  double t0 = now_s();
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(static)
#endif
  for (uint32_t t = 0; t < n_trees; t++)
    ll[t] = cpu_tree_scorer(s, dm_rows + (uint64_t)t * n_nodes * 2u, L);
  double t1 = now_s();
  return (double)n_trees / (t1 - t0);
}
