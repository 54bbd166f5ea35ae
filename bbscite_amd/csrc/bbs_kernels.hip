/*
 * HIP kernels of the chain search for gfx950.
 *
 *   chain_kernel   one wavefront per Markov chain; the chain's two bit
 *                  matrices stay in VGPRs for the whole launch. One
 *                  iteration = the WorkKernel loop body of
 *                  src/Application.hpp:412-487: draw and realise a move,
 *                  update the bit matrices, score the proposal, Metropolis
 *                  step, best tracking.
 *   score_kernel   one wavefront per tree: TreeScorer::logscore_tree
 *                  (src/TreeScorer.hpp:131-210) over many trees.
 *   random_trees_kernel  one wavefront per tree: uniform Pruefer code ->
 *                  parent vector -> ancestor rows
 *                  (src/MutationTree.hpp:379-409,427-510).
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "bbs_device.h"
#include "bbs_kernels.h"

namespace bbs {

/* The register layouts instantiated: (NS node slots, CS cell slots, W gene
 * words) for every gene/cell count up to 127 x 128. */
#define BBS_SHAPES(X)                                                          \
  X(1, 1, 1) X(1, 1, 2) X(1, 2, 1) X(1, 2, 2) X(1, 3, 1) X(1, 3, 2) X(2, 1, 2) \
      X(2, 1, 3) X(2, 1, 4) X(2, 2, 2) X(2, 2, 3) X(2, 2, 4) X(2, 3, 2)         \
      X(2, 3, 3) X(2, 3, 4)
/* the layouts with full cell slots (the counting scorer) */
#define BBS_FULL_SHAPES(X)                                                     \
  X(1, 1, 1) X(1, 1, 2) X(1, 2, 1) X(1, 2, 2) X(2, 1, 2) X(2, 1, 3) X(2, 1, 4) \
      X(2, 2, 2) X(2, 2, 3) X(2, 2, 4)
#define BBS_SHAPE_CODE(NS_, CS_, W_) ((NS_)*100 + (CS_)*10 + (W_))

__device__ void random_tree_wave(const TreeGenArgs &a, uint32_t tree);

/* ------------------------------------------------------------------ moves */

/* The proposal of one move for the rows of this lane
 * (src/MutationTree.hpp:123-307 restated on 64-bit words). */
template <int NS>
struct MoveCtx {
  uint32_t type, v, w;
  bool related;                 /* w is an ancestor of v (Dw[v])          */
  uint64_t Dv[NS], Dw[NS];      /* descendants of v, w                    */
  uint64_t Av[NS], Aw[NS];      /* ancestors of v, w                      */
  uint32_t v_target, w_target;
  uint64_t At_v[NS], At_w[NS];  /* ancestors of v_target, w_target        */
};

template <int NS>
__device__ __forceinline__ void apply_move(const MoveCtx<NS> &m,
                                           const uint64_t (&am)[NS][NS],
                                           const uint64_t (&dm)[NS][NS],
                                           uint64_t (&pam)[NS][NS],
                                           uint64_t (&pdm)[NS][NS]) {
  const uint32_t lane = lane_id();
  const uint64_t lanebit = 1ull << lane;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const uint32_t x = lane + 64u * (uint32_t)k;
    const bool dvx = (m.Dv[k] & lanebit) != 0;
    const bool dwx = (m.Dw[k] & lanebit) != 0;
    switch (m.type) {
    case BBS_MOVE_SWAP_NODES: {
      /* relabel v <-> w: row x takes row sigma(x) with bits v, w exchanged */
      const bool is_v = x == m.v, is_w = x == m.w;
      uint64_t sa[NS], sd[NS];
#pragma unroll
      for (int wd = 0; wd < NS; wd++) {
        sa[wd] = is_v ? m.Dw[wd] : (is_w ? m.Dv[wd] : am[k][wd]);
        sd[wd] = is_v ? m.Aw[wd] : (is_w ? m.Av[wd] : dm[k][wd]);
      }
      const bool a_v = row_bit<NS>(sa, m.v), a_w = row_bit<NS>(sa, m.w);
      const bool d_v = row_bit<NS>(sd, m.v), d_w = row_bit<NS>(sd, m.w);
#pragma unroll
      for (int wd = 0; wd < NS; wd++) {
        const uint64_t mv = ((m.v >> 6) == (uint32_t)wd) ? 1ull << (m.v & 63u) : 0ull;
        const uint64_t mw = ((m.w >> 6) == (uint32_t)wd) ? 1ull << (m.w & 63u) : 0ull;
        pam[k][wd] = (sa[wd] & ~(mv | mw)) | (a_w ? mv : 0ull) | (a_v ? mw : 0ull);
        pdm[k][wd] = (sd[wd] & ~(mv | mw)) | (d_w ? mv : 0ull) | (d_v ? mw : 0ull);
      }
      break;
    }
    case BBS_MOVE_PRUNE_REATTACH: {
      /* x->y for y in subtree(v): x->t or (v->x and x->y)
       * y->x for x in subtree(v): y->t or (y->x and v->y) */
      const bool amx_t = row_bit<NS>(am[k], m.v_target);
#pragma unroll
      for (int wd = 0; wd < NS; wd++) {
        const uint64_t a = am[k][wd], d = dm[k][wd], dv = m.Dv[wd];
        pam[k][wd] = (dv & ((amx_t ? ~0ull : 0ull) | (dvx ? a : 0ull))) | (~dv & a);
        pdm[k][wd] = dvx ? (m.At_v[wd] | (d & dv)) : d;
      }
      break;
    }
    case BBS_MOVE_SWAP_SUBTREES: {
      const bool amx_vt = row_bit<NS>(am[k], m.v_target);
      const bool amx_wt = row_bit<NS>(am[k], m.w_target);
      if (m.related) {
        /* classes: 2 = subtree(v), 1 = subtree(w) \ subtree(v), 0 = rest */
#pragma unroll
        for (int wd = 0; wd < NS; wd++) {
          const uint64_t a = am[k][wd], d = dm[k][wd];
          const uint64_t c2 = m.Dv[wd], c1 = m.Dw[wd] & ~m.Dv[wd];
          const uint64_t c0 = ~(m.Dw[wd] | m.Dv[wd]);
          if (dvx) {
            pam[k][wd] = (a & c2) | (amx_wt ? c1 : 0ull);
            pdm[k][wd] = d & ~c1;
          } else if (dwx) {
            pam[k][wd] = a & c1;
            pdm[k][wd] = (d & ~c2) | (m.At_w[wd] & c2);
          } else {
            pam[k][wd] = a;
            pdm[k][wd] = d & c0;
          }
        }
      } else {
#pragma unroll
        for (int wd = 0; wd < NS; wd++) {
          const uint64_t a = am[k][wd], d = dm[k][wd];
          const uint64_t dv = m.Dv[wd], dw = m.Dw[wd];
          const uint64_t P = dv & ~dw, Q = ~dv & dw, R = ~(P | Q);
          pam[k][wd] = (P & ((amx_vt ? ~0ull : 0ull) | (dvx ? a : 0ull))) |
                       (Q & ((amx_wt ? ~0ull : 0ull) | (dwx ? a : 0ull))) |
                       (R & a);
          if (dvx && !dwx)
            pdm[k][wd] = m.At_v[wd] | (dv & d);
          else if (!dvx && dwx)
            pdm[k][wd] = m.At_w[wd] | (dw & d);
          else
            pdm[k][wd] = d;
        }
      }
      break;
    }
    default: /* ChangeBeta */
#pragma unroll
      for (int wd = 0; wd < NS; wd++) {
        pam[k][wd] = am[k][wd];
        pdm[k][wd] = dm[k][wd];
      }
    }
  }
}

/* ------------------------------------------------------------ proposals */

/* The raw move of the step held by lane sl of the Philox blocks (ra, rb) and
 * the beta jumps of the same steps (src/RawMoveDistribution.hpp:97-155). */
__device__ __forceinline__ bbs_raw_move draw_move(const bbs_u32x4 &ra, const bbs_u32x4 &rb,
                                                  float jumps, uint32_t sl, uint32_t n_nodes,
                                                  const bbs_move_probs &probs) {
  bbs_u32x4 wa, wb;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    wa.x[i] = readlane32(ra.x[i], sl);
    wb.x[i] = readlane32(rb.x[i], sl);
  }
  bbs_raw_move raw = bbs_move_without_jump(wa, wb, n_nodes, probs);
  if (raw.move_type == BBS_MOVE_CHANGE_BETA)
    raw.beta_jump = __uint_as_float(readlane32(__float_as_uint(jumps), sl));
  return raw;
}

/* One step's proposal for the current tree (am, dm) and beta. */
template <int NS>
struct Proposal {
  MoveCtx<NS> m;                  /* m.v, m.w after the swap of v and w    */
  uint32_t nd_v, nd_w;            /* descendant counts of v, w            */
  uint32_t parent_v, parent_w, desc_v, nondesc_v;
  float nb;                       /* the reflected new beta               */
  uint64_t pam[NS][NS], pdm[NS][NS];
};

/* Realisation (src/MutationTree.hpp:323-363) and the proposed rows
 * (src/MutationTree.hpp:123-307). Without TRACE only what the drawn move
 * type uses is computed (a beta move needs no tree data, prune and reattach
 * a random non-descendant, a node swap the four rows, a subtree swap the
 * parents and a random descendant); every value computed equals the full
 * realisation's. */
template <int NS, bool TRACE>
__device__ __forceinline__ void propose(const bbs_raw_move &raw, uint32_t N, float cur_beta,
                                        const uint64_t (&am)[NS][NS],
                                        const uint64_t (&dm)[NS][NS], Proposal<NS> &P) {
  MoveCtx<NS> &m = P.m;
  m.type = raw.move_type;
  m.related = false;
  m.v_target = 0u;
  m.w_target = 0u;
  uint32_t v = raw.raw_v, w = raw.raw_w;
  uint32_t nd_v = 0u, nd_w = 0u;
  uint32_t parent_v = 0u, parent_w = 0u, desc_v = 0u, nondesc_v = 0u;
  const bool subtrees = m.type == BBS_MOVE_SWAP_SUBTREES;
  const bool prune = m.type == BBS_MOVE_PRUNE_REATTACH;
  if (TRACE || m.type != BBS_MOVE_CHANGE_BETA) {
    /* the four rows at once, then the swap of src/MutationTree.hpp:326-330
     * (v an ancestor of w) as a selection */
    uint64_t D0[NS], D1[NS], A0[NS], A1[NS];
    fetch_row<NS>(am, v, D0);
    fetch_row<NS>(am, w, D1);
    fetch_row<NS>(dm, v, A0);
    fetch_row<NS>(dm, w, A1);
    const bool sw = row_bit<NS>(D0, w);
    const uint32_t t = sw ? w : v;
    w = sw ? v : w;
    v = t;
#pragma unroll
    for (int wd = 0; wd < NS; wd++) {
      m.Dv[wd] = sw ? D1[wd] : D0[wd];
      m.Dw[wd] = sw ? D0[wd] : D1[wd];
      m.Av[wd] = sw ? A1[wd] : A0[wd];
      m.Aw[wd] = sw ? A0[wd] : A1[wd];
      nd_v += (uint32_t)__popcll(m.Dv[wd]);
    }
  } else {
#pragma unroll
    for (int wd = 0; wd < NS; wd++) m.Dv[wd] = m.Dw[wd] = m.Av[wd] = m.Aw[wd] = 0ull;
  }
  m.v = v;
  m.w = w;
  if (TRACE || subtrees) {
#pragma unroll
    for (int wd = 0; wd < NS; wd++) nd_w += (uint32_t)__popcll(m.Dw[wd]);
    m.related = row_bit<NS>(m.Dw, v);
    parent_w = parent_of<NS>(dm, m.Aw, w, N);
  }
  if (TRACE || (subtrees && !m.related)) parent_v = parent_of<NS>(dm, m.Av, v, N);
  if (TRACE || (subtrees && m.related)) {
    const uint32_t i_d = (uint32_t)floorf(raw.raw_descendant_of_v * (float)nd_v);
    desc_v = select_bit<NS>(m.Dv, i_d, N);
  }
  if (TRACE || prune) {
    const uint32_t nn_v = N - nd_v;
    const uint32_t i_n = (uint32_t)floorf(raw.raw_nondescendant_of_v * (float)nn_v);
    uint64_t ndv[NS];
#pragma unroll
    for (int wd = 0; wd < NS; wd++) ndv[wd] = ~m.Dv[wd];
    nondesc_v = select_bit<NS>(ndv, i_n, N);
  }
  float nb = cur_beta + raw.beta_jump;
  if (nb < 0.0f) nb = fabsf(nb);
  if (nb > 1.0f) nb = nb - 2.0f * (nb - 1.0f);

  if (subtrees) {
    m.v_target = parent_w;
    m.w_target = m.related ? desc_v : parent_v;
  } else if (prune) {
    m.v_target = nondesc_v;
  }
  if (subtrees || prune) fetch_row<NS>(dm, m.v_target, m.At_v);
  if (subtrees) fetch_row<NS>(dm, m.w_target, m.At_w);

  apply_move<NS>(m, am, dm, P.pam, P.pdm);
  P.nd_v = nd_v;
  P.nd_w = nd_w;
  P.parent_v = parent_v;
  P.parent_w = parent_w;
  P.desc_v = desc_v;
  P.nondesc_v = nondesc_v;
  P.nb = nb;
}

/* The rows of this lane's nodes (n_nodes of them) to rows[2 x][2]. */
template <int NS>
__device__ __forceinline__ void write_rows(uint64_t *am_out, uint64_t *dm_out,
                                           const uint64_t (&am)[NS][NS],
                                           const uint64_t (&dm)[NS][NS], uint32_t N) {
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const uint32_t x = lane_id() + 64u * (uint32_t)k;
    if (x < N) {
#pragma unroll
      for (int wd = 0; wd < NS; wd++) {
        am_out[2u * x + (uint32_t)wd] = am[k][wd];
        dm_out[2u * x + (uint32_t)wd] = dm[k][wd];
      }
      if (NS == 1) {
        am_out[2u * x + 1u] = 0ull;
        dm_out[2u * x + 1u] = 0ull;
      }
    }
  }
}

/* The Hastings correction of src/Application.hpp:443-450: nd(v)/nd(w) for a
 * subtree swap with w an ancestor of v (after the swap), else 1. */
template <int NS>
__device__ __forceinline__ float acceptance_corr(const Proposal<NS> &P) {
  return (P.m.type == BBS_MOVE_SWAP_SUBTREES && P.m.related) ? (float)P.nd_v / (float)P.nd_w
                                                             : 1.0f;
}

template <int NS>
__device__ __forceinline__ bbs_step_trace step_trace(const Proposal<NS> &P, float prop_score,
                                                     bool accept, bool new_max) {
  bbs_step_trace tr;
  tr.move_type = P.m.type;
  tr.v = P.m.v;
  tr.w = P.m.w;
  tr.parent_of_v = P.parent_v;
  tr.parent_of_w = P.parent_w;
  tr.descendant_of_v = P.desc_v;
  tr.nondescendant_of_v = P.nondesc_v;
  tr.new_beta = P.nb;
  tr.proposed_score = prop_score;
  tr.flags = (accept ? 1u : 0u) | (new_max ? 2u : 0u);
  return tr;
}

/* ------------------------------------------------------------ chain body */

/* Everything one wavefront needs to advance one chain; all pointers point
 * at this chain's own rows / scalars. */
struct ChainView {
  const uint64_t *cells;
  uint64_t *am, *dm, *best_am, *best_dm; /* n_nodes rows each             */
  float *beta, *score, *best_beta, *best_score;
  uint32_t *best_step;
  bbs_step_trace *trace; /* [step] or null                        */
  uint64_t *prop_am, *prop_dm; /* [step][n_nodes][2] or null: proposals */
  uint64_t seed;
  uint32_t job, chain; /* Philox counter words                          */
  uint32_t step_begin, step_end;
  uint32_t n_cells, n_nodes;
  bool is_first_chain; /* its step 0 seeds the maximum (i == 0)        */
  bool score_init;
  float L00, L10, gamma;
  PriorArgs prior;
  bbs_move_probs probs;
};

/* Steps [step_begin, step_end) of one chain: the WorkKernel loop body of
 * src/Application.hpp:412-487 with the state resident in VGPRs. With INCR
 * the pair scores of the current tree are kept in the LDS table `tab`: a
 * tree move rescores only the nodes whose ancestor row changed and takes
 * the other nodes' scores from the table, which is updated when the move is
 * accepted; beta moves rescore every pair. With MFMA those full scorings
 * count pairs on the matrix cores where the dataset fills its tiles (the
 * plain kernel; the batched kernel, mostly small datasets, keeps its
 * registers for the vector path). */
template <int NS, int CS, int W, bool TRACE, bool INCR, bool MFMA, bool SPL = false,
          bool HALF = false>
__device__ __forceinline__ void run_chain(const ChainView &a, WaveLds &lds, TableView tab,
                                          const CellSplit &split = CellSplit{nullptr, 0u, 0u},
                                          bool leader = true) {
  static_assert(!SPL || (CS == 1 && INCR && !MFMA), "a split chain: one cell slot, table");
  static_assert(!HALF || (CS == 2 && INCR && !MFMA && !SPL),
                "a half table: two full cell slots, the table of slot 0");
  const uint32_t lane = lane_id();
  const uint32_t N = a.n_nodes;
  /* SPL: this wave scores the cells a.cells[0, a.n_cells) of a chain split
   * over the waves of its workgroup; every wave runs the same move, and only
   * the leader writes to memory. Each scoring sums in the next of the two
   * shared buffers (split.shared is the first). */
  uint32_t flip = 0u;
  auto next_split = [&]() {
    CellSplit sp = split;
    sp.shared = split.shared + 128u * (flip & 1u);
    flip++;
    return sp;
  };

  Cells<CS, W> cells;
  load_cells<CS, W>(a.cells, a.n_cells, cells);
  /* With the table, full scorings (the first one and beta moves) count
   * pairs on the matrix cores when the dataset fills its 32 x 32 tiles
   * well; padded pairs cost the vector unit as much as real ones. */
  const uint32_t tiles_area = 1024u * ((N + 31u) >> 5) * ((a.n_cells + 31u) >> 5);
  const bool use_mfma = MFMA && INCR && 10u * N * a.n_cells >= 7u * tiles_area;
  MfmaCells<CS, W> mc;
  if (use_mfma) load_mfma_cells<CS, W>(a.cells, a.n_cells, mc);

  uint64_t am[NS][NS], dm[NS][NS];
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const uint32_t x = lane + 64u * (uint32_t)k;
#pragma unroll
    for (int wd = 0; wd < NS; wd++) {
      am[k][wd] = x < N ? a.am[2u * x + (uint32_t)wd] : 0ull;
      dm[k][wd] = x < N ? a.dm[2u * x + (uint32_t)wd] : 0ull;
    }
  }

  float cur_beta = *a.beta;
  LogProbs L;
  L.L00 = a.L00;
  L.L10 = a.L10;
  L.L01 = bbs_logf(cur_beta);
  L.L11 = bbs_logf_of_double(1.0 - (double)cur_beta);
  float cur_prior = logscore_beta(a.prior, cur_beta);
  uint32_t unused01[2], unused11[2];
  float cur_score;
  if (a.score_init || INCR) {
    float t;
    if (use_mfma) {
      store_rows<NS>(dm, N, lds);
      t = mfma_maxima<NS, CS, W, true>(mc, N, a.n_cells, L, lds, tab);
    } else {
      t = score_tree<NS, CS, W, false, INCR, SPL, HALF>(
          dm, cells, N, a.n_cells, L, lds, unused01, unused11,
          INCR ? tab : TableView{nullptr, 0u, 0u}, SPL ? next_split() : split);
    }
    cur_score = a.score_init ? t + cur_prior : *a.score;
  } else {
    cur_score = *a.score;
  }

  /* the best tree's rows stay in memory: a new maximum overwrites them
   * there at once (new maxima are rare after the first steps, and the rows
   * then cost no registers) */
  float best_score = *a.best_score;
  float best_beta = *a.best_beta;
  uint32_t best_step = *a.best_step;

  /* the Philox blocks of 64 consecutive steps, one step per lane,
   * computed on the vector unit; each step reads its 8 words by readlane.
   * The beta jumps (a normal draw, in double) of the same 64 steps are
   * computed once per block, one per lane, whatever the move type: a beta
   * move reads its own, other moves ignore theirs. */
  uint32_t rbase = a.step_begin;
  bbs_u32x4 ra = bbs_philox_block(a.seed, a.job, a.chain, rbase + lane, BBS_PURPOSE_MOVE_A);
  bbs_u32x4 rb = bbs_philox_block(a.seed, a.job, a.chain, rbase + lane, BBS_PURPOSE_MOVE_B);
  float jumps = bbs_beta_jump(rb, a.probs);
  const Exp2Lanes exp2t = exp2_lanes();
  for (uint32_t step = a.step_begin; step < a.step_end; step++) {
    if (step - rbase == 64u) {
      rbase = step;
      ra = bbs_philox_block(a.seed, a.job, a.chain, rbase + lane, BBS_PURPOSE_MOVE_A);
      rb = bbs_philox_block(a.seed, a.job, a.chain, rbase + lane, BBS_PURPOSE_MOVE_B);
      jumps = bbs_beta_jump_at(rb, a.probs, bbs_coef_base());
    }
    const bbs_raw_move raw = draw_move(ra, rb, jumps, step - rbase, N, a.probs);
    Proposal<NS> P;
    propose<NS, TRACE>(raw, N, cur_beta, am, dm, P);
    const MoveCtx<NS> &m = P.m;
    const float nb = P.nb;
    const uint64_t(&pam)[NS][NS] = P.pam;
    const uint64_t(&pdm)[NS][NS] = P.pdm;

    float prop_beta = cur_beta;
    LogProbs Lp = L;
    float prop_prior = cur_prior;
    if (m.type == BBS_MOVE_CHANGE_BETA) {
      prop_beta = nb;
      const uint32_t kb = bbs_coef_base();
      Lp.L01 = bbs_logf_at(nb, kb);
      Lp.L11 = bbs_logf_of_double_at(1.0 - (double)nb, kb);
      prop_prior = logscore_beta(a.prior, nb, kb);
    }
    uint64_t changed[NS];
    float prop_raw;
    if (INCR && m.type != BBS_MOVE_CHANGE_BETA) {
      bool diff[NS];
#pragma unroll
      for (int k = 0; k < NS; k++) {
        diff[k] = false;
#pragma unroll
        for (int wd = 0; wd < NS; wd++) diff[k] |= pdm[k][wd] != dm[k][wd];
        changed[k] = ballot(diff[k]);
      }
      store_pen<NS>(diff, N, lds);
      store_rows<NS>(pdm, N, lds);
      prop_raw = score_changed<NS, CS, W, SPL, HALF>(changed, cells, N, a.n_cells, Lp, lds, tab,
                                                     SPL ? next_split() : split);
    } else {
      /* beta moves (or full scoring): with INCR the proposal's pair scores
       * go straight into the table, since most beta moves are accepted; a
       * rejection rebuilds the table from the current beta */
      if (use_mfma) {
        store_rows<NS>(pdm, N, lds);
        prop_raw = mfma_maxima<NS, CS, W, true>(mc, N, a.n_cells, Lp, lds, tab);
      } else {
        prop_raw = score_tree<NS, CS, W, false, INCR, SPL, HALF>(
            pdm, cells, N, a.n_cells, Lp, lds, unused01, unused11,
            INCR ? tab : TableView{nullptr, 0u, 0u}, SPL ? next_split() : split);
      }
    }
    const float prop_score = prop_raw + prop_prior;

    const float acc_prob = acceptance_corr(P) * expf_lanes((prop_score - cur_score) * a.gamma, exp2t);
    const bool accept = acc_prob > raw.acceptance_level;
    const bool new_max = (a.is_first_chain && step == 0u) || prop_score > best_score;

    if (TRACE && a.trace != nullptr && lane == 0u && leader)
      a.trace[step] = step_trace(P, prop_score, accept, new_max);
    if (TRACE && a.prop_am != nullptr && leader)
      write_rows<NS>(a.prop_am + 2ull * step * N, a.prop_dm + 2ull * step * N, pam, pdm, N);

    if (new_max) {
      best_score = prop_score;
      best_beta = prop_beta;
      best_step = step;
      if (leader) write_rows<NS>(a.best_am, a.best_dm, pam, pdm, N);
    }
    if (INCR && m.type == BBS_MOVE_CHANGE_BETA && !accept) {
      /* the table holds the rejected beta's scores: back to the current
       * beta (the rows in lds.rows are the current tree's) */
      if (use_mfma) {
        mfma_maxima<NS, CS, W, false>(mc, N, a.n_cells, L, lds, tab);
      } else {
        float best[2];
        bool n0[2];
        node_maxima<NS, CS, W, false, true, 0, HALF>(cells, N, L, lds, tab, best, n0, unused01,
                                                     unused11);
      }
      wave_lds_sync();
    }
    if (accept) {
      if (INCR && m.type != BBS_MOVE_CHANGE_BETA) {
        /* the table follows the accepted tree (its rows are in lds.rows) */
        float R[2];
        bool n0[2];
        changed_nodes<NS, CS, W, true, HALF>(changed, cells, Lp, lds, tab, R, n0);
        wave_lds_sync();
      }
      cur_score = prop_score;
      cur_beta = prop_beta;
      L = Lp;
      cur_prior = prop_prior;
#pragma unroll
      for (int k = 0; k < NS; k++)
#pragma unroll
        for (int wd = 0; wd < NS; wd++) {
          am[k][wd] = pam[k][wd];
          dm[k][wd] = pdm[k][wd];
        }
    }
  }

  if (!leader) return;
  write_rows<NS>(a.am, a.dm, am, dm, N);
  if (lane == 0u) {
    *a.beta = cur_beta;
    *a.score = cur_score;
    *a.best_score = best_score;
    *a.best_beta = best_beta;
    *a.best_step = best_step;
  }
}

/* ------------------------------------------------------------ plain kernel */

/* One dataset per launch, one wavefront per chain (the drop-in form). */
/* Dynamic LDS of an incremental chain wave, in floats: the pair-score
 * table (a row per cell of the dataset and the spare rows), then the tree rows
 * (16 bytes per node), the per-cell sums of the layout's cell slots and the
 * node penalties. Every part starts on a 16-byte boundary and the total is a
 * multiple of 4 floats, so consecutive slabs stay aligned. The scan's unused
 * reads past the end (TableView::scan_reach) stay inside the slab: after the
 * last table row come at least 4 n_nodes + 64 floats, and the penalties get
 * at least scan_reach floats. */
struct ChainLds {
  uint32_t table, rows, sums, pen, total;
  /* HALF: a table of the first 64 cells only (two full cell slots, the
   * second rescored every step) */
  template <bool HALF = false>
  __host__ __device__ static ChainLds of(uint32_t n_nodes, uint32_t n_cells, int cs) {
    const uint32_t stride = TableView::stride_for(n_nodes);
    const uint32_t slots = cs == 1 ? 64u : cs == 3 ? 96u : 128u;
    const uint32_t pen_floats = TableView::pen_len(n_nodes) > TableView::scan_reach(n_nodes)
                                    ? TableView::pen_len(n_nodes)
                                    : TableView::scan_reach(n_nodes);
    ChainLds l;
    l.table = 0u;
    if constexpr (HALF)
      l.rows = 64u * stride;
    else
      l.rows = TableView::floats_for(n_nodes, n_cells, slots);
    l.sums = l.rows + 4u * n_nodes;
    l.pen = l.sums + slots;
    l.total = l.pen + pen_floats;
    return l;
  }
};

template <int NS, int CS, int W, bool TRACE, bool INCR, bool HALF = false>
__global__ __launch_bounds__(64) void chain_kernel(ChainArgs a) {
  extern __shared__ __align__(16) float dyn[]; /* ChainLds with INCR */
  const uint32_t chain = blockIdx.x;
  if (chain >= a.n_chains) return;
  const uint64_t rows = 2ull * chain * a.n_nodes;
  ChainView v;
  v.cells = a.cells;
  v.am = a.am + rows;
  v.dm = a.dm + rows;
  v.best_am = a.best_am + rows;
  v.best_dm = a.best_dm + rows;
  v.beta = a.beta + chain;
  v.score = a.score + chain;
  v.best_beta = a.best_beta + chain;
  v.best_score = a.best_score + chain;
  v.best_step = a.best_step + chain;
  v.trace = (TRACE && chain < a.n_traced) ? a.trace + (uint64_t)chain * a.trace_stride
                                          : nullptr;
  v.prop_am = (TRACE && chain < a.n_prop_traced) ? a.prop_am + rows * a.trace_stride : nullptr;
  v.prop_dm = (TRACE && chain < a.n_prop_traced) ? a.prop_dm + rows * a.trace_stride : nullptr;
  v.seed = a.seed;
  v.job = a.job;
  v.chain = a.chain_offset + chain;
  v.step_begin = a.step_begin;
  v.step_end = a.step_end;
  v.n_cells = a.n_cells;
  v.n_nodes = a.n_nodes;
  v.is_first_chain = a.chain_offset + chain == 0u;
  v.score_init = a.score_init != 0u;
  v.L00 = a.L00;
  v.L10 = a.L10;
  v.gamma = a.gamma;
  v.prior = a.prior;
  v.probs = a.probs;
  if constexpr (INCR && HALF) {
    /* the table of cells 0..63 (slot 0); slot 1 is rescored every step */
    const ChainLds l = ChainLds::of<true>(a.n_nodes, a.n_cells, CS);
    WaveLds lds{reinterpret_cast<uint4 *>(dyn + l.rows), dyn + l.sums, dyn + l.pen};
    run_chain<NS, CS, W, TRACE, true, false, false, true>(
        v, lds, TableView{dyn + l.table, TableView::stride_for(a.n_nodes), 64u});
  } else if constexpr (INCR) {
    const ChainLds l = ChainLds::of(a.n_nodes, a.n_cells, CS);
    WaveLds lds{reinterpret_cast<uint4 *>(dyn + l.rows), dyn + l.sums, dyn + l.pen};
    run_chain<NS, CS, W, TRACE, true, true>(
        v, lds, TableView{dyn + l.table, TableView::stride_for(a.n_nodes), a.n_cells});
  } else {
    __shared__ WaveLdsStore store;
    WaveLds lds = store.view();
    run_chain<NS, CS, W, TRACE, false, false>(v, lds, TableView{nullptr, 0u, 0u});
  }
}

/* ----------------------------------------------------- speculative chain */

/* One speculative proposal's result, exchanged between the waves of a
 * speculative chain. */
struct SpecSlot {
  float score, beta, L01, L11, prior;
  uint32_t accept;
  uint32_t pad[2];
};

/* Dynamic LDS of a speculative chain of k waves, in floats: two pair-score
 * tables (as ChainLds's: the current tree's, and one a beta proposal writes
 * its scores to), the exchange slots (2 rounds x k), the rows of an
 * accepted proposal (n_nodes x (am, dm)), then per wave its proposal's tree
 * rows, per-cell sums and penalties. The scan's unused reads past a table's
 * last row stay inside what follows it; each wave's penalties get
 * scan_reach floats. */
struct SpecLds {
  uint32_t table, table_floats, slots, xrows, wave0, per_wave, sums_off, pen_off, total;
  __host__ __device__ static SpecLds of(uint32_t n_nodes, uint32_t n_cells, int cs, uint32_t k) {
    const uint32_t cslots = cs == 1 ? 64u : cs == 3 ? 96u : 128u;
    const uint32_t pen = TableView::pen_len(n_nodes) > TableView::scan_reach(n_nodes)
                             ? TableView::pen_len(n_nodes)
                             : TableView::scan_reach(n_nodes);
    SpecLds l;
    l.table = 0u;
    l.table_floats = (TableView::floats_for(n_nodes, n_cells, cslots) + 3u) & ~3u;
    l.slots = 2u * l.table_floats;
    l.xrows = l.slots + 2u * k * (uint32_t)(sizeof(SpecSlot) / 4u);
    l.wave0 = l.xrows + 8u * n_nodes; /* 2 rows of 2 u64 per node */
    l.sums_off = 4u * n_nodes;
    l.pen_off = l.sums_off + cslots;
    l.per_wave = (l.pen_off + pen + 3u) & ~3u;
    l.total = l.wave0 + k * l.per_wave;
    return l;
  }
};

/*
 * A chain on k waves that evaluate consecutive steps at once, for launches
 * with SIMDs to spare (few chains). Each round, wave i proposes step s + i
 * against the same current tree, with the shared pair-score table. That is
 * the sequential chain's proposal of step s + i whenever steps s .. s + i - 1
 * are all rejected, and most steps are. After one barrier every wave reads
 * the k accept decisions. Steps up to and including the first accepted one
 * are the sequential chain's; the rest are dropped and proposed again from
 * the new tree. The accepting wave writes its tree to the exchange rows and
 * updates the table (changed nodes, or every pair for a new beta); after a
 * second barrier all waves take the new tree. Trajectories, traces and
 * bests equal the sequential chain's step for step; only wave 0 writes the
 * final state.
 */
template <int NS, int CS, int W, bool TRACE>
__device__ __forceinline__ void run_chain_spec(const ChainView &a, float *dyn, const SpecLds &l,
                                               uint32_t wave, uint32_t K) {
  const uint32_t lane = lane_id();
  const uint32_t N = a.n_nodes;
  float *wbase = dyn + l.wave0 + wave * l.per_wave;
  WaveLds lds{reinterpret_cast<uint4 *>(wbase), wbase + l.sums_off, wbase + l.pen_off};
  /* table(cur) holds the current tree's pair scores (a base computed, not
   * an indexed pair of views, which would live in scratch memory) */
  const uint32_t stride = TableView::stride_for(N);
  auto table = [&](uint32_t which) {
    return TableView{dyn + l.table + (which ? l.table_floats : 0u), stride, a.n_cells};
  };
  uint32_t cur = 0u;
  SpecSlot *slots = reinterpret_cast<SpecSlot *>(dyn + l.slots);
  uint64_t *xam = reinterpret_cast<uint64_t *>(dyn + l.xrows);
  uint64_t *xdm = xam + 2u * N;

  Cells<CS, W> cells;
  load_cells<CS, W>(a.cells, a.n_cells, cells);
  uint64_t am[NS][NS], dm[NS][NS];
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const uint32_t x = lane + 64u * (uint32_t)k;
#pragma unroll
    for (int wd = 0; wd < NS; wd++) {
      am[k][wd] = x < N ? a.am[2u * x + (uint32_t)wd] : 0ull;
      dm[k][wd] = x < N ? a.dm[2u * x + (uint32_t)wd] : 0ull;
    }
  }
  float cur_beta = *a.beta;
  LogProbs L;
  L.L00 = a.L00;
  L.L10 = a.L10;
  L.L01 = bbs_logf(cur_beta);
  L.L11 = bbs_logf_of_double(1.0 - (double)cur_beta);
  float cur_prior = logscore_beta(a.prior, cur_beta);
  uint32_t unused01[2], unused11[2];
  /* wave 0 builds the shared table; the others score without writing it */
  float t;
  if (wave == 0u)
    t = score_tree<NS, CS, W, false, true>(dm, cells, N, a.n_cells, L, lds, unused01, unused11,
                                           table(0u));
  else
    t = score_tree<NS, CS, W, false, false>(dm, cells, N, a.n_cells, L, lds, unused01,
                                            unused11);
  float cur_score = a.score_init ? t + cur_prior : *a.score;
  __syncthreads();

  float best_score = *a.best_score;
  float best_beta = *a.best_beta;
  uint32_t best_step = *a.best_step;
  const Exp2Lanes exp2t = exp2_lanes();
  /* Every wave holds the same two Philox blocks, steps [rbase, rbase + 128),
   * one step per lane in each, and which of those steps change beta (bit i
   * of bm0 / bm1: step rbase + i / rbase + 64 + i). A round starts at
   * s < rbase + 64, so its k steps are all there. */
  uint32_t rbase = a.step_begin;
  bbs_u32x4 ra0 = bbs_philox_block(a.seed, a.job, a.chain, rbase + lane, BBS_PURPOSE_MOVE_A);
  bbs_u32x4 rb0 = bbs_philox_block(a.seed, a.job, a.chain, rbase + lane, BBS_PURPOSE_MOVE_B);
  bbs_u32x4 ra1 = bbs_philox_block(a.seed, a.job, a.chain, rbase + 64u + lane, BBS_PURPOSE_MOVE_A);
  bbs_u32x4 rb1 = bbs_philox_block(a.seed, a.job, a.chain, rbase + 64u + lane, BBS_PURPOSE_MOVE_B);
  float jumps0 = bbs_beta_jump(rb0, a.probs), jumps1 = bbs_beta_jump(rb1, a.probs);
  uint64_t bm0 = ballot(bbs_move_type_of(bbs_u01f(ra0.x[0]), a.probs) == BBS_MOVE_CHANGE_BETA);
  uint64_t bm1 = ballot(bbs_move_type_of(bbs_u01f(ra1.x[0]), a.probs) == BBS_MOVE_CHANGE_BETA);
  uint32_t round = 0u;
  for (uint32_t s = a.step_begin; s < a.step_end; round++) {
    const uint32_t my = s + wave; /* past step_end: proposed, never used */
    if (s - rbase >= 64u) {
      rbase += 64u;
      ra0 = ra1;
      rb0 = rb1;
      jumps0 = jumps1;
      bm0 = bm1;
      ra1 = bbs_philox_block(a.seed, a.job, a.chain, rbase + 64u + lane, BBS_PURPOSE_MOVE_A);
      rb1 = bbs_philox_block(a.seed, a.job, a.chain, rbase + 64u + lane, BBS_PURPOSE_MOVE_B);
      jumps1 = bbs_beta_jump_at(rb1, a.probs, bbs_coef_base());
      bm1 = ballot(bbs_move_type_of(bbs_u01f(ra1.x[0]), a.probs) == BBS_MOVE_CHANGE_BETA);
    }
    /* the round's first beta move (bit i: step s + i) */
    const uint32_t off = s - rbase;
    const uint64_t betas = ((bm0 >> off) | (off ? bm1 << (64u - off) : 0ull)) &
                           (K < 64u ? (1ull << K) - 1ull : ~0ull);
    const uint32_t first_beta = betas ? (uint32_t)__ffsll((long long)betas) - 1u : K;
    const uint32_t sl = my - rbase;
    const bbs_raw_move raw = sl < 64u ? draw_move(ra0, rb0, jumps0, sl, N, a.probs)
                                      : draw_move(ra1, rb1, jumps1, sl - 64u, N, a.probs);
    Proposal<NS> P;
    propose<NS, TRACE>(raw, N, cur_beta, am, dm, P);
    const bool beta_move = P.m.type == BBS_MOVE_CHANGE_BETA;
    const TableView tab = table(cur);
    float prop_beta = cur_beta;
    LogProbs Lp = L;
    float prop_prior = cur_prior;
    if (beta_move) {
      prop_beta = P.nb;
      const uint32_t kb = bbs_coef_base();
      Lp.L01 = bbs_logf_at(P.nb, kb);
      Lp.L11 = bbs_logf_of_double_at(1.0 - (double)P.nb, kb);
      prop_prior = logscore_beta(a.prior, P.nb, kb);
    }
    /* scored against the current tree's table, which no wave writes now */
    uint64_t changed[NS];
    float prop_raw;
    if (!beta_move) {
      bool diff[NS];
#pragma unroll
      for (int k = 0; k < NS; k++) {
        diff[k] = false;
#pragma unroll
        for (int wd = 0; wd < NS; wd++) diff[k] |= P.pdm[k][wd] != dm[k][wd];
        changed[k] = ballot(diff[k]);
      }
      store_pen<NS>(diff, N, lds);
      store_rows<NS>(P.pdm, N, lds);
      prop_raw = score_changed<NS, CS, W>(changed, cells, N, a.n_cells, Lp, lds, tab);
    } else if (wave == first_beta) {
      /* the round's first beta move writes its pair scores to the spare
       * table, which becomes the current one if this step is accepted */
      prop_raw = score_tree<NS, CS, W, false, true>(P.pdm, cells, N, a.n_cells, Lp, lds,
                                                    unused01, unused11, table(cur ^ 1u));
    } else {
      prop_raw = score_tree<NS, CS, W, false, false>(P.pdm, cells, N, a.n_cells, Lp, lds,
                                                     unused01, unused11);
    }
    const float prop_score = prop_raw + prop_prior;
    const float acc_prob =
        acceptance_corr(P) * expf_lanes((prop_score - cur_score) * a.gamma, exp2t);
    const bool accept = acc_prob > raw.acceptance_level;

    SpecSlot *sr = slots + (round & 1u) * K; /* a round's slots are rewritten two rounds on */
    if (lane == 0u) {
      SpecSlot o;
      o.score = prop_score;
      o.beta = prop_beta;
      o.L01 = Lp.L01;
      o.L11 = Lp.L11;
      o.prior = prop_prior;
      o.accept = accept ? 1u : 0u;
      o.pad[0] = o.pad[1] = 0u;
      sr[wave] = o;
    }
    __syncthreads();
    const uint32_t n_act = a.step_end - s < K ? a.step_end - s : K;
    uint32_t j = n_act; /* the first accepted step of the round, if any */
    for (uint32_t i = 0; i < n_act; i++)
      if (sr[i].accept != 0u) {
        j = i;
        break;
      }
    const uint32_t committed = j < n_act ? j + 1u : n_act;
    /* best tracking over the committed proposals in step order
     * (src/Application.hpp:456) */
    uint32_t last_max = K;
    bool my_max = false;
    for (uint32_t i = 0; i < committed; i++) {
      const float ps = sr[i].score;
      if ((a.is_first_chain && s + i == 0u) || ps > best_score) {
        best_score = ps;
        best_beta = sr[i].beta;
        best_step = s + i;
        last_max = i;
        if (i == wave) my_max = true;
      }
    }
    if (TRACE && a.trace != nullptr && lane == 0u && wave < committed)
      a.trace[my] = step_trace(P, prop_score, wave == j, my_max);
    if (TRACE && a.prop_am != nullptr && wave < committed)
      write_rows<NS>(a.prop_am + 2ull * my * N, a.prop_dm + 2ull * my * N, P.pam, P.pdm, N);
    if (wave == last_max) write_rows<NS>(a.best_am, a.best_dm, P.pam, P.pdm, N);
    if (j < n_act) {
      /* the accepting wave publishes its tree and brings the table to it:
       * the changed nodes; for a new beta, the spare table already holds it
       * when this is the round's first beta move, else every pair is
       * rescored into the current table (its rows are the tree's, in its
       * LDS) */
      if (wave == j) {
        write_rows<NS>(xam, xdm, P.pam, P.pdm, N);
        float R[2];
        bool n0[2];
        if (!beta_move)
          changed_nodes<NS, CS, W, true>(changed, cells, Lp, lds, tab, R, n0);
        else if (j != first_beta)
          node_maxima<NS, CS, W, false, true>(cells, N, Lp, lds, tab, R, n0, unused01,
                                               unused11);
      }
      if (j == first_beta) cur ^= 1u;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < NS; k++) {
        const uint32_t x = lane + 64u * (uint32_t)k;
#pragma unroll
        for (int wd = 0; wd < NS; wd++) {
          am[k][wd] = x < N ? xam[2u * x + (uint32_t)wd] : 0ull;
          dm[k][wd] = x < N ? xdm[2u * x + (uint32_t)wd] : 0ull;
        }
      }
      const SpecSlot &o = sr[j];
      cur_score = o.score;
      cur_beta = o.beta;
      L.L01 = o.L01;
      L.L11 = o.L11;
      cur_prior = o.prior;
    }
    s += committed;
  }
  if (wave != 0u) return;
  write_rows<NS>(a.am, a.dm, am, dm, N);
  if (lane == 0u) {
    *a.beta = cur_beta;
    *a.score = cur_score;
    *a.best_score = best_score;
    *a.best_beta = best_beta;
    *a.best_step = best_step;
  }
}

/* One chain per workgroup of k = blockDim / 64 waves running it
 * speculatively (run_chain_spec). */
template <int NS, int CS, int W, bool TRACE>
__global__ __launch_bounds__(512) void chain_spec_kernel(ChainArgs a) {
  extern __shared__ __align__(16) float dyn[];
  const uint32_t chain = blockIdx.x;
  if (chain >= a.n_chains) return;
  const uint32_t K = blockDim.x >> 6, wave = threadIdx.x >> 6;
  const uint64_t rows = 2ull * chain * a.n_nodes;
  ChainView v;
  v.cells = a.cells;
  v.am = a.am + rows;
  v.dm = a.dm + rows;
  v.best_am = a.best_am + rows;
  v.best_dm = a.best_dm + rows;
  v.beta = a.beta + chain;
  v.score = a.score + chain;
  v.best_beta = a.best_beta + chain;
  v.best_score = a.best_score + chain;
  v.best_step = a.best_step + chain;
  v.trace = (TRACE && chain < a.n_traced) ? a.trace + (uint64_t)chain * a.trace_stride
                                          : nullptr;
  v.prop_am = (TRACE && chain < a.n_prop_traced) ? a.prop_am + rows * a.trace_stride : nullptr;
  v.prop_dm = (TRACE && chain < a.n_prop_traced) ? a.prop_dm + rows * a.trace_stride : nullptr;
  v.seed = a.seed;
  v.job = a.job;
  v.chain = a.chain_offset + chain;
  v.step_begin = a.step_begin;
  v.step_end = a.step_end;
  v.n_cells = a.n_cells;
  v.n_nodes = a.n_nodes;
  v.is_first_chain = a.chain_offset + chain == 0u;
  v.score_init = a.score_init != 0u;
  v.L00 = a.L00;
  v.L10 = a.L10;
  v.gamma = a.gamma;
  v.prior = a.prior;
  v.probs = a.probs;
  run_chain_spec<NS, CS, W, TRACE>(v, dyn, SpecLds::of(a.n_nodes, a.n_cells, CS, K), wave, K);
}

/* One chain per workgroup of SPLIT waves, split by cells: wave i scores
 * cells [i * cs, (i + 1) * cs), cs = ceil(n_cells / SPLIT), with its own
 * table; every wave computes the same move, acceptance and ordered cell sum.
 * For launches with fewer chains than SIMDs, where a step's latency, not
 * the device's width, sets the rate. Dynamic LDS: the two shared sum buffers
 * (2 x 128 floats), then one ChainLds slab of cs cells per wave. */
template <int NS, int W, bool TRACE, int SPLIT>
__global__ __launch_bounds__(64 * SPLIT) void chain_split_kernel(ChainArgs a) {
  extern __shared__ __align__(16) float dyn[];
  const uint32_t chain = blockIdx.x;
  if (chain >= a.n_chains) return;
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t cs = (a.n_cells + (uint32_t)SPLIT - 1u) / (uint32_t)SPLIT;
  const uint32_t c0 = wave * cs;
  const uint32_t mine = a.n_cells - c0 < cs ? a.n_cells - c0 : cs; /* >= 1 (host check) */
  /* the sums past n_cells stay +0.0 */
  for (uint32_t i = threadIdx.x; i < 256u; i += 64u * (uint32_t)SPLIT) dyn[i] = 0.0f;
  __syncthreads();
  const uint64_t rows = 2ull * chain * a.n_nodes;
  ChainView v;
  v.cells = a.cells + 4ull * c0;
  v.am = a.am + rows;
  v.dm = a.dm + rows;
  v.best_am = a.best_am + rows;
  v.best_dm = a.best_dm + rows;
  v.beta = a.beta + chain;
  v.score = a.score + chain;
  v.best_beta = a.best_beta + chain;
  v.best_score = a.best_score + chain;
  v.best_step = a.best_step + chain;
  v.trace = (TRACE && chain < a.n_traced) ? a.trace + (uint64_t)chain * a.trace_stride
                                          : nullptr;
  v.prop_am = (TRACE && chain < a.n_prop_traced) ? a.prop_am + rows * a.trace_stride : nullptr;
  v.prop_dm = (TRACE && chain < a.n_prop_traced) ? a.prop_dm + rows * a.trace_stride : nullptr;
  v.seed = a.seed;
  v.job = a.job;
  v.chain = a.chain_offset + chain;
  v.step_begin = a.step_begin;
  v.step_end = a.step_end;
  v.n_cells = mine;
  v.n_nodes = a.n_nodes;
  v.is_first_chain = a.chain_offset + chain == 0u;
  v.score_init = a.score_init != 0u;
  v.L00 = a.L00;
  v.L10 = a.L10;
  v.gamma = a.gamma;
  v.prior = a.prior;
  v.probs = a.probs;
  const ChainLds l = ChainLds::of(a.n_nodes, cs, 1);
  float *slab = dyn + 256u + wave * l.total;
  WaveLds lds{reinterpret_cast<uint4 *>(slab + l.rows), slab + l.sums, slab + l.pen};
  run_chain<NS, 1, W, TRACE, true, false, true>(
      v, lds, TableView{slab + l.table, TableView::stride_for(a.n_nodes), mine},
      CellSplit{dyn, c0, a.n_cells}, wave == 0u);
}

/* ---------------------------------------------------------- batched kernel */

/*
 * Batched form with one wavefront per (job, chain) item: the items of one
 * register layout (shape) are launched together, largest work first, so
 * the hardware dispatcher hands the next item to whichever SIMD frees up
 * first. waves_per_group wavefronts share a workgroup and run independent
 * items (run_chain has no workgroup barrier). WPS is the register budget:
 * 1 lets the compiler use every register for one wave per SIMD (the most
 * instruction-level parallelism); 2 holds a wave to 256 registers so that
 * two share a SIMD (latency hidden by the other wave, at the price of a
 * few spills in the larger layouts). Results are identical.
 */
template <int NS, int CS, int W, bool INCR, int WPS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPS))) void
batch_items_kernel(BatchArgs a, uint32_t first, uint32_t count, uint32_t wave_floats) {
  extern __shared__ __align__(16) float dyn[]; /* INCR: one ChainLds slab per wave */
  const uint32_t wave = threadIdx.x >> 6;
  const uint32_t k = blockIdx.x * (blockDim.x >> 6) + wave;
  if (k >= count) return;
  const uint32_t packed = a.items[first + k];
  const uint32_t j = packed >> 16, c = packed & 0xffffu;
  const BatchJob &J = a.jobs[j];
  const uint64_t rows = 2ull * ((uint64_t)J.row_offset + (uint64_t)c * J.n_nodes);
  const uint32_t gc = J.chain_offset + c;
  ChainView v;
  v.cells = a.cells + 4ull * J.cell_offset;
  v.am = a.am + rows;
  v.dm = a.dm + rows;
  v.best_am = a.best_am + rows;
  v.best_dm = a.best_dm + rows;
  v.beta = a.beta + gc;
  v.score = a.score + gc;
  v.best_beta = a.best_beta + gc;
  v.best_score = a.best_score + gc;
  v.best_step = a.best_step + gc;
  v.trace = nullptr;
  v.prop_am = v.prop_dm = nullptr;
  v.seed = J.seed;
  v.job = J.job_tag;
  v.chain = c;
  v.step_begin = J.step_begin;
  v.step_end = J.step_end;
  v.n_cells = J.n_cells;
  v.n_nodes = J.n_nodes;
  v.is_first_chain = c == 0u;
  v.score_init = J.score_init != 0u;
  v.L00 = J.L00;
  v.L10 = J.L10;
  v.gamma = J.gamma;
  v.prior = J.prior;
  v.probs = J.probs;
  if constexpr (INCR) {
    float *slab = dyn + wave * wave_floats;
    const ChainLds l = ChainLds::of(J.n_nodes, J.n_cells, CS);
    WaveLds lds{reinterpret_cast<uint4 *>(slab + l.rows), slab + l.sums, slab + l.pen};
    /* layout 233 (65..96 genes, 65..96 cells): its LDS slabs hold it to
     * about one wave per SIMD, so the matrix operands' registers cost it
     * little concurrency, and its beta moves count pairs on the matrix
     * cores where the dataset fills its tiles (measured faster) */
    run_chain<NS, CS, W, false, true, NS == 2 && CS == 3 && W == 3 && WPS == 1>(
        v, lds, TableView{slab + l.table, TableView::stride_for(J.n_nodes), J.n_cells});
  } else {
    __shared__ WaveLdsStore lds_all[4];
    WaveLds lds = lds_all[wave].view();
    run_chain<NS, CS, W, false, false, false>(v, lds, TableView{nullptr, 0u, 0u});
  }
}

/* ------------------------------------------------------------ score kernel */

template <int NS, int CS, int W, bool COUNTS>
__global__ __launch_bounds__(64) void score_kernel(ScoreArgs a) {
  __shared__ WaveLdsStore store;
  WaveLds lds = store.view();
  const uint32_t lane = lane_id();
  const uint32_t N = a.n_nodes;
  Cells<CS, W> cells;
  load_cells<CS, W>(a.cells, a.n_cells, cells);
  for (uint32_t tree = blockIdx.x; tree < a.n_trees; tree += gridDim.x) {
    uint64_t dm[NS][NS];
    const uint64_t base = (uint64_t)tree * N;
#pragma unroll
    for (int k = 0; k < NS; k++) {
      const uint32_t x = lane + 64u * (uint32_t)k;
#pragma unroll
      for (int wd = 0; wd < NS; wd++)
        dm[k][wd] = x < N ? a.dm[2u * (base + x) + (uint32_t)wd] : 0ull;
    }
    LogProbs L;
    L.L00 = a.L00;
    L.L10 = a.L10;
    float prior;
    if (a.betas) {
      const float beta = a.betas[tree];
      L.L01 = bbs_logf(beta);
      L.L11 = bbs_logf_of_double(1.0 - (double)beta);
      prior = logscore_beta(a.prior, beta);
    } else {
      L.L01 = a.L01;
      L.L11 = a.L11;
      prior = a.prior_value;
    }
    uint32_t arg01[2], arg11[2];
    float T = score_tree<NS, CS, W, COUNTS>(dm, cells, N, a.n_cells, L, lds,
                                            arg01, arg11);
    if (lane == 0u) a.ll[tree] = a.with_prior ? T + prior : T;
    if (COUNTS) {
#pragma unroll
      for (int j = 0; j < Slots<CS>::n; j++) {
        const uint32_t cell = lane + 64u * (uint32_t)j;
        if (cell < a.n_cells) {
          const uint32_t n01 = arg01[j], n11 = arg11[j];
          const uint32_t n00 = (uint32_t)cells.fa[j] - n01;
          const uint32_t n10 = (uint32_t)cells.fb[j] - n11;
          const uint32_t packed = n00 | (n01 << 8) | (n10 << 16) | (n11 << 24);
          reinterpret_cast<uint32_t *>(a.counts)[(uint64_t)tree * a.n_cells + cell] =
              packed;
        }
      }
    }
  }
}

/* ------------------------------------------- bulk score, lane = tree */

/*
 * Bulk TreeScorer over many trees with one tree per lane: the 64 lanes of a
 * wave score 64 different trees against the same data, so every lane does
 * useful work whatever the cell count. The cell data are wave-uniform
 * (scalar loads, SGPR operands); each lane keeps its per-cell running
 * maxima in VGPRs (CMAX of them) and streams its tree's ancestor rows, one
 * node per iteration. Cells beyond n_cells carry no known entry: their
 * maximum is -0.0 and adding it leaves the ordered cell sum unchanged.
 * cells32: [CMAX][10] words: K&M words 0..3, K&~M words 0..3, |K&~M|,
 * |K&M| (as floats).
 */
template <int CMAX, int W>
__global__ __launch_bounds__(256) void score_lanes_kernel(ScoreArgs a,
                                                          const uint32_t *__restrict__ cells32) {
  /* per cell: (K&M w0..w2, |K&~M|) (K&~M w0..w2, |K&M|) for W <= 3, and a
   * third quad (K&M w3, K&~M w3, 0, 0) for W = 4 */
  constexpr int Q = W <= 3 ? 2 : 3;
  __shared__ uint4 cl[CMAX * Q];
  for (uint32_t c = threadIdx.x; c < (uint32_t)CMAX; c += blockDim.x) {
    const uint32_t *g = cells32 + 10u * c;
    cl[Q * c] = make_uint4(g[0], g[1], g[2], g[8]);
    cl[Q * c + 1] = make_uint4(g[4], g[5], g[6], g[9]);
    if (Q == 3) cl[Q * c + 2] = make_uint4(g[3], g[7], 0u, 0u);
  }
  __syncthreads();
  const uint32_t N = a.n_nodes;
  const uint32_t lanes_total = gridDim.x * blockDim.x;
  for (uint32_t tree0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
       tree0 < a.n_trees; tree0 += lanes_total) {
    const uint32_t tree = tree0 + lane_id();
    const bool live = tree < a.n_trees;
    const uint32_t t = live ? tree : a.n_trees - 1u;
    LogProbs L;
    L.L00 = a.L00;
    L.L10 = a.L10;
    float prior;
    if (a.betas) {
      const float beta = a.betas[t];
      L.L01 = bbs_logf(beta);
      L.L11 = bbs_logf_of_double(1.0 - (double)beta);
      prior = logscore_beta(a.prior, beta);
    } else {
      L.L01 = a.L01;
      L.L11 = a.L11;
      prior = a.prior_value;
    }
    const uint4 *rows = reinterpret_cast<const uint4 *>(a.dm) + (uint64_t)t * N;
    float best[CMAX];
    /* one (cell, node) pair: the reference's four masked popcounts */
    auto pair = [&](const uint4 *cp, const uint4 &A) {
      const uint4 q0 = cp[0], q1 = cp[1];
      uint32_t n11 = bcnt_acc(q0.x & A.x, 0u);
      uint32_t n01 = bcnt_acc(q1.x & A.x, 0u);
      if (W > 1) {
        n11 = bcnt_acc(q0.y & A.y, n11);
        n01 = bcnt_acc(q1.y & A.y, n01);
      }
      if (W > 2) {
        n11 = bcnt_acc(q0.z & A.z, n11);
        n01 = bcnt_acc(q1.z & A.z, n01);
      }
      if (W > 3) {
        const uint4 q2 = cp[2];
        n11 = bcnt_acc(q2.x & A.w, n11);
        n01 = bcnt_acc(q2.y & A.w, n01);
      }
      return pair_score(n01, n11, __uint_as_float(q0.w), __uint_as_float(q1.w), L);
    };
    const uint4 A = rows[0];
#pragma unroll
    for (int c = 0; c < CMAX; c++) best[c] = pair(cl + Q * c, A);
    /* four nodes per pass share each cell's LDS reads; past the last node
     * the clamped index repeats node N - 1, which cannot change a maximum
     * (only node 0's NaN is special, handled below) */
    const uint32_t last = N - 1u;
    for (uint32_t v = 1; v < N; v += 4) {
      const uint4 A0 = rows[v];
      const uint4 A1 = rows[v + 1u < last ? v + 1u : last];
      const uint4 A2 = rows[v + 2u < last ? v + 2u : last];
      const uint4 A3 = rows[v + 3u < last ? v + 3u : last];
      /* an opaque base keeps the cell reads inside the node loop: hoisting
       * all of them would need CMAX * 4 Q registers */
      uint32_t base = 0u;
      asm volatile("" : "+v"(base));
      const uint4 *cb = reinterpret_cast<const uint4 *>(
          reinterpret_cast<const char *>(cl) + base);
#pragma unroll
      for (int c = 0; c < CMAX; c++) {
        const uint4 *cp = cb + Q * c;
        const float m01 = vmax(pair(cp, A0), pair(cp, A1));
        const float m23 = vmax(pair(cp, A2), pair(cp, A3));
        best[c] = vmax(best[c], vmax(m01, m23));
      }
    }
    /* A NaN at node 0 sticks in the reference's maximum; it needs an
     * infinite log-probability, i.e. beta (or alpha) exactly 0 or 1. */
    const bool inf_l = !(fabsf(L.L01) <= 3.4e38f) || !(fabsf(L.L11) <= 3.4e38f) ||
                       !(fabsf(L.L00) <= 3.4e38f) || !(fabsf(L.L10) <= 3.4e38f);
    if (inf_l) {
      const uint4 A0 = rows[0];
#pragma unroll
      for (int c = 0; c < CMAX; c++) {
        const float s0 = pair(cl + Q * c, A0);
        if (s0 != s0) best[c] = s0;
      }
    }
    float T = 0.0f;
#pragma unroll
    for (int c = 0; c < CMAX; c++)
      if ((uint32_t)c < a.n_cells) T = T + best[c];
    if (live) a.ll[tree] = a.with_prior ? T + prior : T;
  }
}

/* ------------------------------------------ bulk score, matrix cores */

/*
 * TreeScorer::logscore_tree over many trees, one tree per wavefront, with
 * the pair counts on the matrix cores: n11 = (K&M) A^T and n01 = (K&~M) A^T
 * are products of 0/1 matrices (genes as the inner dimension), computed by
 * v_mfma_i32_32x32x32_i8 on 32 x 32 tiles (rows = nodes, columns = cells).
 * The float part of each pair stays on the vector unit in the reference's
 * order (pair_score). Operands, per lane l (h = l / 32):
 *   A (tree, per node tile): node 32 mt + l % 32, genes 32 kb + 16 h + b
 *   B (data, per cell tile): cell 32 jt + l % 32, the same genes
 *   D: cell 32 jt + l % 32, node 32 mt + 8 (r / 4) + 4 h + r % 4 in reg r.
 * A and B put gene g at the same (half, byte), which is all a product needs.
 * Each lane thus holds one cell's pairs for 16 nodes per tile: the maximum
 * over nodes stays in the lane, the other 16 nodes are in lane l ^ 32.
 * Padding nodes (>= n_nodes) are left out of the maximum; padding cells have
 * no known entries and add +0.0. A NaN at node 0 sticks, as in the
 * reference's first-wins maximum; later NaNs are ignored.
 */
template <int KB, int JT>
__global__ __launch_bounds__(256) void score_mfma_kernel(ScoreArgs a,
                                                         const uint32_t *__restrict__ cells32) {
  __shared__ __align__(16) float sums[4][128];
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6, h = lane >> 5;
  const uint32_t N = a.n_nodes;
  float *const S = sums[wave];
  /* the data operands and per-cell popcounts, once per wave */
  bbs_v4i bkm[JT][KB], bknm[JT][KB];
  float fa[JT], fb[JT];
#pragma unroll
  for (int jt = 0; jt < JT; jt++) {
    const uint32_t *g = cells32 + 10u * (32u * (uint32_t)jt + (lane & 31u));
#pragma unroll
    for (int kb = 0; kb < KB; kb++) {
      bkm[jt][kb] = bits_to_bytes16((g[kb] >> (16u * h)) & 0xFFFFu);
      bknm[jt][kb] = bits_to_bytes16((g[4 + kb] >> (16u * h)) & 0xFFFFu);
    }
    fa[jt] = __uint_as_float(g[8]);
    fb[jt] = __uint_as_float(g[9]);
  }
  const uint32_t waves_total = gridDim.x * 4u;
  for (uint32_t tree = blockIdx.x * 4u + wave; tree < a.n_trees; tree += waves_total) {
    LogProbs L;
    L.L00 = a.L00;
    L.L10 = a.L10;
    float prior;
    if (a.betas) {
      const float beta = a.betas[tree];
      L.L01 = bbs_logf(beta);
      L.L11 = bbs_logf_of_double(1.0 - (double)beta);
      prior = logscore_beta(a.prior, beta);
    } else {
      L.L01 = a.L01;
      L.L11 = a.L11;
      prior = a.prior_value;
    }
    /* the tree operand: node rows as bytes, node tiles x gene blocks */
    bbs_v4i at[KB][KB];
#pragma unroll
    for (int mt = 0; mt < KB; mt++) {
      uint32_t n = 32u * (uint32_t)mt + (lane & 31u);
      n = n < N ? n : N - 1u;
      const uint4 row = reinterpret_cast<const uint4 *>(a.dm)[(uint64_t)tree * N + n];
      const uint32_t w[4] = {row.x, row.y, row.z, row.w};
#pragma unroll
      for (int kb = 0; kb < KB; kb++) at[mt][kb] = bits_to_bytes16((w[kb] >> (16u * h)) & 0xFFFFu);
    }
    const float ninf = __int_as_float(0xff800000);
#pragma unroll
    for (int jt = 0; jt < JT; jt++) {
      float best = ninf;
      bool nan0 = false;
#pragma unroll
      for (int mt = 0; mt < KB; mt++) {
        bbs_v16i c11 = {}, c01 = {};
#pragma unroll
        for (int kb = 0; kb < KB; kb++) {
          c11 = __builtin_amdgcn_mfma_i32_32x32x32_i8(at[mt][kb], bkm[jt][kb], c11, 0, 0, 0);
          c01 = __builtin_amdgcn_mfma_i32_32x32x32_i8(at[mt][kb], bknm[jt][kb], c01, 0, 0, 0);
        }
        /* padding nodes exist only in a tile reaching past n_nodes */
        auto tile_pairs = [&](bool full) {
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const uint32_t node = 32u * (uint32_t)mt + 8u * (uint32_t)(r / 4) + 4u * h +
                                  (uint32_t)(r % 4);
            const float sc = pair_score((uint32_t)c01[r], (uint32_t)c11[r], fa[jt], fb[jt], L);
            if (mt == 0 && r == 0) nan0 = h == 0u && sc != sc;
            best = vmax(best, (full || node < N) ? sc : ninf);
          }
        };
        if (32u * (uint32_t)(mt + 1) <= N)
          tile_pairs(true);
        else
          tile_pairs(false);
      }
      /* the cell's other 16 nodes per tile are in lane l ^ 32 */
      const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(best),
                                                       __float_as_uint(best), false, false);
      const float other = __uint_as_float(sw[1]);
      const float qnan = __int_as_float(0x7fc00000);
      const uint32_t cell = 32u * (uint32_t)jt + lane;
      if (lane < 32u) S[cell] = cell < a.n_cells ? (nan0 ? qnan : vmax(best, other)) : 0.0f;
    }
    wave_lds_sync();
    /* ((0 + x[0]) + x[1]) + ... in cell order */
    const float T = sum_padded<JT * 8>(S);
    wave_lds_sync();
    if (lane == 0u) a.ll[tree] = a.with_prior ? T + prior : T;
  }
}

/* ------------------------------------------------------- random trees */

/* One uniform random tree from a Pruefer code drawn from the Philox stream
 * (chain word = chain_offset + tree), built by one wavefront: rows
 * [tree][n_nodes][2] of am (nullable) and dm, and beta (nullable). */
__device__ void random_tree_wave(const TreeGenArgs &a, uint32_t tree) {
  __shared__ uint32_t degree[128];
  __shared__ uint32_t parent[128];
  __shared__ uint32_t code[128];
  __shared__ uint64_t rows[128][2];
  const uint32_t lane = lane_id();
  const uint32_t N = a.n_nodes;
  const uint32_t chain = a.chain_offset + tree;
  /* code[i] ~ U{0..N-1}, i < N-2 (src/MutationTree.hpp:500-507) */
  for (uint32_t i = lane; i + 2u < N; i += 64u) {
    bbs_u32x4 b = bbs_philox_block(a.seed, a.job, chain, i >> 2, a.purpose);
    code[i] = bbs_uint_below(b.x[i & 3u], N);
  }
  for (uint32_t i = lane; i < 128u; i += 64u) {
    degree[i] = 1u;
    parent[i] = N - 1u;
  }
  __syncthreads();
  if (lane == 0u)
    for (uint32_t i = 0; i + 2u < N; i++) degree[code[i]]++;
  __syncthreads();
  /* decode: repeatedly attach the smallest leaf (src/MutationTree.hpp:454-463) */
  for (uint32_t i = 0; i + 2u < N; i++) {
    bool leaf0 = lane < N && degree[lane] == 1u;
    bool leaf1 = lane + 64u < N && degree[lane + 64u] == 1u;
    uint64_t m0 = ballot(leaf0), m1 = ballot(leaf1);
    uint32_t j = m0 ? (uint32_t)__ffsll((long long)m0) - 1u
                    : 64u + (uint32_t)__ffsll((long long)m1) - 1u;
    __syncthreads();
    if (lane == 0u) {
      uint32_t c = code[i];
      parent[j] = c;
      degree[c]--;
      degree[j]--;
    }
    __syncthreads();
  }
  /* the remaining leaf other than the root hangs below the root: parent
   * already initialised to N-1 (src/MutationTree.hpp:467-473) */
  /* ancestor rows: dm[x] = {x} | dm[parent[x]], iterated to a fixed point */
  uint64_t d[2][2];
  for (int k = 0; k < 2; k++) {
    const uint32_t x = lane + 64u * (uint32_t)k;
    d[k][0] = (x < N && x < 64u) ? (1ull << x) : 0ull;
    d[k][1] = (x < N && x >= 64u) ? (1ull << (x - 64u)) : 0ull;
    if (x < N) {
      const uint32_t r = N - 1u;
      d[k][r >> 6] |= 1ull << (r & 63u);
    }
  }
  for (uint32_t iter = 0; iter < N; iter++) {
    for (int k = 0; k < 2; k++) {
      const uint32_t x = lane + 64u * (uint32_t)k;
      rows[x][0] = d[k][0];
      rows[x][1] = d[k][1];
    }
    __syncthreads();
    bool changed = false;
    for (int k = 0; k < 2; k++) {
      const uint32_t x = lane + 64u * (uint32_t)k;
      if (x < N && x != N - 1u) {
        const uint32_t p = parent[x];
        const uint64_t n0 = d[k][0] | rows[p][0], n1 = d[k][1] | rows[p][1];
        changed = changed || n0 != d[k][0] || n1 != d[k][1];
        d[k][0] = n0;
        d[k][1] = n1;
      }
    }
    __syncthreads();
    if (ballot(changed) == 0ull) break;
  }
  const uint64_t base = (uint64_t)tree * N;
  for (int k = 0; k < 2; k++) {
    const uint32_t x = lane + 64u * (uint32_t)k;
    if (x < N) {
      a.dm[2u * (base + x)] = d[k][0];
      a.dm[2u * (base + x) + 1u] = d[k][1];
    }
  }
  if (a.am) {
    /* am[x] bit y = dm[y] bit x: transpose through LDS */
    for (int k = 0; k < 2; k++) {
      const uint32_t x = lane + 64u * (uint32_t)k;
      rows[x][0] = d[k][0];
      rows[x][1] = d[k][1];
    }
    __syncthreads();
    for (int k = 0; k < 2; k++) {
      const uint32_t x = lane + 64u * (uint32_t)k;
      uint64_t r0 = 0ull, r1 = 0ull;
      if (x < N) {
        for (uint32_t y = 0; y < N; y++) {
          const bool bit = (rows[y][x >> 6] >> (x & 63u)) & 1u;
          if (bit) {
            if (y < 64u)
              r0 |= 1ull << y;
            else
              r1 |= 1ull << (y - 64u);
          }
        }
        a.am[2u * (base + x)] = r0;
        a.am[2u * (base + x) + 1u] = r1;
      }
    }
  }
  if (a.beta && lane == 0u) a.beta[tree] = a.beta_value;
  __syncthreads();
}

__global__ __launch_bounds__(64) void random_trees_kernel(TreeGenArgs a) {
  for (uint32_t tree = blockIdx.x; tree < a.n_trees; tree += gridDim.x)
    random_tree_wave(a, tree);
}

/* ------------------------------------------------ allocation-free staging */

/* The staging of bbs_score_trees_noalloc: one fixed array per GPU, part of
 * this code object, so a call needs no device allocation. In host code its
 * name is only a stand-in: kernels and copies use the address that
 * hipGetSymbolAddress returns for the current device. */
static __device__ __align__(256) unsigned char scratch[BBS_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
} // namespace

hipError_t noalloc_scratch(int device, void **ptr) {
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *ptr = g_scratch_ptr[device];
  return hipSuccess;
}

/* ------------------------------------------------------ VALU peak probe */

/* Integer VALU throughput calibration: 8 independent and/popcount chains per
 * lane (2 int32 operations per step), long enough to hide all latency. */
__global__ __launch_bounds__(256) void valu_peak_kernel(const uint32_t *in, uint32_t *out,
                                                        uint32_t iters) {
  const uint32_t tid = blockIdx.x * 256u + threadIdx.x;
  uint32_t x[8];
#pragma unroll
  for (int k = 0; k < 8; k++) x[k] = in[(threadIdx.x * 8u + (uint32_t)k) & 1023u];
  const uint32_t m = in[1024u + (threadIdx.x & 7u)];
  for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      uint32_t t;
      asm volatile("v_and_b32 %0, %1, %2" : "=v"(t) : "v"(x[k]), "v"(m));
      asm volatile("v_bcnt_u32_b32 %0, %1, %2" : "=v"(x[k]) : "v"(t), "v"(x[k]));
    }
  }
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) acc += x[k];
  out[tid] = acc;
}

hipError_t launch_valu_peak(const uint32_t *in, uint32_t *out, uint32_t blocks,
                            uint32_t iters, hipStream_t stream) {
  hipLaunchKernelGGL(valu_peak_kernel, dim3(blocks), dim3(256), 0, stream, in, out,
                     iters);
  return hipGetLastError();
}

/* ------------------------------------------------------------ libm probe */

/* Device values of the three libm restatements, for comparison with glibc.
 * expf is taken as the chain takes it: expf_lanes of a wave-uniform
 * argument, each lane's value in turn. */
__global__ __launch_bounds__(256) void libm_kernel(const float *x, uint32_t n,
                                                   float *e, float *l, float *l1m) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const Exp2Lanes t = exp2_lanes();
  const float v = i < n ? x[i] : 0.0f;
  float ev = 0.0f;
  for (uint32_t j = 0; j < 64u; j++) {
    const float r = expf_lanes(__uint_as_float(readlane32(__float_as_uint(v), j)), t);
    if (lane_id() == j) ev = r;
  }
  if (i >= n) return;
  e[i] = ev;
  l[i] = bbs_logf(v);
  l1m[i] = bbs_logf_of_double(1.0 - (double)v);
}

hipError_t launch_libm(const float *x, uint32_t n, float *e, float *l, float *l1m,
                       hipStream_t stream) {
  if (n == 0u) return hipSuccess;
  hipLaunchKernelGGL(libm_kernel, dim3((n + 255u) / 256u), dim3(256), 0, stream, x,
                     n, e, l, l1m);
  return hipGetLastError();
}

/* ------------------------------------------------------------ dispatch */

void shape_of(uint32_t n_genes, uint32_t n_cells, int *ns, int *cs, int *w,
              bool paired) {
  *ns = (n_genes + 1u) <= 64u ? 1 : 2;
  *cs = n_cells <= 64u ? 1 : (paired && n_cells <= 96u ? 3 : 2);
  *w = (int)((n_genes + 31u) / 32u);
  if (*w < 1) *w = 1;
  if (*ns == 2 && *w < 2) *w = 2;
}

uint32_t chain_table_bytes(uint32_t n_nodes, uint32_t n_cells) {
  int ns, cs, w;
  shape_of(n_nodes - 1u, n_cells, &ns, &cs, &w, true);
  const uint32_t bytes = 4u * ChainLds::of(n_nodes, n_cells, cs).total;
  return bytes <= kMaxGroupLds ? bytes : 0u;
}

uint32_t chain_half_table_bytes(uint32_t n_nodes, uint32_t n_cells) {
  int ns, cs, w;
  shape_of(n_nodes - 1u, n_cells, &ns, &cs, &w, true);
  if (cs != 2) return 0u;
  const uint32_t bytes = 4u * ChainLds::of<true>(n_nodes, n_cells, cs).total;
  return bytes <= kMaxGroupLds ? bytes : 0u;
}

uint32_t chain_static_lds() { return 0u; }

template <int NS, int CS, int W>
static hipError_t launch_chain_shape(const ChainArgs &a, bool trace, uint32_t table_bytes,
                                     bool half, hipStream_t stream) {
  const dim3 grid(a.n_chains), block(64);
  if (half) {
    if constexpr (CS == 2) {
      if (trace)
        hipLaunchKernelGGL((chain_kernel<NS, CS, W, true, true, true>), grid, block, table_bytes,
                           stream, a);
      else
        hipLaunchKernelGGL((chain_kernel<NS, CS, W, false, true, true>), grid, block,
                           table_bytes, stream, a);
      return hipGetLastError();
    }
    return hipErrorInvalidValue;
  }
  if (table_bytes) {
    if (trace)
      hipLaunchKernelGGL((chain_kernel<NS, CS, W, true, true>), grid, block, table_bytes, stream,
                         a);
    else
      hipLaunchKernelGGL((chain_kernel<NS, CS, W, false, true>), grid, block, table_bytes,
                         stream, a);
    return hipGetLastError();
  }
  if (trace)
    hipLaunchKernelGGL((chain_kernel<NS, CS, W, true, false>), grid, block, 0, stream, a);
  else
    hipLaunchKernelGGL((chain_kernel<NS, CS, W, false, false>), grid, block, 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_chain(const ChainArgs &a, bool trace, uint32_t table_bytes, bool half,
                        hipStream_t stream) {
  int ns, cs, w;
  shape_of(a.n_nodes - 1u, a.n_cells, &ns, &cs, &w, true);
  const uint32_t need = half ? ChainLds::of<true>(a.n_nodes, a.n_cells, cs).total
                             : ChainLds::of(a.n_nodes, a.n_cells, cs).total;
  if ((half && (cs != 2 || table_bytes == 0u)) || (table_bytes && table_bytes < 4u * need))
    return hipErrorInvalidValue;
#define X(NS_, CS_, W_)                                                        \
  if (ns == NS_ && cs == CS_ && w == W_)                                       \
    return launch_chain_shape<NS_, CS_, W_>(a, trace, table_bytes, half, stream);
  BBS_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

uint32_t chain_split_bytes(uint32_t n_nodes, uint32_t n_cells, uint32_t split) {
  if (split != 2u && split != 4u) return 0u;
  const uint32_t cs = (n_cells + split - 1u) / split;
  if (cs == 0u || (split - 1u) * cs >= n_cells) return 0u; /* a wave without cells */
  const uint32_t bytes = 4u * (256u + split * ChainLds::of(n_nodes, cs, 1).total);
  return bytes <= kMaxGroupLds ? bytes : 0u;
}

template <int NS, int W>
static hipError_t launch_chain_split_shape(const ChainArgs &a, bool trace, uint32_t split,
                                           uint32_t bytes, hipStream_t stream) {
  const dim3 grid(a.n_chains), block(64u * split);
#define XS(SP_)                                                                \
  if (split == SP_) {                                                          \
    if (trace)                                                                 \
      hipLaunchKernelGGL((chain_split_kernel<NS, W, true, SP_>), grid, block,  \
                         bytes, stream, a);                                    \
    else                                                                       \
      hipLaunchKernelGGL((chain_split_kernel<NS, W, false, SP_>), grid, block, \
                         bytes, stream, a);                                    \
    return hipGetLastError();                                                  \
  }
  XS(2u) XS(4u)
#undef XS
  return hipErrorInvalidValue;
}

hipError_t launch_chain_split(const ChainArgs &a, bool trace, uint32_t split,
                              hipStream_t stream) {
  const uint32_t bytes = chain_split_bytes(a.n_nodes, a.n_cells, split);
  if (bytes == 0u) return hipErrorInvalidValue;
  int ns, cs, w;
  shape_of(a.n_nodes - 1u, a.n_cells, &ns, &cs, &w, true);
#define X2(NS_, W_)                                                            \
  if (ns == NS_ && w == W_) return launch_chain_split_shape<NS_, W_>(a, trace, split, bytes, stream);
  X2(1, 1) X2(1, 2) X2(2, 2) X2(2, 3) X2(2, 4)
#undef X2
  return hipErrorInvalidValue;
}

uint32_t chain_spec_bytes(uint32_t n_nodes, uint32_t n_cells, uint32_t k) {
  if (k != 2u && k != 4u && k != 8u) return 0u;
  int ns, cs, w;
  shape_of(n_nodes - 1u, n_cells, &ns, &cs, &w, true);
  const uint32_t bytes = 4u * SpecLds::of(n_nodes, n_cells, cs, k).total;
  return bytes <= kMaxGroupLds ? bytes : 0u;
}

template <int NS, int CS, int W>
static hipError_t launch_chain_spec_shape(const ChainArgs &a, bool trace, uint32_t k,
                                          uint32_t bytes, hipStream_t stream) {
  const dim3 grid(a.n_chains), block(64u * k);
  if (trace)
    hipLaunchKernelGGL((chain_spec_kernel<NS, CS, W, true>), grid, block, bytes, stream, a);
  else
    hipLaunchKernelGGL((chain_spec_kernel<NS, CS, W, false>), grid, block, bytes, stream, a);
  return hipGetLastError();
}

hipError_t launch_chain_spec(const ChainArgs &a, bool trace, uint32_t k, hipStream_t stream) {
  const uint32_t bytes = chain_spec_bytes(a.n_nodes, a.n_cells, k);
  if (bytes == 0u) return hipErrorInvalidValue;
  int ns, cs, w;
  shape_of(a.n_nodes - 1u, a.n_cells, &ns, &cs, &w, true);
#define XP(NS_, CS_, W_)                                                       \
  if (ns == NS_ && cs == CS_ && w == W_)                                       \
    return launch_chain_spec_shape<NS_, CS_, W_>(a, trace, k, bytes, stream);
  BBS_SHAPES(XP)
#undef XP
  return hipErrorInvalidValue;
}

hipError_t launch_score(const ScoreArgs &a, bool counts, bool mfma, hipStream_t stream,
                        const uint32_t *cells32) {
  int ns, cs, w;
  shape_of(a.n_nodes - 1u, a.n_cells, &ns, &cs, &w, false);
  if (!counts && cells32 != nullptr) {
    /* matrix-core counts: one tree per wave, gene blocks x cell tiles */
    if (mfma) {
      const int kb = (int)((a.n_nodes + 31u) / 32u), jt = (int)((a.n_cells + 31u) / 32u);
      uint32_t groups = (a.n_trees + 3u) / 4u;
      if (groups > 16384u) groups = 16384u;
      if (groups == 0u) return hipSuccess;
#define XM(KB_, JT_)                                                           \
  if (kb == KB_ && jt == JT_) {                                                \
    hipLaunchKernelGGL((score_mfma_kernel<KB_, JT_>), dim3(groups), dim3(256), \
                       0, stream, a, cells32);                                 \
    return hipGetLastError();                                                  \
  }
      XM(1, 1) XM(1, 2) XM(1, 3) XM(1, 4) XM(2, 1) XM(2, 2) XM(2, 3) XM(2, 4)
      XM(3, 1) XM(3, 2) XM(3, 3) XM(3, 4) XM(4, 1) XM(4, 2) XM(4, 3) XM(4, 4)
#undef XM
    }
    const int cmax = a.n_cells <= 32u ? 32 : a.n_cells <= 64u ? 64 : a.n_cells <= 96u ? 96 : 128;
    const int wq = (int)((a.n_nodes - 1u + 31u) / 32u);
    uint32_t groups = (a.n_trees + 255u) / 256u;
    if (groups > 16384u) groups = 16384u;
    if (groups == 0u) return hipSuccess;
#define XL(CM_, W_)                                                            \
  if (cmax == CM_ && wq == W_) {                                               \
    hipLaunchKernelGGL((score_lanes_kernel<CM_, W_>), dim3(groups), dim3(256), \
                       0, stream, a, cells32);                                 \
    return hipGetLastError();                                                  \
  }
    XL(32, 1) XL(32, 2) XL(32, 3) XL(32, 4) XL(64, 1) XL(64, 2) XL(64, 3) XL(64, 4)
    XL(96, 1) XL(96, 2) XL(96, 3) XL(96, 4) XL(128, 1) XL(128, 2) XL(128, 3) XL(128, 4)
#undef XL
  }
  uint32_t blocks = a.n_trees < 65536u ? a.n_trees : 65536u;
  if (blocks == 0u) return hipSuccess;
  dim3 grid(blocks), block(64);
#define X(NS_, CS_, W_)                                                        \
  if (ns == NS_ && cs == CS_ && w == W_) {                                     \
    if (counts)                                                                \
      hipLaunchKernelGGL((score_kernel<NS_, CS_, W_, true>), grid, block, 0,   \
                         stream, a);                                           \
    else                                                                       \
      hipLaunchKernelGGL((score_kernel<NS_, CS_, W_, false>), grid, block, 0,  \
                         stream, a);                                           \
    return hipGetLastError();                                                  \
  }
  BBS_FULL_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}

uint32_t batch_wave_floats(uint32_t shape, uint32_t max_nodes, uint32_t max_cells) {
  const int cs = (int)(shape / 10u % 10u);
  return ChainLds::of(max_nodes, max_cells, cs).total;
}

template <int NS, int CS, int W>
static hipError_t launch_items_shape(const BatchArgs &a, uint32_t first, uint32_t count,
                                     uint32_t wpg, uint32_t wave_floats, uint32_t wps,
                                     hipStream_t stream) {
  const uint32_t groups = (count + wpg - 1u) / wpg;
  if (wave_floats) {
    if (wps == 2u)
      hipLaunchKernelGGL((batch_items_kernel<NS, CS, W, true, 2>), dim3(groups),
                         dim3(64u * wpg), 4u * wave_floats * wpg, stream, a, first, count,
                         wave_floats);
    else
      hipLaunchKernelGGL((batch_items_kernel<NS, CS, W, true, 1>), dim3(groups),
                         dim3(64u * wpg), 4u * wave_floats * wpg, stream, a, first, count,
                         wave_floats);
    return hipGetLastError();
  }
  /* the full scorer fits two waves per SIMD in every layout as it is */
  hipLaunchKernelGGL((batch_items_kernel<NS, CS, W, false, 1>), dim3(groups), dim3(64u * wpg),
                     0, stream, a, first, count, 0u);
  return hipGetLastError();
}

template <int NS, int CS, int W>
static hipError_t items_occupancy_shape(uint32_t wpg, uint32_t wave_floats, uint32_t wps,
                                        int *groups) {
  const size_t lds = 4ull * wave_floats * wpg;
  if (!wave_floats)
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        groups, batch_items_kernel<NS, CS, W, false, 1>, (int)(64u * wpg), 0);
  if (wps == 2u)
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        groups, batch_items_kernel<NS, CS, W, true, 2>, (int)(64u * wpg), lds);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      groups, batch_items_kernel<NS, CS, W, true, 1>, (int)(64u * wpg), lds);
}

hipError_t batch_items_occupancy(uint32_t shape, uint32_t waves_per_group, uint32_t wave_floats,
                                 uint32_t waves_per_simd, uint32_t *waves_per_cu) {
  int groups = 0;
  hipError_t e = hipErrorInvalidValue;
#define XO(NS_, CS_, W_)                                                       \
  if (shape == (uint32_t)BBS_SHAPE_CODE(NS_, CS_, W_))                        \
    e = items_occupancy_shape<NS_, CS_, W_>(waves_per_group, wave_floats, waves_per_simd, &groups);
  BBS_SHAPES(XO)
#undef XO
  *waves_per_cu = e == hipSuccess ? (uint32_t)groups * waves_per_group : 0u;
  return e;
}

hipError_t launch_batch_items(const BatchArgs &a, uint32_t shape, uint32_t first,
                              uint32_t count, uint32_t waves_per_group, uint32_t wave_floats,
                              uint32_t waves_per_simd, hipStream_t stream) {
  if (count == 0u) return hipSuccess;
  if (waves_per_group < 1u || waves_per_group > 4u) return hipErrorInvalidValue;
  if (waves_per_simd < 1u || waves_per_simd > 2u) return hipErrorInvalidValue;
  if (4ull * wave_floats * waves_per_group > kMaxGroupLds) return hipErrorInvalidValue;
#define XB(NS_, CS_, W_)                                                       \
  if (shape == (uint32_t)BBS_SHAPE_CODE(NS_, CS_, W_))                        \
    return launch_items_shape<NS_, CS_, W_>(a, first, count, waves_per_group, wave_floats,  \
                                            waves_per_simd, stream);
  BBS_SHAPES(XB)
#undef XB
  return hipErrorInvalidValue;
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t *p, uint32_t value, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = value;
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t value, uint32_t n, hipStream_t stream) {
  if (n == 0u) return hipSuccess;
  uint32_t blocks = (n + 255u) / 256u;
  if (blocks > 1024u) blocks = 1024u;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, stream, p, value, n);
  return hipGetLastError();
}

hipError_t launch_random_trees(const TreeGenArgs &a, hipStream_t stream) {
  uint32_t blocks = a.n_trees < 65536u ? a.n_trees : 65536u;
  if (blocks == 0u) return hipSuccess;
  hipLaunchKernelGGL(random_trees_kernel, dim3(blocks), dim3(64), 0, stream, a);
  return hipGetLastError();
}

} // namespace bbs
