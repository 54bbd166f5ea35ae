/*
 * C-ABI of the MI355X chain search (include/bbscite.h): context lifetime,
 * device memory, kernel launches and the host-side helpers (CSV parsing,
 * encoding, synthetic data, Pruefer decoding, Graphviz/Newick text).
 *
 * Reference counterparts are named on each entry point in the header; the
 * host-side restatements here follow src/main.cpp:52-144,
 * src/Parameters.hpp:39-45,451-612, src/TreeScorer.hpp:92-126 and
 * src/MutationTree.hpp:379-510,549-596,799-894.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <thread>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/bbscite.h"
#include "bbs_kernels.h"
#include "bbs_libm.h"
#include "bbs_rng.h"

namespace {

thread_local std::string g_create_error;

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

} // namespace

struct bbs_ctx {
  bbs_params params;
  uint32_t n_cells = 0, n_genes = 0, n_nodes = 0;
  float L00 = 0.f, L10 = 0.f;
  bbs::PriorArgs prior{};
  bbs_move_probs probs{};
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf cells, cells32, am, dm, beta, score, best_am, best_dm, best_beta,
      best_score, best_step, trace, prop_am, prop_dm;
  uint32_t n_traced = 0;
  uint32_t n_prop_traced = 0; /* chains whose proposals are recorded */
  int scoring_mode = BBS_SCORING_AUTO;
  uint32_t chain_split = 0u; /* waves per chain: 0 auto, 1, 2, 4 */
  uint32_t speculation = 0u; /* speculative waves per chain: 0 auto, 1, 2, 4, 8 */
  uint32_t n_cus = 0, lds_per_cu = 0;
  bool states_set = false;   /* initial states present on the device */
  bool score_known = false;  /* initial scores present on the device  */
  bool begun = false;        /* bests reset, steps counted from 0     */
  uint32_t steps_done = 0;
  bool ran = false;
  void *scratch = nullptr;   /* this device's bbs_score_trees_noalloc staging */
  std::string scratch_err;   /* why scratch is null                   */
  std::string err;
};

namespace {

int fail(bbs_ctx *ctx, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx)
    ctx->err = buf;
  else
    g_create_error = buf;
  return code;
}

#define HIP_TRY(ctx, expr)                                                     \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail((ctx), BBS_ERR_HIP, "%s failed: %s", #expr,                  \
                  hipGetErrorString(e_));                                      \
  } while (0)

int dev_alloc(bbs_ctx *ctx, DevBuf &b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return BBS_OK;
  b.release();
  if (bytes == 0) return BBS_OK;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(ctx, e == hipErrorOutOfMemory ? BBS_ERR_BUDGET : BBS_ERR_HIP,
                "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  b.bytes = bytes;
  return BBS_OK;
}

/* log(tgamma(x)) as the reference computes it in float
 * (src/TreeScorer.hpp:123-125). Where tgammaf overflows (x > ~35.04) the
 * reference's prior is inf - inf = NaN; this library uses log-gamma in double
 * instead (the only deliberate numeric deviation; same rule in the oracle). */
float lgamma_term(float x) {
  float g = tgammaf(x);
  if (std::isinf(g) && g > 0 && x > 0) return (float)lgamma((double)x);
  return logf(g);
}

bool is_gfx950(int dev) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

/* ---------------- host tree helpers ---------------- */

int pruefer_to_parent(const uint32_t *code, uint32_t len, uint32_t *parent) {
  uint32_t n = len + 2u;
  if (n > BBS_MAX_NODES) return BBS_ERR_SIZE;
  uint32_t degree[BBS_MAX_NODES];
  for (uint32_t i = 0; i < n; i++) degree[i] = 1u;
  for (uint32_t i = 0; i < len; i++) {
    if (code[i] >= n) return BBS_ERR_ARG;
    degree[code[i]]++;
  }
  for (uint32_t i = 0; i < n; i++) parent[i] = n - 1u;
  /* attach the smallest remaining leaf to the next code entry
   * (src/MutationTree.hpp:454-463) */
  for (uint32_t i = 0; i < len; i++) {
    for (uint32_t j = 0; j < n; j++) {
      if (degree[j] == 1u) {
        parent[j] = code[i];
        degree[code[i]]--;
        degree[j]--;
        break;
      }
    }
  }
  return BBS_OK;
}

int parent_to_rows(const uint32_t *parent, uint32_t n, uint64_t *am,
                   uint64_t *dm) {
  if (n < 1u || n > BBS_MAX_NODES) return BBS_ERR_SIZE;
  const uint32_t root = n - 1u;
  if (am) std::memset(am, 0, sizeof(uint64_t) * 2u * n);
  std::vector<uint64_t> d(2u * n, 0ull);
  for (uint32_t i = 0; i < n; i++) {
    uint32_t x = i, guard = 0;
    while (x != root) {
      if (x >= n || guard++ > n) return BBS_ERR_ARG; /* not a tree */
      d[2u * i + (x >> 6)] |= 1ull << (x & 63u);
      x = parent[x];
    }
    d[2u * i + (root >> 6)] |= 1ull << (root & 63u);
  }
  if (dm) std::memcpy(dm, d.data(), sizeof(uint64_t) * 2u * n);
  if (am)
    for (uint32_t y = 0; y < n; y++)
      for (uint32_t x = 0; x < n; x++)
        if ((d[2u * y + (x >> 6)] >> (x & 63u)) & 1u)
          am[2u * x + (y >> 6)] |= 1ull << (y & 63u);
  return BBS_OK;
}

inline bool rbit(const uint64_t *row, uint32_t i) {
  return (row[i >> 6] >> (i & 63u)) & 1u;
}

/* get_parent of src/MutationTree.hpp:549-596 on dm rows. */
uint32_t rows_parent(const uint64_t *dm, uint32_t n, uint32_t child) {
  uint32_t parent = 0;
  for (uint32_t p = 0; p < n; p++) {
    const uint64_t *pa = dm + 2u * p, *ca = dm + 2u * child;
    if (!rbit(ca, p)) continue;
    if (p == child) {
      if (child == n - 1u) parent = p;
      continue;
    }
    bool same = true;
    for (uint32_t i = 0; i < n && same; i++)
      if (i != child && rbit(pa, i) != rbit(ca, i)) same = false;
    if (same) parent = p;
  }
  return parent;
}

struct TextOut {
  char *buf;
  uint64_t cap, len;
  void put(const std::string &s) {
    for (char c : s) {
      if (len + 1u < cap) buf[len] = c;
      len++;
    }
  }
  void finish() {
    if (cap) buf[len + 1u < cap ? len : cap - 1u] = 0;
  }
};

void newick_rec(const std::vector<std::vector<uint32_t>> &children, uint32_t node,
                TextOut &out, uint32_t depth) {
  if (depth > BBS_MAX_NODES) return;
  const auto &ch = children[node];
  if (!ch.empty()) {
    out.put("(");
    for (size_t i = 0; i < ch.size(); i++) {
      newick_rec(children, ch[i], out, depth + 1u);
      if (i + 1u != ch.size()) out.put(",");
    }
    out.put(")");
  }
  out.put(std::to_string(node));
}

void gen_tree_parent(uint64_t seed, uint32_t job, uint32_t chain,
                     uint32_t purpose, uint32_t n, uint32_t *parent) {
  uint32_t code[BBS_MAX_NODES];
  for (uint32_t i = 0; i + 2u < n; i++) {
    bbs_u32x4 b = bbs_philox_block(seed, job, chain, i >> 2, purpose);
    code[i] = bbs_uint_below(b.x[i & 3u], n);
  }
  pruefer_to_parent(code, n - 2u, parent);
}

} // namespace

/* ======================================================================= */

extern "C" {

void bbs_params_default(bbs_params *p) {
  if (!p) return;
  /* src/Parameters.hpp:39-45 */
  p->alpha_mean = 6.04e-5f;
  p->beta_mean = 0.4309f;
  p->beta_sd = 0.1f;
  p->gamma = 1.0f;
  p->prob_beta_change = 0.0f;
  p->prob_prune_n_reattach = 0.55f;
  p->prob_swap_nodes = 0.4f;
  p->prob_swap_subtrees = 0.05f;
  p->beta_jump_scaling_chi = 10.0f;
  p->n_chains = 0;
  p->n_steps = 0;
  p->seed = 0;
}

const char *bbs_last_error(const bbs_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int bbs_create(const bbs_params *params, uint32_t n_cells, uint32_t n_genes,
               const uint64_t *is_mutated, const uint64_t *is_known,
               bbs_ctx **out) {
  if (!out) return fail(nullptr, BBS_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!params || !is_mutated || !is_known)
    return fail(nullptr, BBS_ERR_ARG, "params and data must not be NULL");
  if (n_genes < 2u || n_genes > BBS_MAX_GENES)
    return fail(nullptr, BBS_ERR_SIZE,
                "the library supports 2 to %u genes; the input has %u",
                BBS_MAX_GENES, n_genes);
  if (n_cells < 1u || n_cells > BBS_MAX_CELLS)
    return fail(nullptr, BBS_ERR_SIZE,
                "the library supports 1 to %u cells; the input has %u",
                BBS_MAX_CELLS, n_cells);
  const bbs_params &p = *params;
  if (!(p.alpha_mean > 0.f && p.alpha_mean < 1.f))
    return fail(nullptr, BBS_ERR_ARG, "alpha (-fd) must lie in (0, 1)");
  if (!(p.beta_mean > 0.f && p.beta_mean < 1.f))
    return fail(nullptr, BBS_ERR_ARG, "beta (-ad) must lie in (0, 1)");
  if (!(p.beta_sd > 0.f))
    return fail(nullptr, BBS_ERR_ARG, "beta sd (-sd) must be positive");
  if (!(p.prob_beta_change >= 0.f && p.prob_prune_n_reattach >= 0.f &&
        p.prob_swap_nodes >= 0.f && p.prob_swap_subtrees >= 0.f))
    return fail(nullptr, BBS_ERR_ARG, "Move probabilities may not be negative.");
  if (p.prob_beta_change + p.prob_prune_n_reattach + p.prob_swap_nodes +
          p.prob_swap_subtrees ==
      0.f)
    return fail(nullptr, BBS_ERR_ARG,
                "The sum of the move type probabilities is zero.");
  if (!(p.beta_jump_scaling_chi > 0.f))
    return fail(nullptr, BBS_ERR_ARG, "chi (-x) must be positive");
  if (p.n_chains == 0u)
    return fail(nullptr, BBS_ERR_ARG, "at least one chain (-r) is needed");

  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0)
    return fail(nullptr, BBS_ERR_NODEV, "no HIP device available");
  int dev = -1;
  for (int d = 0; d < n_dev && dev < 0; d++)
    if (is_gfx950(d)) dev = d;
  if (dev < 0) return fail(nullptr, BBS_ERR_NODEV, "no gfx950 device found");

  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return fail(nullptr, BBS_ERR_NODEV, "hipGetDeviceProperties failed");

  bbs_ctx *ctx = new bbs_ctx();
  ctx->params = p;
  ctx->n_cells = n_cells;
  ctx->n_genes = n_genes;
  ctx->n_nodes = n_genes + 1u;
  ctx->device = dev;
  ctx->n_cus = (uint32_t)prop.multiProcessorCount;
  ctx->lds_per_cu = (uint32_t)prop.maxSharedMemoryPerMultiProcessor;

  /* TreeScorer constructor (src/TreeScorer.hpp:92-117) */
  ctx->L00 = (float)std::log(1.0 - (double)p.alpha_mean);
  ctx->L10 = logf(p.alpha_mean);
  double a = ((double)(1.0f - p.beta_mean) * std::pow((double)p.beta_mean, 2.0) /
              std::pow((double)p.beta_sd, 2.0)) -
             (double)p.beta_mean;
  float bpa = (float)a;
  float bpb = bpa * ((1.0f / p.beta_mean) - 1.0f);
  ctx->prior.am1 = bpa - 1.0f;
  ctx->prior.bm1 = bpb - 1.0f;
  ctx->prior.lg_ab = lgamma_term(bpa + bpb);
  ctx->prior.lg_a = lgamma_term(bpa);
  ctx->prior.lg_b = lgamma_term(bpb);

  /* normalised move probabilities (src/Parameters.hpp:451-541,609-612) */
  float sum = p.prob_beta_change + p.prob_prune_n_reattach + p.prob_swap_nodes +
              p.prob_swap_subtrees;
  float pb = p.prob_beta_change / sum, pp = p.prob_prune_n_reattach / sum,
        ps = p.prob_swap_nodes / sum;
  ctx->probs.t_beta = pb;
  ctx->probs.t_prune = pb + pp;
  ctx->probs.t_swap = pb + pp + ps;
  ctx->probs.beta_jump_sd = p.beta_sd / p.beta_jump_scaling_chi;

  /* per-cell K&M and K&~M restricted to the genes (src/main.cpp:103-144) */
  std::vector<uint64_t> cells(4u * n_cells);
  uint64_t gmask[2] = {n_genes >= 64u ? ~0ull : ((1ull << n_genes) - 1ull),
                       n_genes <= 64u ? 0ull : ((1ull << (n_genes - 64u)) - 1ull)};
  for (uint32_t c = 0; c < n_cells; c++) {
    for (int wd = 0; wd < 2; wd++) {
      uint64_t K = is_known[2u * c + wd] & gmask[wd];
      uint64_t M = is_mutated[2u * c + wd];
      cells[4u * c + wd] = K & M;
      cells[4u * c + 2u + wd] = K & ~M;
    }
  }

  /* the same data as [128][10] words for the lane-per-tree scorer */
  std::vector<uint32_t> cells32(128u * 10u, 0u);
  for (uint32_t c = 0; c < n_cells; c++) {
    uint32_t a = 0, b = 0;
    for (int w = 0; w < 4; w++) {
      uint32_t km = (uint32_t)(cells[4u * c + (w >> 1)] >> (32 * (w & 1)));
      uint32_t knm = (uint32_t)(cells[4u * c + 2u + (w >> 1)] >> (32 * (w & 1)));
      cells32[10u * c + w] = km;
      cells32[10u * c + 4u + w] = knm;
      b += (uint32_t)__builtin_popcount(km);
      a += (uint32_t)__builtin_popcount(knm);
    }
    float fa = (float)a, fb = (float)b;
    std::memcpy(&cells32[10u * c + 8u], &fa, 4);
    std::memcpy(&cells32[10u * c + 9u], &fb, 4);
  }

  int rc;
  hipError_t e = hipSetDevice(dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
  if (e != hipSuccess) {
    g_create_error = std::string("HIP setup failed: ") + hipGetErrorString(e);
    bbs_destroy(ctx);
    return BBS_ERR_HIP;
  }
  /* the allocation-free scorer's staging address on this device (looked up
   * once per device; only bbs_score_trees_noalloc needs it) */
  {
    const hipError_t es = bbs::noalloc_scratch(dev, &ctx->scratch);
    if (es != hipSuccess) {
      ctx->scratch = nullptr;
      ctx->scratch_err = hipGetErrorString(es);
      (void)hipGetLastError();
    }
  }
  const size_t rows = (size_t)p.n_chains * ctx->n_nodes * 2u * sizeof(uint64_t);
  const size_t per_chain = (size_t)p.n_chains * sizeof(float);
  if ((rc = dev_alloc(ctx, ctx->cells, cells.size() * sizeof(uint64_t))) ||
      (rc = dev_alloc(ctx, ctx->cells32, cells32.size() * sizeof(uint32_t))) ||
      (rc = dev_alloc(ctx, ctx->am, rows)) || (rc = dev_alloc(ctx, ctx->dm, rows)) ||
      (rc = dev_alloc(ctx, ctx->best_am, rows)) ||
      (rc = dev_alloc(ctx, ctx->best_dm, rows)) ||
      (rc = dev_alloc(ctx, ctx->beta, per_chain)) ||
      (rc = dev_alloc(ctx, ctx->score, per_chain)) ||
      (rc = dev_alloc(ctx, ctx->best_beta, per_chain)) ||
      (rc = dev_alloc(ctx, ctx->best_score, per_chain)) ||
      (rc = dev_alloc(ctx, ctx->best_step, per_chain))) {
    g_create_error = ctx->err;
    bbs_destroy(ctx);
    return rc;
  }
  e = hipMemcpy(ctx->cells.p, cells.data(), cells.size() * sizeof(uint64_t),
                hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(ctx->cells32.p, cells32.data(), cells32.size() * sizeof(uint32_t),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    g_create_error = std::string("hipMemcpy failed: ") + hipGetErrorString(e);
    bbs_destroy(ctx);
    return BBS_ERR_HIP;
  }
  *out = ctx;
  return BBS_OK;
}

void bbs_destroy(bbs_ctx *ctx) {
  if (!ctx) return;
  if (ctx->device >= 0) (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  DevBuf *bufs[] = {&ctx->cells,     &ctx->cells32,    &ctx->am,         &ctx->dm,
                    &ctx->beta,      &ctx->score,      &ctx->best_am,
                    &ctx->best_dm,   &ctx->best_beta,  &ctx->best_score,
                    &ctx->best_step, &ctx->trace,     &ctx->prop_am,
                    &ctx->prop_dm};
  for (DevBuf *b : bufs) b->release();
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int bbs_set_initial_states(bbs_ctx *ctx, const uint64_t *am, const uint64_t *dm,
                           const float *beta, const float *score) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (!am || !dm || !beta)
    return fail(ctx, BBS_ERR_ARG, "am, dm and beta must not be NULL");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t rows = (size_t)ctx->params.n_chains * ctx->n_nodes * 2u * sizeof(uint64_t);
  const size_t per_chain = (size_t)ctx->params.n_chains * sizeof(float);
  HIP_TRY(ctx, hipMemcpy(ctx->am.p, am, rows, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->dm.p, dm, rows, hipMemcpyHostToDevice));
  HIP_TRY(ctx, hipMemcpy(ctx->beta.p, beta, per_chain, hipMemcpyHostToDevice));
  if (score)
    HIP_TRY(ctx, hipMemcpy(ctx->score.p, score, per_chain, hipMemcpyHostToDevice));
  ctx->states_set = true;
  ctx->score_known = score != nullptr;
  ctx->begun = false;
  ctx->ran = false;
  return BBS_OK;
}

int bbs_set_trace(bbs_ctx *ctx, uint32_t n_traced_chains) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (n_traced_chains > ctx->params.n_chains)
    return fail(ctx, BBS_ERR_ARG, "cannot trace %u of %u chains", n_traced_chains,
                ctx->params.n_chains);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  int rc = dev_alloc(ctx, ctx->trace,
                     (size_t)n_traced_chains * ctx->params.n_steps *
                         sizeof(bbs_step_trace));
  if (rc) return rc;
  ctx->n_traced = n_traced_chains;
  return BBS_OK;
}

int bbs_set_proposal_trace(bbs_ctx *ctx, uint32_t n_traced_chains) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (n_traced_chains > ctx->params.n_chains)
    return fail(ctx, BBS_ERR_ARG, "cannot trace %u of %u chains", n_traced_chains,
                ctx->params.n_chains);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t bytes = (size_t)n_traced_chains * ctx->params.n_steps * ctx->n_nodes * 2u *
                       sizeof(uint64_t);
  int rc;
  if ((rc = dev_alloc(ctx, ctx->prop_am, bytes)) || (rc = dev_alloc(ctx, ctx->prop_dm, bytes)))
    return rc;
  ctx->n_prop_traced = n_traced_chains;
  return BBS_OK;
}

/* Start a run: initial states (drawn from the seed unless set), no bests. */
int bbs_begin(bbs_ctx *ctx) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const uint32_t R = ctx->params.n_chains;
  if (!ctx->states_set) {
    bbs::TreeGenArgs g{};
    g.am = (uint64_t *)ctx->am.p;
    g.dm = (uint64_t *)ctx->dm.p;
    g.beta = (float *)ctx->beta.p;
    g.beta_value = ctx->params.beta_mean;
    g.seed = ctx->params.seed;
    g.job = 0u;
    g.chain_offset = 0u;
    g.purpose = BBS_PURPOSE_INIT_TREE;
    g.n_trees = R;
    g.n_nodes = ctx->n_nodes;
    HIP_TRY(ctx, bbs::launch_random_trees(g, ctx->stream));
    ctx->states_set = true;
    ctx->score_known = false;
  }
  std::vector<uint32_t> steps(R, 0xffffffffu);
  std::vector<float> scores(R, -INFINITY);
  HIP_TRY(ctx, hipMemcpyAsync(ctx->best_step.p, steps.data(), R * sizeof(uint32_t),
                              hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->best_score.p, scores.data(), R * sizeof(float),
                              hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->best_beta.p, scores.data(), R * sizeof(float),
                              hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  ctx->begun = true;
  ctx->steps_done = 0;
  ctx->ran = false;
  return BBS_OK;
}

/* Advance every chain by n steps (continuing the current run). */
/* LDS bytes of the pair-score table for this context's launches, or 0 for
 * full rescoring. Automatic mode takes the table when it leaves every CU
 * room for as many chain waves as the launch can give it (at most one per
 * SIMD), so the table never costs concurrency. */
/* The pair-score table of this context's plain launches: its LDS bytes per
 * chain wave (0: every step rescores every pair) and whether it holds the
 * first 64 cells only (half; the second cell slot is then rescored every
 * step). AUTO keeps the whole table when it costs no concurrent chains (as
 * many chains run at once with it as without, one per SIMD), else the half
 * table on the same terms (over 96 cells) when every chain fits one per
 * SIMD, else none. Beyond that, full rescoring runs two waves per SIMD
 * without LDS: at 127 x 128 with 2,048 chains it took 228.5 ms per 10,000
 * steps against 383.8 ms with the half table. */
struct TableChoice {
  uint32_t bytes;
  bool half;
};

static TableChoice chain_table_choice(const bbs_ctx *ctx) {
  if (ctx->scoring_mode == BBS_SCORING_FULL) return {0u, false};
  if (ctx->scoring_mode == BBS_SCORING_HALF_TABLE)
    return {bbs::chain_half_table_bytes(ctx->n_nodes, ctx->n_cells), true};
  const uint32_t t = bbs::chain_table_bytes(ctx->n_nodes, ctx->n_cells);
  if (t == 0u || ctx->scoring_mode == BBS_SCORING_INCREMENTAL) return {t, false};
  const uint32_t cus = ctx->n_cus ? ctx->n_cus : 1u;
  uint32_t want = (ctx->params.n_chains + cus - 1u) / cus;
  if (want > 4u) want = 4u;
  if (want == 0u) want = 1u;
  auto fits = [&](uint32_t bytes) {
    return (uint64_t)(bytes + bbs::chain_static_lds()) * want <= ctx->lds_per_cu;
  };
  if (fits(t)) return {t, false};
  const uint32_t h = bbs::chain_half_table_bytes(ctx->n_nodes, ctx->n_cells);
  if (h != 0u && ctx->params.n_chains <= 4u * cus && fits(h)) return {h, true};
  return {0u, false};
}

/* The whole table's bytes if this context's launches keep one, else 0. */
static uint32_t chain_table(const bbs_ctx *ctx) {
  const TableChoice c = chain_table_choice(ctx);
  return c.half ? 0u : c.bytes;
}

int bbs_set_scoring_mode(bbs_ctx *ctx, int mode) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (mode != BBS_SCORING_AUTO && mode != BBS_SCORING_FULL && mode != BBS_SCORING_INCREMENTAL &&
      mode != BBS_SCORING_HALF_TABLE)
    return fail(ctx, BBS_ERR_ARG, "unknown scoring mode %d", mode);
  if (mode == BBS_SCORING_HALF_TABLE &&
      bbs::chain_half_table_bytes(ctx->n_nodes, ctx->n_cells) == 0u)
    return fail(ctx, BBS_ERR_SIZE,
                "no half table for %u cells (it serves datasets of over 96 cells)", ctx->n_cells);
  if (mode == BBS_SCORING_INCREMENTAL && bbs::chain_table_bytes(ctx->n_nodes, ctx->n_cells) == 0u)
    return fail(ctx, BBS_ERR_SIZE,
                "no incremental scorer for %u genes x %u cells (table over the LDS of a workgroup)",
                ctx->n_genes, ctx->n_cells);
  ctx->scoring_mode = mode;
  return BBS_OK;
}

/* Waves per chain of this context's launches (bbs_set_chain_split). */
static uint32_t chain_split(const bbs_ctx *ctx) {
  /* the split keeps a whole table */
  if (ctx->scoring_mode == BBS_SCORING_FULL || ctx->scoring_mode == BBS_SCORING_HALF_TABLE)
    return 1u;
  if (ctx->chain_split != 0u) return ctx->chain_split;
  if (chain_table(ctx) == 0u) return 1u;
  /* Splitting the cells shortens a wave's step only where one wave would
   * hold two cell slots (over 64 cells): lanes already score cells in
   * parallel, so fewer cells per wave alone removes no instruction. Measured
   * with 8 chains: 63 x 64 loses 11 % with 2 waves, 94 x 95 gains 5 %,
   * 127 x 128 17 %; 4 waves never beat 2. Speculation gains more where it
   * fits; the split serves launches of 257-512 chains over 64 cells. */
  const uint64_t simds = 4ull * (ctx->n_cus ? ctx->n_cus : 1u);
  if (ctx->n_cells > 64u && (uint64_t)ctx->params.n_chains * 2u <= simds &&
      bbs::chain_split_bytes(ctx->n_nodes, ctx->n_cells, 2u) != 0u)
    return 2u;
  return 1u;
}

/* Speculative waves per chain of this context's launches (bbs_set_speculation).
 * Automatically the largest k of 8, 4 and 2 whose waves run at most two per
 * SIMD, with every chain's LDS at once, while the table is in use. Measured
 * per launch of 10,000-20,000 steps with 10 % beta moves, against one wave
 * per chain: 63 x 64 with 128 / 256 / 512 / 1,024 chains 28 / 29 / 25 / 4 %
 * faster (k = 8 / 8 / 4 / 2), 94 x 95 with 128 / 256 chains 28 / 29 %;
 * where two chains' slabs do not fit a CU (94 x 95 and 127 x 128 from 512
 * chains) every k was slower, since fewer chains then run at once. */
static uint32_t chain_speculation(const bbs_ctx *ctx) {
  /* speculation keeps a whole table */
  if (ctx->scoring_mode == BBS_SCORING_FULL || ctx->scoring_mode == BBS_SCORING_HALF_TABLE)
    return 1u;
  if (ctx->speculation != 0u) return ctx->speculation;
  if (chain_table(ctx) == 0u) return 1u;
  const uint64_t cus = ctx->n_cus ? ctx->n_cus : 1u;
  const uint64_t chains = ctx->params.n_chains;
  const uint64_t per_cu = (chains + cus - 1u) / cus;
  for (uint32_t k = 8u; k >= 2u; k /= 2u) {
    const uint64_t bytes = bbs::chain_spec_bytes(ctx->n_nodes, ctx->n_cells, k);
    if (bytes != 0u && chains * k <= 8u * cus && per_cu * k <= 8u &&
        per_cu * bytes <= ctx->lds_per_cu)
      return k;
  }
  return 1u;
}

int bbs_set_speculation(bbs_ctx *ctx, uint32_t waves_per_chain) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  const uint32_t k = waves_per_chain;
  if (k != 0u && k != 1u && k != 2u && k != 4u && k != 8u)
    return fail(ctx, BBS_ERR_ARG, "speculative waves per chain must be 0 (auto), 1, 2, 4 or 8, not %u",
                k);
  if (k > 1u && bbs::chain_spec_bytes(ctx->n_nodes, ctx->n_cells, k) == 0u)
    return fail(ctx, BBS_ERR_SIZE,
                "no %u-wave speculative chain for %u genes x %u cells (over the LDS of a "
                "workgroup)",
                k, ctx->n_genes, ctx->n_cells);
  ctx->speculation = k;
  return BBS_OK;
}

int bbs_set_chain_split(bbs_ctx *ctx, uint32_t waves_per_chain) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  const uint32_t k = waves_per_chain;
  if (k != 0u && k != 1u && k != 2u && k != 4u)
    return fail(ctx, BBS_ERR_ARG, "waves per chain must be 0 (auto), 1, 2 or 4, not %u", k);
  if (k > 1u && bbs::chain_split_bytes(ctx->n_nodes, ctx->n_cells, k) == 0u)
    return fail(ctx, BBS_ERR_SIZE,
                "no %u-wave split for %u genes x %u cells (LDS of a workgroup, or a wave "
                "without cells)",
                k, ctx->n_genes, ctx->n_cells);
  ctx->chain_split = k;
  return BBS_OK;
}

int bbs_advance(bbs_ctx *ctx, uint32_t n, float *makespan_ms) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (!ctx->begun) {
    int rc = bbs_begin(ctx);
    if (rc) return rc;
  }
  if ((uint64_t)ctx->steps_done + n > ctx->params.n_steps)
    return fail(ctx, BBS_ERR_ARG, "advance beyond n_steps (%u + %u > %u)",
                ctx->steps_done, n, ctx->params.n_steps);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  bbs::ChainArgs a{};
  a.cells = (const uint64_t *)ctx->cells.p;
  a.am = (uint64_t *)ctx->am.p;
  a.dm = (uint64_t *)ctx->dm.p;
  a.beta = (float *)ctx->beta.p;
  a.score = (float *)ctx->score.p;
  a.best_am = (uint64_t *)ctx->best_am.p;
  a.best_dm = (uint64_t *)ctx->best_dm.p;
  a.best_beta = (float *)ctx->best_beta.p;
  a.best_score = (float *)ctx->best_score.p;
  a.best_step = (uint32_t *)ctx->best_step.p;
  a.trace = (bbs_step_trace *)ctx->trace.p;
  a.n_traced = ctx->n_traced;
  a.trace_stride = ctx->params.n_steps;
  a.prop_am = (uint64_t *)ctx->prop_am.p;
  a.prop_dm = (uint64_t *)ctx->prop_dm.p;
  a.n_prop_traced = ctx->n_prop_traced;
  a.seed = ctx->params.seed;
  a.job = 0u;
  a.chain_offset = 0u;
  a.n_chains = ctx->params.n_chains;
  a.step_begin = ctx->steps_done;
  a.step_end = ctx->steps_done + n;
  a.n_cells = ctx->n_cells;
  a.n_nodes = ctx->n_nodes;
  a.score_init = (ctx->steps_done == 0u && !ctx->score_known) ? 1u : 0u;
  a.L00 = ctx->L00;
  a.L10 = ctx->L10;
  a.gamma = ctx->params.gamma;
  a.prior = ctx->prior;
  a.probs = ctx->probs;
  const TableChoice table = chain_table_choice(ctx);
  const uint32_t spec = chain_speculation(ctx);
  const uint32_t split = spec > 1u ? 1u : chain_split(ctx);
  const bool traced = ctx->n_traced > 0u || ctx->n_prop_traced > 0u;
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (a.n_chains > 0u && spec > 1u)
    HIP_TRY(ctx, bbs::launch_chain_spec(a, traced, spec, ctx->stream));
  else if (a.n_chains > 0u && split > 1u)
    HIP_TRY(ctx, bbs::launch_chain_split(a, traced, split, ctx->stream));
  else if (a.n_chains > 0u)
    HIP_TRY(ctx, bbs::launch_chain(a, traced, table.bytes, table.half, ctx->stream));
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (makespan_ms) *makespan_ms = ms;
  ctx->steps_done += n;
  ctx->score_known = true;
  ctx->ran = ctx->steps_done == ctx->params.n_steps;
  return BBS_OK;
}

int bbs_run(bbs_ctx *ctx, float *makespan_ms) {
  int rc = bbs_begin(ctx);
  if (rc) return rc;
  return bbs_advance(ctx, ctx->params.n_steps, makespan_ms);
}

int bbs_get_chain_results(bbs_ctx *ctx, float *best_score, uint32_t *best_step,
                          float *final_score, float *final_beta) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (!ctx->begun) return fail(ctx, BBS_ERR_STATE, "no run yet");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t n = ctx->params.n_chains;
  if (best_score)
    HIP_TRY(ctx, hipMemcpy(best_score, ctx->best_score.p, n * 4u, hipMemcpyDeviceToHost));
  if (best_step)
    HIP_TRY(ctx, hipMemcpy(best_step, ctx->best_step.p, n * 4u, hipMemcpyDeviceToHost));
  if (final_score)
    HIP_TRY(ctx, hipMemcpy(final_score, ctx->score.p, n * 4u, hipMemcpyDeviceToHost));
  if (final_beta)
    HIP_TRY(ctx, hipMemcpy(final_beta, ctx->beta.p, n * 4u, hipMemcpyDeviceToHost));
  return BBS_OK;
}

int bbs_get_best(bbs_ctx *ctx, uint64_t *am, uint64_t *dm, float *beta,
                 float *score) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (!ctx->begun || ctx->steps_done == 0u)
    return fail(ctx, BBS_ERR_STATE, "no steps have run");
  const uint32_t R = ctx->params.n_chains;
  std::vector<float> s(R), b(R);
  std::vector<uint32_t> st(R);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMemcpy(s.data(), ctx->best_score.p, R * 4u, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(st.data(), ctx->best_step.p, R * 4u, hipMemcpyDeviceToHost));
  HIP_TRY(ctx, hipMemcpy(b.data(), ctx->best_beta.p, R * 4u, hipMemcpyDeviceToHost));
  /* Chain-major order (src/Application.hpp:408-456, P = 1): chain 0's first
   * proposal seeds the maximum; a later chain replaces it only with a
   * strictly greater proposal, its first such one. */
  uint32_t best = 0u;
  for (uint32_t c = 1; c < R; c++)
    if (st[c] != 0xffffffffu && s[c] > s[best]) best = c;
  const size_t row_bytes = (size_t)ctx->n_nodes * 2u * sizeof(uint64_t);
  if (am)
    HIP_TRY(ctx, hipMemcpy(am, (const char *)ctx->best_am.p + best * row_bytes,
                           row_bytes, hipMemcpyDeviceToHost));
  if (dm)
    HIP_TRY(ctx, hipMemcpy(dm, (const char *)ctx->best_dm.p + best * row_bytes,
                           row_bytes, hipMemcpyDeviceToHost));
  if (beta) *beta = b[best];
  if (score) *score = s[best];
  return BBS_OK;
}

int bbs_get_chain_states(bbs_ctx *ctx, uint64_t *am, uint64_t *dm) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (!ctx->states_set) return fail(ctx, BBS_ERR_STATE, "no chain states yet");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t rows = (size_t)ctx->params.n_chains * ctx->n_nodes * 2u * sizeof(uint64_t);
  if (am) HIP_TRY(ctx, hipMemcpy(am, ctx->am.p, rows, hipMemcpyDeviceToHost));
  if (dm) HIP_TRY(ctx, hipMemcpy(dm, ctx->dm.p, rows, hipMemcpyDeviceToHost));
  return BBS_OK;
}

int bbs_get_trace(bbs_ctx *ctx, bbs_step_trace *out) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (!out) return fail(ctx, BBS_ERR_ARG, "out is NULL");
  if (!ctx->ran) return fail(ctx, BBS_ERR_STATE, "run not finished");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMemcpy(out, ctx->trace.p,
                         (size_t)ctx->n_traced * ctx->params.n_steps *
                             sizeof(bbs_step_trace),
                         hipMemcpyDeviceToHost));
  return BBS_OK;
}

int bbs_get_proposal_trace(bbs_ctx *ctx, uint64_t *am, uint64_t *dm) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (!ctx->ran) return fail(ctx, BBS_ERR_STATE, "run not finished");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t bytes = (size_t)ctx->n_prop_traced * ctx->params.n_steps * ctx->n_nodes * 2u *
                       sizeof(uint64_t);
  if (am && bytes) HIP_TRY(ctx, hipMemcpy(am, ctx->prop_am.p, bytes, hipMemcpyDeviceToHost));
  if (dm && bytes) HIP_TRY(ctx, hipMemcpy(dm, ctx->prop_dm.p, bytes, hipMemcpyDeviceToHost));
  return BBS_OK;
}

/* Fill a ScoreArgs for this context; betas == nullptr means beta_mean. */
static void score_args(bbs_ctx *ctx, bbs::ScoreArgs &s, uint32_t flags) {
  s.cells = (const uint64_t *)ctx->cells.p;
  s.n_cells = ctx->n_cells;
  s.n_nodes = ctx->n_nodes;
  s.with_prior = (flags & BBS_SCORE_NO_PRIOR) ? 0u : 1u;
  s.L00 = ctx->L00;
  s.L10 = ctx->L10;
  const float b = ctx->params.beta_mean;
  s.L01 = bbs_logf(b);
  s.L11 = bbs_logf_of_double(1.0 - (double)b);
  s.prior = ctx->prior;
  s.prior_value = ctx->prior.lg_ab + ctx->prior.am1 * bbs_logf(b) +
                  ctx->prior.bm1 * bbs_logf(1.0f - b) - ctx->prior.lg_a -
                  ctx->prior.lg_b;
}

int bbs_score_trees(bbs_ctx *ctx, const uint64_t *dm_rows, uint32_t n_trees,
                    const float *betas, uint32_t flags, float *ll,
                    uint8_t *counts) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (n_trees == 0u) return BBS_OK;
  if (!dm_rows || !ll) return fail(ctx, BBS_ERR_ARG, "dm_rows and ll must not be NULL");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t rows = (size_t)n_trees * ctx->n_nodes * 2u * sizeof(uint64_t);
  DevBuf d_rows, d_betas, d_ll, d_counts;
  int rc = dev_alloc(ctx, d_rows, rows);
  if (!rc && betas) rc = dev_alloc(ctx, d_betas, (size_t)n_trees * 4u);
  if (!rc) rc = dev_alloc(ctx, d_ll, (size_t)n_trees * 4u);
  if (!rc && counts) rc = dev_alloc(ctx, d_counts, (size_t)n_trees * ctx->n_cells * 4u);
  if (rc) {
    d_rows.release();
    d_betas.release();
    d_ll.release();
    d_counts.release();
    return rc;
  }
  bbs::ScoreArgs s{};
  score_args(ctx, s, flags);
  s.dm = (const uint64_t *)d_rows.p;
  s.betas = betas ? (const float *)d_betas.p : nullptr;
  s.ll = (float *)d_ll.p;
  s.counts = counts ? (uint8_t *)d_counts.p : nullptr;
  s.n_trees = n_trees;
  hipError_t e = hipMemcpy(d_rows.p, dm_rows, rows, hipMemcpyHostToDevice);
  if (e == hipSuccess && betas)
    e = hipMemcpy(d_betas.p, betas, (size_t)n_trees * 4u, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = bbs::launch_score(s, counts != nullptr, !(flags & BBS_SCORE_VALU_LANES), ctx->stream,
                          (flags & BBS_SCORE_CELL_LANES) ? nullptr
                                                         : (const uint32_t *)ctx->cells32.p);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpy(ll, d_ll.p, (size_t)n_trees * 4u, hipMemcpyDeviceToHost);
  if (e == hipSuccess && counts)
    e = hipMemcpy(counts, d_counts.p, (size_t)n_trees * ctx->n_cells * 4u,
                  hipMemcpyDeviceToHost);
  d_rows.release();
  d_betas.release();
  d_ll.release();
  d_counts.release();
  if (e != hipSuccess)
    return fail(ctx, BBS_ERR_HIP, "bbs_score_trees: %s", hipGetErrorString(e));
  return BBS_OK;
}

int bbs_score_trees_noalloc(bbs_ctx *ctx, const uint64_t *dm_rows, uint32_t n_trees,
                            const float *betas, uint32_t flags, float *ll,
                            uint8_t *counts) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (n_trees == 0u) return BBS_OK;
  if (!dm_rows || !ll) return fail(ctx, BBS_ERR_ARG, "dm_rows and ll must not be NULL");
  /* staging layout in the scratch array: rows, betas, scores, counts, each
   * rounded up to 256 bytes */
  auto up = [](uint64_t x) { return (x + 255u) & ~(uint64_t)255u; };
  const uint64_t rows = (uint64_t)n_trees * ctx->n_nodes * 16u;
  const uint64_t o_betas = up(rows);
  const uint64_t o_ll = o_betas + (betas ? up(4ull * n_trees) : 0u);
  const uint64_t o_counts = o_ll + up(4ull * n_trees);
  const uint64_t need = o_counts + (counts ? up(4ull * n_trees * ctx->n_cells) : 0u);
  if (need > BBS_NOALLOC_SCRATCH_BYTES)
    return fail(ctx, BBS_ERR_SIZE,
                "bbs_score_trees_noalloc: %u trees need %llu bytes of staging, more than "
                "its fixed %u bytes; score them in smaller calls",
                n_trees, (unsigned long long)need, (unsigned)BBS_NOALLOC_SCRATCH_BYTES);
  if (!ctx->scratch)
    return fail(ctx, BBS_ERR_HIP, "bbs_score_trees_noalloc: no staging address on this device (%s)",
                ctx->scratch_err.c_str());
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  char *base = (char *)ctx->scratch;
  bbs::ScoreArgs s{};
  score_args(ctx, s, flags);
  s.dm = (const uint64_t *)base;
  s.betas = betas ? (const float *)(base + o_betas) : nullptr;
  s.ll = (float *)(base + o_ll);
  s.counts = counts ? (uint8_t *)(base + o_counts) : nullptr;
  s.n_trees = n_trees;
  hipError_t e = hipMemcpyAsync(base, dm_rows, rows, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && betas)
    e = hipMemcpyAsync(base + o_betas, betas, 4ull * n_trees, hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess)
    e = bbs::launch_score(s, counts != nullptr, !(flags & BBS_SCORE_VALU_LANES), ctx->stream,
                          (flags & BBS_SCORE_CELL_LANES) ? nullptr
                                                         : (const uint32_t *)ctx->cells32.p);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ll, base + o_ll, 4ull * n_trees, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && counts)
    e = hipMemcpyAsync(counts, base + o_counts, 4ull * n_trees * ctx->n_cells,
                       hipMemcpyDeviceToHost, ctx->stream);
  const hipError_t es = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = es;
  if (e != hipSuccess)
    return fail(ctx, BBS_ERR_HIP, "bbs_score_trees_noalloc: %s", hipGetErrorString(e));
  return BBS_OK;
}

/* Device-resident bulk scoring of n_trees uniform random trees (drawn on
 * the device from seed), reps timed launches. kernel_ms[reps] receives each
 * launch's HIP-event time on the launch stream; ll_head receives the first
 * n_head scores (for checking against host-generated trees). */
int bbs_bench_score_random(bbs_ctx *ctx, uint64_t seed, uint32_t n_trees,
                           uint32_t reps, uint32_t flags, float *kernel_ms,
                           float *ll_head, uint32_t n_head) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (n_head > n_trees) return fail(ctx, BBS_ERR_ARG, "n_head > n_trees");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  DevBuf d_rows, d_ll;
  int rc = dev_alloc(ctx, d_rows, (size_t)n_trees * ctx->n_nodes * 16u);
  if (!rc) rc = dev_alloc(ctx, d_ll, (size_t)n_trees * 4u);
  if (rc) {
    d_rows.release();
    d_ll.release();
    return rc;
  }
  bbs::TreeGenArgs g{};
  g.dm = (uint64_t *)d_rows.p;
  g.seed = seed;
  g.purpose = BBS_PURPOSE_TREE_SAMPLE;
  g.n_trees = n_trees;
  g.n_nodes = ctx->n_nodes;
  bbs::ScoreArgs s{};
  score_args(ctx, s, flags);
  s.dm = (const uint64_t *)d_rows.p;
  s.ll = (float *)d_ll.p;
  s.n_trees = n_trees;
  hipError_t e = bbs::launch_random_trees(g, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  for (uint32_t r = 0; r < reps && e == hipSuccess; r++) {
    e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e == hipSuccess)
      e = bbs::launch_score(s, false, !(flags & BBS_SCORE_VALU_LANES), ctx->stream,
                            (flags & BBS_SCORE_CELL_LANES) ? nullptr
                                                           : (const uint32_t *)ctx->cells32.p);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (kernel_ms) kernel_ms[r] = ms;
  }
  if (e == hipSuccess && n_head && ll_head)
    e = hipMemcpy(ll_head, d_ll.p, (size_t)n_head * 4u, hipMemcpyDeviceToHost);
  d_rows.release();
  d_ll.release();
  if (e != hipSuccess)
    return fail(ctx, BBS_ERR_HIP, "bbs_bench_score_random: %s", hipGetErrorString(e));
  return BBS_OK;
}

/* Device-generated random trees copied to the host (dm rows; am nullable). */
int bbs_random_trees_device(bbs_ctx *ctx, uint64_t seed, uint32_t purpose,
                            uint32_t n_trees, uint64_t *am, uint64_t *dm) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const size_t bytes = (size_t)n_trees * ctx->n_nodes * 16u;
  DevBuf d_am, d_dm;
  int rc = dev_alloc(ctx, d_dm, bytes);
  if (!rc && am) rc = dev_alloc(ctx, d_am, bytes);
  if (rc) {
    d_am.release();
    d_dm.release();
    return rc;
  }
  bbs::TreeGenArgs g{};
  g.am = am ? (uint64_t *)d_am.p : nullptr;
  g.dm = (uint64_t *)d_dm.p;
  g.seed = seed;
  g.purpose = purpose;
  g.n_trees = n_trees;
  g.n_nodes = ctx->n_nodes;
  hipError_t e = bbs::launch_random_trees(g, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && dm) e = hipMemcpy(dm, d_dm.p, bytes, hipMemcpyDeviceToHost);
  if (e == hipSuccess && am) e = hipMemcpy(am, d_am.p, bytes, hipMemcpyDeviceToHost);
  d_am.release();
  d_dm.release();
  if (e != hipSuccess)
    return fail(ctx, BBS_ERR_HIP, "bbs_random_trees_device: %s", hipGetErrorString(e));
  return BBS_OK;
}

int bbs_measure_valu_peak(bbs_ctx *ctx, double *int32_ops_per_s,
                          double *valu_insts_per_s) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipDeviceProp_t prop;
  HIP_TRY(ctx, hipGetDeviceProperties(&prop, ctx->device));
  const uint32_t blocks = (uint32_t)prop.multiProcessorCount * 16u;
  const uint32_t iters = 4096u;
  DevBuf din, dout;
  int rc = dev_alloc(ctx, din, 2048u * 4u);
  if (!rc) rc = dev_alloc(ctx, dout, (size_t)blocks * 256u * 4u);
  if (rc) {
    din.release();
    dout.release();
    return rc;
  }
  std::vector<uint32_t> h(2048);
  for (uint32_t i = 0; i < 2048u; i++) h[i] = 0x9E3779B9u * (i + 1u);
  hipError_t e = hipMemcpy(din.p, h.data(), 2048u * 4u, hipMemcpyHostToDevice);
  float best = 1e30f;
  for (int r = 0; r < 6 && e == hipSuccess; r++) {
    e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e == hipSuccess)
      e = bbs::launch_valu_peak((const uint32_t *)din.p, (uint32_t *)dout.p, blocks,
                                iters, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (r > 0 && ms < best) best = ms; /* first launch warms up */
  }
  din.release();
  dout.release();
  if (e != hipSuccess)
    return fail(ctx, BBS_ERR_HIP, "bbs_measure_valu_peak: %s", hipGetErrorString(e));
  const double lanes = (double)blocks * 256.0;
  const double ops = lanes * iters * 16.0; /* 8 chains x (and + popcount) */
  if (int32_ops_per_s) *int32_ops_per_s = ops / (best * 1e-3);
  if (valu_insts_per_s) *valu_insts_per_s = ops / 64.0 / (best * 1e-3);
  return BBS_OK;
}

int bbs_eval_libm(bbs_ctx *ctx, const float *x, uint32_t n, float *exp_out,
                  float *log_out, float *log1m_out) {
  if (!ctx) return fail(nullptr, BBS_ERR_ARG, "ctx is NULL");
  if (!x || !exp_out || !log_out || !log1m_out)
    return fail(ctx, BBS_ERR_ARG, "NULL argument");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  DevBuf dx, de, dl, dm;
  const size_t bytes = (size_t)n * 4u;
  int rc = dev_alloc(ctx, dx, bytes);
  if (!rc) rc = dev_alloc(ctx, de, bytes);
  if (!rc) rc = dev_alloc(ctx, dl, bytes);
  if (!rc) rc = dev_alloc(ctx, dm, bytes);
  hipError_t e = hipSuccess;
  if (!rc) {
    e = hipMemcpy(dx.p, x, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = bbs::launch_libm((const float *)dx.p, n, (float *)de.p, (float *)dl.p,
                           (float *)dm.p, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(exp_out, de.p, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(log_out, dl.p, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(log1m_out, dm.p, bytes, hipMemcpyDeviceToHost);
  }
  dx.release();
  de.release();
  dl.release();
  dm.release();
  if (rc) return rc;
  if (e != hipSuccess) return fail(ctx, BBS_ERR_HIP, "bbs_eval_libm: %s", hipGetErrorString(e));
  return BBS_OK;
}

/* ---------------------------------------------------------------- batch */

} /* extern "C" */

struct bbs_batch {
  int device = -1;
  hipStream_t stream = nullptr;           /* stream 0 */
  hipStream_t side[BBS_MAX_STREAMS - 1] = {};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t side_done[BBS_MAX_STREAMS - 1] = {};
  std::vector<bbs::BatchJob> jobs;
  std::vector<uint32_t> items;
  struct Range {
    uint32_t shape, first, count;
    uint32_t max_floats; /* largest incremental slab of its datasets    */
    double work;
  };
  /* items of one shape, heaviest range first */
  std::vector<Range> ranges;
  bbs_batch_config cfg{};
  uint32_t n_cus = 0;
  uint64_t total_steps = 0;
  DevBuf d_jobs, d_items, cells, am, dm, best_am, best_dm, beta, score,
      best_beta, best_score, best_step;
  hipEvent_t ready = nullptr;       /* stream 0 -> side streams */
  hipGraph_t graph = nullptr;       /* graph mode: the captured run */
  hipGraphExec_t exec = nullptr;
  bool ran = false;
  std::string err;
};

namespace {

int bfail(bbs_batch *b, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b)
    b->err = buf;
  else
    g_create_error = buf;
  return code;
}

#define BHIP_TRY(b, expr)                                                      \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return bfail((b), BBS_ERR_HIP, "%s failed: %s", #expr,                   \
                   hipGetErrorString(e_));                                     \
  } while (0)

int balloc(bbs_batch *b, DevBuf &d, size_t bytes) {
  d.release();
  if (bytes == 0) return BBS_OK;
  hipError_t e = hipMalloc(&d.p, bytes);
  if (e != hipSuccess) {
    d.p = nullptr;
    return bfail(b, e == hipErrorOutOfMemory ? BBS_ERR_BUDGET : BBS_ERR_HIP,
                 "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  d.bytes = bytes;
  return BBS_OK;
}

/* Scorer constants of one job, as the TreeScorer constructor computes them
 * (src/TreeScorer.hpp:92-117), and its move thresholds. */
int fill_job(const bbs_job &in, bbs::BatchJob &J, std::string &why) {
  const bbs_params &p = in.params;
  if (in.n_genes < 2u || in.n_genes > BBS_MAX_GENES || in.n_cells < 1u ||
      in.n_cells > BBS_MAX_CELLS) {
    why = "unsupported size";
    return BBS_ERR_SIZE;
  }
  if (!(p.alpha_mean > 0.f && p.alpha_mean < 1.f) ||
      !(p.beta_mean > 0.f && p.beta_mean < 1.f) || !(p.beta_sd > 0.f) ||
      !(p.beta_jump_scaling_chi > 0.f) || p.prob_beta_change < 0.f ||
      p.prob_prune_n_reattach < 0.f || p.prob_swap_nodes < 0.f ||
      p.prob_swap_subtrees < 0.f ||
      p.prob_beta_change + p.prob_prune_n_reattach + p.prob_swap_nodes +
              p.prob_swap_subtrees ==
          0.f ||
      p.n_chains == 0u || p.n_chains > 65535u || !in.is_mutated || !in.is_known) {
    why = "invalid parameters";
    return BBS_ERR_ARG;
  }
  J = bbs::BatchJob{};
  J.seed = p.seed;
  J.job_tag = in.job_tag;
  J.n_nodes = in.n_genes + 1u;
  J.n_cells = in.n_cells;
  int ns, cs, w;
  bbs::shape_of(in.n_genes, in.n_cells, &ns, &cs, &w, true);
  J.shape = (uint32_t)(ns * 100 + cs * 10 + w);
  J.n_chains = p.n_chains;
  J.step_begin = 0u;
  J.step_end = p.n_steps;
  J.score_init = 1u;
  J.L00 = (float)std::log(1.0 - (double)p.alpha_mean);
  J.L10 = logf(p.alpha_mean);
  J.gamma = p.gamma;
  J.beta_mean = p.beta_mean;
  double a = ((double)(1.0f - p.beta_mean) * std::pow((double)p.beta_mean, 2.0) /
              std::pow((double)p.beta_sd, 2.0)) -
             (double)p.beta_mean;
  float bpa = (float)a, bpb = bpa * ((1.0f / p.beta_mean) - 1.0f);
  J.prior.am1 = bpa - 1.0f;
  J.prior.bm1 = bpb - 1.0f;
  J.prior.lg_ab = lgamma_term(bpa + bpb);
  J.prior.lg_a = lgamma_term(bpa);
  J.prior.lg_b = lgamma_term(bpb);
  float sum = p.prob_beta_change + p.prob_prune_n_reattach + p.prob_swap_nodes +
              p.prob_swap_subtrees;
  float pb = p.prob_beta_change / sum, pp = p.prob_prune_n_reattach / sum,
        ps = p.prob_swap_nodes / sum;
  J.probs.t_beta = pb;
  J.probs.t_prune = pb + pp;
  J.probs.t_swap = pb + pp + ps;
  J.probs.beta_jump_sd = p.beta_sd / p.beta_jump_scaling_chi;
  return BBS_OK;
}

/* Relative device cost of one step of a job (VALU work of the scorer). */
double step_cost(const bbs::BatchJob &J) {
  const int cs = (J.shape / 10) % 10, w = J.shape % 10;
  const double slots = cs == 1 ? 1.0 : cs == 2 ? 2.0 : 1.5;
  return (double)J.n_nodes * slots * (4.0 * w + 12.0) + 600.0;
}

/* Register budget of a layout's incremental scorer: waves per SIMD. */
uint32_t layout_wps(uint32_t shape, uint32_t two_wave_layouts) {
  const uint32_t ns = shape / 100u, cs = shape / 10u % 10u, w = shape % 10u;
  return (two_wave_layouts & BBS_LAYOUT_BIT(ns, cs, w)) ? 2u : 1u;
}

/* LDS floats of a job's incremental slab. */
uint32_t job_floats(const bbs::BatchJob &J) {
  return bbs::batch_wave_floats(J.shape, J.n_nodes, J.n_cells);
}

/* Incremental slab of a layout (floats per wave), or 0 where the table
 * would cost concurrency: a range of `count` items runs on as many waves at
 * once with the slabs (LDS-bound, slabs per CU) as without (one per SIMD),
 * as the plain kernel's automatic rule decides. count = UINT32_MAX asks for
 * a saturated device (the server). */
uint32_t range_slab(uint32_t f, uint32_t count, uint32_t n_cus) {
  const uint64_t per_cu = bbs::kMaxGroupLds / (4ull * f);
  const uint64_t cus = n_cus ? n_cus : 1u;
  const uint64_t with_table = std::min<uint64_t>(count, cus * per_cu);
  const uint64_t without = std::min<uint64_t>(count, cus * 4u);
  return per_cu > 0u && with_table >= without ? f : 0u;
}

/* Launch shape of a range: each wave's incremental LDS slab (0: full
 * rescoring) and the waves per workgroup, halved from the configured value
 * until the workgroup's slabs fit in the LDS one workgroup may hold. */
void range_launch(uint32_t max_floats, uint32_t wpg_cfg, bool full, uint32_t count,
                  uint32_t n_cus, uint32_t &wave_floats, uint32_t &wpg) {
  wpg = wpg_cfg;
  wave_floats = full ? 0u : range_slab(max_floats, count, n_cus);
  while (wave_floats && wpg > 1u && 4ull * wave_floats * wpg > bbs::kMaxGroupLds) wpg /= 2u;
  if (4ull * wave_floats * wpg > bbs::kMaxGroupLds) wave_floats = 0u;
}

/* Chain waves per CU of a launch of `count` items of a layout whose largest
 * slab is max_floats, in the launch shape range_launch picks (0 if the
 * occupancy cannot be queried). */
uint32_t launch_waves_per_cu(uint32_t shape, uint32_t max_floats, uint32_t count,
                             uint32_t wpg_cfg, bool full, uint32_t n_cus,
                             uint32_t two_wave_layouts) {
  uint32_t f = 0u, wpg = 1u, w = 0u;
  range_launch(max_floats, wpg_cfg, full, count, n_cus, f, wpg);
  if (bbs::batch_items_occupancy(shape, wpg, f, layout_wps(shape, two_wave_layouts), &w) !=
      hipSuccess) {
    (void)hipGetLastError();
    return 0u;
  }
  return w;
}

/* Splits the jobs of one layout (`items` chains in all) into the class whose
 * own slabs let a CU hold the most chain waves and the rest: small[k] = 1 for
 * the first. Returns false (no split) when every job is in the first class,
 * when it has fewer chains than the device has CUs, or with full scoring. */
bool slab_classes(uint32_t shape, const std::vector<const bbs::BatchJob *> &js, uint32_t items,
                  uint32_t wpg_cfg, bool full, uint32_t n_cus, uint32_t two_wave_layouts,
                  std::vector<char> &small) {
  small.assign(js.size(), 1);
  if (full || js.size() < 2u) return false;
  std::vector<uint32_t> occ(js.size());
  uint32_t top = 0u;
  for (size_t k = 0; k < js.size(); k++) {
    occ[k] = launch_waves_per_cu(shape, job_floats(*js[k]), items, wpg_cfg, false, n_cus,
                                 two_wave_layouts);
    top = std::max(top, occ[k]);
  }
  uint32_t small_items = 0u;
  bool any_rest = false;
  for (size_t k = 0; k < js.size(); k++) {
    small[k] = occ[k] == top ? 1 : 0;
    if (small[k])
      small_items += js[k]->n_chains;
    else
      any_rest = true;
  }
  if (!any_rest || small_items < n_cus) {
    small.assign(js.size(), 1);
    return false;
  }
  return true;
}

/* Items grouped by register layout, one launch (range) per layout and at
 * most two per layout: when some datasets' own slabs let a CU hold more
 * chain waves than the layout's largest slab does, those datasets (at least
 * one item per CU of them) get a range of their own, with a smaller slab.
 * Inside a range the largest chain comes first (a greedy longest-
 * processing-time schedule for the hardware dispatcher) unless order == 1
 * (job order). Ranges are sorted heaviest first. */
void order_items(bbs_batch *b) {
  b->items.clear();
  b->ranges.clear();
  std::vector<uint32_t> order(b->jobs.size());
  for (uint32_t j = 0; j < order.size(); j++) order[j] = j;
  auto work = [&](uint32_t x) {
    const auto &A = b->jobs[x];
    return step_cost(A) * (A.step_end - A.step_begin);
  };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    if (b->jobs[x].shape != b->jobs[y].shape) return b->jobs[x].shape < b->jobs[y].shape;
    return b->cfg.order == 0u ? work(x) > work(y) : false;
  });
  auto emit = [&](const std::vector<uint32_t> &js) {
    if (js.empty()) return;
    bbs_batch::Range r{b->jobs[js[0]].shape, (uint32_t)b->items.size(), 0u, 0u, 0.0};
    for (uint32_t j : js) {
      r.max_floats = std::max(r.max_floats, job_floats(b->jobs[j]));
      for (uint32_t c = 0; c < b->jobs[j].n_chains; c++) b->items.push_back((j << 16) | c);
      r.work += work(j) * b->jobs[j].n_chains;
    }
    r.count = (uint32_t)b->items.size() - r.first;
    if (r.count) b->ranges.push_back(r);
  };
  for (uint32_t idx = 0; idx < order.size();) {
    const uint32_t shape = b->jobs[order[idx]].shape;
    std::vector<uint32_t> group;
    uint32_t items = 0u;
    for (; idx < order.size() && b->jobs[order[idx]].shape == shape; idx++) {
      group.push_back(order[idx]);
      items += b->jobs[order[idx]].n_chains;
    }
    std::vector<const bbs::BatchJob *> js;
    for (uint32_t j : group) js.push_back(&b->jobs[j]);
    std::vector<char> cls;
    slab_classes(shape, js, items, b->cfg.waves_per_group, b->cfg.scoring == BBS_SCORING_FULL,
                 b->n_cus, b->cfg.two_wave_layouts, cls);
    std::vector<uint32_t> small, rest;
    for (size_t k = 0; k < group.size(); k++) (cls[k] ? small : rest).push_back(group[k]);
    emit(small);
    emit(rest);
  }
  std::stable_sort(b->ranges.begin(), b->ranges.end(),
                   [](const bbs_batch::Range &x, const bbs_batch::Range &y) {
                     return x.work > y.work;
                   });
}

bbs::BatchArgs batch_args(bbs_batch *b) {
  bbs::BatchArgs a{};
  a.jobs = (const bbs::BatchJob *)b->d_jobs.p;
  a.items = (const uint32_t *)b->d_items.p;
  a.n_items = (uint32_t)b->items.size();
  a.cells = (const uint64_t *)b->cells.p;
  a.am = (uint64_t *)b->am.p;
  a.dm = (uint64_t *)b->dm.p;
  a.best_am = (uint64_t *)b->best_am.p;
  a.best_dm = (uint64_t *)b->best_dm.p;
  a.beta = (float *)b->beta.p;
  a.score = (float *)b->score.p;
  a.best_beta = (float *)b->best_beta.p;
  a.best_score = (float *)b->best_score.p;
  a.best_step = (uint32_t *)b->best_step.p;
  return a;
}

} // namespace

extern "C" {

void bbs_batch_config_default(bbs_batch_config *cfg) {
  if (!cfg) return;
  cfg->waves_per_group = 1u;
  cfg->n_streams = 16u;
  cfg->order = 0u;
  cfg->persistent = 0u;
  cfg->scoring = BBS_SCORING_AUTO;
  cfg->graph = 0u;
  cfg->two_wave_layouts = BBS_TWO_WAVE_DEFAULT;
}

const char *bbs_batch_last_error(const bbs_batch *b) {
  return b ? b->err.c_str() : g_create_error.c_str();
}

void bbs_batch_destroy(bbs_batch *b) {
  if (!b) return;
  if (b->device >= 0) (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  DevBuf *bufs[] = {&b->d_jobs,  &b->d_items,   &b->cells,
                    &b->am,      &b->dm,        &b->best_am,   &b->best_dm,
                    &b->beta,    &b->score,     &b->best_beta, &b->best_score,
                    &b->best_step};
  for (DevBuf *d : bufs) d->release();
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->ready) (void)hipEventDestroy(b->ready);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  for (int i = 0; i < BBS_MAX_STREAMS - 1; i++) {
    if (b->side[i]) (void)hipStreamSynchronize(b->side[i]);
    if (b->side_done[i]) (void)hipEventDestroy(b->side_done[i]);
    if (b->side[i]) (void)hipStreamDestroy(b->side[i]);
  }
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int bbs_batch_create(const bbs_job *jobs, uint32_t n_jobs, const bbs_batch_config *cfg,
                     bbs_batch **out) {
  if (!out) return bfail(nullptr, BBS_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!jobs || n_jobs == 0u) return bfail(nullptr, BBS_ERR_ARG, "no jobs");
  if (n_jobs > 65535u) return bfail(nullptr, BBS_ERR_SIZE, "at most 65535 jobs per batch");
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0)
    return bfail(nullptr, BBS_ERR_NODEV, "no HIP device available");
  int dev = -1;
  for (int d = 0; d < n_dev && dev < 0; d++)
    if (is_gfx950(d)) dev = d;
  if (dev < 0) return bfail(nullptr, BBS_ERR_NODEV, "no gfx950 device found");

  bbs_batch *b = new bbs_batch();
  b->device = dev;
  if (cfg)
    b->cfg = *cfg;
  else
    bbs_batch_config_default(&b->cfg);
  if (b->cfg.waves_per_group < 1u || b->cfg.waves_per_group > 4u ||
      b->cfg.n_streams < 1u || b->cfg.n_streams > (uint32_t)BBS_MAX_STREAMS ||
      (b->cfg.scoring != BBS_SCORING_AUTO && b->cfg.scoring != BBS_SCORING_FULL) ||
      b->cfg.graph > 1u) {
    bbs_batch_destroy(b);
    return bfail(nullptr, BBS_ERR_ARG, "invalid batch configuration");
  }
  b->jobs.resize(n_jobs);
  uint64_t cells_total = 0, chains_total = 0, rows_total = 0;
  std::vector<uint64_t> cells;
  for (uint32_t j = 0; j < n_jobs; j++) {
    std::string why;
    int rc = fill_job(jobs[j], b->jobs[j], why);
    if (rc) {
      bbs_batch_destroy(b);
      return bfail(nullptr, rc, "job %u: %s", j, why.c_str());
    }
    bbs::BatchJob &J = b->jobs[j];
    J.cell_offset = (uint32_t)cells_total;
    J.chain_offset = (uint32_t)chains_total;
    J.row_offset = rows_total;
    const uint32_t G = jobs[j].n_genes;
    uint64_t gmask[2] = {G >= 64u ? ~0ull : ((1ull << G) - 1ull),
                         G <= 64u ? 0ull : ((1ull << (G - 64u)) - 1ull)};
    for (uint32_t c = 0; c < J.n_cells; c++) {
      uint64_t q[4];
      for (int wd = 0; wd < 2; wd++) {
        uint64_t K = jobs[j].is_known[2u * c + wd] & gmask[wd];
        uint64_t M = jobs[j].is_mutated[2u * c + wd];
        q[wd] = K & M;
        q[2 + wd] = K & ~M;
      }
      cells.insert(cells.end(), q, q + 4);
    }
    cells_total += J.n_cells;
    chains_total += J.n_chains;
    rows_total += (uint64_t)J.n_chains * J.n_nodes;
    b->total_steps += (uint64_t)J.n_chains * (J.step_end - J.step_begin);
  }
  hipDeviceProp_t prop;
  hipError_t e = hipSetDevice(dev);
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ready, hipEventDisableTiming);
  for (uint32_t i = 0; i + 1u < b->cfg.n_streams && e == hipSuccess; i++) {
    e = hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->side_done[i], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    bbs_batch_destroy(b);
    return bfail(nullptr, BBS_ERR_HIP, "HIP setup failed: %s", hipGetErrorString(e));
  }
  b->n_cus = (uint32_t)prop.multiProcessorCount;
  order_items(b); /* needs the device current (occupancy queries) */
  const size_t rows_b = rows_total * 16u, sc_b = chains_total * 4u;
  int rc;
  if ((rc = balloc(b, b->d_jobs, b->jobs.size() * sizeof(bbs::BatchJob))) ||
      (rc = balloc(b, b->d_items, b->items.size() * 4u)) ||
      (rc = balloc(b, b->cells, cells.size() * 8u)) || (rc = balloc(b, b->am, rows_b)) ||
      (rc = balloc(b, b->dm, rows_b)) || (rc = balloc(b, b->best_am, rows_b)) ||
      (rc = balloc(b, b->best_dm, rows_b)) || (rc = balloc(b, b->beta, sc_b)) ||
      (rc = balloc(b, b->score, sc_b)) || (rc = balloc(b, b->best_beta, sc_b)) ||
      (rc = balloc(b, b->best_score, sc_b)) || (rc = balloc(b, b->best_step, sc_b))) {
    std::string m = b->err;
    bbs_batch_destroy(b);
    return bfail(nullptr, rc, "%s", m.c_str());
  }
  e = hipMemcpy(b->d_jobs.p, b->jobs.data(), b->jobs.size() * sizeof(bbs::BatchJob),
                hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_items.p, b->items.data(), b->items.size() * 4u,
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->cells.p, cells.data(), cells.size() * 8u, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    bbs_batch_destroy(b);
    return bfail(nullptr, BBS_ERR_HIP, "upload failed: %s", hipGetErrorString(e));
  }
  *out = b;
  return BBS_OK;
}

/* Everything one run puts on the streams: initial trees and cleared
 * maxima on stream 0, then the layout ranges, heaviest first, spread over
 * the streams, joined back into stream 0. Kernels only (no memset, no host
 * copies), so it can be captured in a hipGraph. */
static hipError_t enqueue_run(bbs_batch *b) {
  const bbs::BatchArgs a = batch_args(b);
  const uint32_t n_chains =
      b->jobs.empty() ? 0u : b->jobs.back().chain_offset + b->jobs.back().n_chains;
  hipError_t e = hipSuccess;
  for (size_t k = 0; k < b->jobs.size() && e == hipSuccess; k++) {
    const bbs::BatchJob &J = b->jobs[k];
    bbs::TreeGenArgs g{};
    g.am = a.am + 2ull * J.row_offset;
    g.dm = a.dm + 2ull * J.row_offset;
    g.beta = a.beta + J.chain_offset;
    g.beta_value = J.beta_mean;
    g.seed = J.seed;
    g.job = J.job_tag;
    g.chain_offset = 0u;
    g.purpose = BBS_PURPOSE_INIT_TREE;
    g.n_trees = J.n_chains;
    g.n_nodes = J.n_nodes;
    e = bbs::launch_random_trees(g, b->stream);
  }
  if (e == hipSuccess) e = bbs::launch_fill_u32(a.best_step, 0xffffffffu, n_chains, b->stream);
  if (e == hipSuccess)
    e = bbs::launch_fill_u32((uint32_t *)a.best_score, 0xff800000u, n_chains, b->stream);
  if (e == hipSuccess)
    e = bbs::launch_fill_u32((uint32_t *)a.best_beta, 0xff800000u, n_chains, b->stream);
  const uint32_t ns = b->cfg.n_streams;
  if (e == hipSuccess) e = hipEventRecord(b->ready, b->stream);
  for (uint32_t i = 0; i + 1u < ns && e == hipSuccess; i++)
    e = hipStreamWaitEvent(b->side[i], b->ready, 0);
  for (size_t r = 0; r < b->ranges.size() && e == hipSuccess; r++) {
    const auto &R = b->ranges[r];
    hipStream_t st = (r % ns) == 0u ? b->stream : b->side[(r % ns) - 1u];
    uint32_t f = 0u, wpg = 1u;
    range_launch(R.max_floats, b->cfg.waves_per_group, b->cfg.scoring == BBS_SCORING_FULL,
                 R.count, b->n_cus, f, wpg);
    e = bbs::launch_batch_items(a, R.shape, R.first, R.count, wpg, f,
                                layout_wps(R.shape, b->cfg.two_wave_layouts), st);
  }
  for (uint32_t i = 0; i + 1u < ns && e == hipSuccess; i++) {
    e = hipEventRecord(b->side_done[i], b->side[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->side_done[i], 0);
  }
  return e;
}

int bbs_batch_run(bbs_batch *b, float *makespan_ms) {
  if (!b) return bfail(nullptr, BBS_ERR_ARG, "batch is NULL");
  if (b->cfg.persistent)
    return bfail(b, BBS_ERR_STATE,
                 "the persistent kernel form was retired (its first GPU test hung, no "
                 "cause was found); per-layout launches, optionally replayed as a "
                 "hipGraph, are the batched form");
  BHIP_TRY(b, hipSetDevice(b->device));
  hipError_t e = hipSuccess;
  if (b->cfg.graph && !b->exec) {
    e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const hipError_t e1 = enqueue_run(b);
      hipGraph_t g = nullptr;
      const hipError_t e2 = hipStreamEndCapture(b->stream, &g);
      e = e1 != hipSuccess ? e1 : e2;
      if (e == hipSuccess) {
        b->graph = g;
        e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
    }
    if (e != hipSuccess)
      return bfail(b, BBS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
  }
  e = hipEventRecord(b->ev0, b->stream);
  if (e == hipSuccess) e = b->exec ? hipGraphLaunch(b->exec, b->stream) : enqueue_run(b);
  if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
  if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
  if (e != hipSuccess) return bfail(b, BBS_ERR_HIP, "batch run failed: %s", hipGetErrorString(e));
  float ms = 0.f;
  BHIP_TRY(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  if (makespan_ms) *makespan_ms = ms;
  b->ran = true;
  return BBS_OK;
}

uint64_t bbs_batch_total_steps(const bbs_batch *b) { return b ? b->total_steps : 0u; }

int bbs_batch_get_chain_results(bbs_batch *b, uint32_t job, float *best_score,
                                uint32_t *best_step, float *final_score,
                                float *final_beta) {
  if (!b) return bfail(nullptr, BBS_ERR_ARG, "batch is NULL");
  if (job >= b->jobs.size()) return bfail(b, BBS_ERR_ARG, "job %u out of range", job);
  if (!b->ran) return bfail(b, BBS_ERR_STATE, "batch has not run");
  BHIP_TRY(b, hipSetDevice(b->device));
  const bbs::BatchJob &J = b->jobs[job];
  const size_t off = (size_t)J.chain_offset * 4u, n = (size_t)J.n_chains * 4u;
  if (best_score)
    BHIP_TRY(b, hipMemcpy(best_score, (char *)b->best_score.p + off, n, hipMemcpyDeviceToHost));
  if (best_step)
    BHIP_TRY(b, hipMemcpy(best_step, (char *)b->best_step.p + off, n, hipMemcpyDeviceToHost));
  if (final_score)
    BHIP_TRY(b, hipMemcpy(final_score, (char *)b->score.p + off, n, hipMemcpyDeviceToHost));
  if (final_beta)
    BHIP_TRY(b, hipMemcpy(final_beta, (char *)b->beta.p + off, n, hipMemcpyDeviceToHost));
  return BBS_OK;
}

int bbs_batch_get_best(bbs_batch *b, uint32_t job, uint64_t *am, uint64_t *dm,
                       float *beta, float *score) {
  if (!b) return bfail(nullptr, BBS_ERR_ARG, "batch is NULL");
  if (job >= b->jobs.size()) return bfail(b, BBS_ERR_ARG, "job %u out of range", job);
  const bbs::BatchJob &J = b->jobs[job];
  if (J.step_end == 0u) return bfail(b, BBS_ERR_STATE, "job has no steps");
  std::vector<float> s(J.n_chains), bb(J.n_chains);
  std::vector<uint32_t> st(J.n_chains);
  int rc = bbs_batch_get_chain_results(b, job, s.data(), st.data(), nullptr, nullptr);
  if (rc) return rc;
  BHIP_TRY(b, hipMemcpy(bb.data(), (char *)b->best_beta.p + (size_t)J.chain_offset * 4u,
                        (size_t)J.n_chains * 4u, hipMemcpyDeviceToHost));
  /* chain-major order as in bbs_get_best (src/Application.hpp:456) */
  uint32_t best = 0u;
  for (uint32_t c = 1; c < J.n_chains; c++)
    if (st[c] != 0xffffffffu && s[c] > s[best]) best = c;
  const size_t row_bytes = (size_t)J.n_nodes * 16u;
  const size_t off = (size_t)(J.row_offset + (uint64_t)best * J.n_nodes) * 16u;
  if (am)
    BHIP_TRY(b, hipMemcpy(am, (char *)b->best_am.p + off, row_bytes, hipMemcpyDeviceToHost));
  if (dm)
    BHIP_TRY(b, hipMemcpy(dm, (char *)b->best_dm.p + off, row_bytes, hipMemcpyDeviceToHost));
  if (beta) *beta = bb[best];
  if (score) *score = s[best];
  return BBS_OK;
}

/* Best of two runs of a batch (the first warms up); 1e30 if it fails. */
static int time_batch(const std::vector<bbs_job> &jobs, const bbs_batch_config &c, float &t) {
  bbs_batch *b = nullptr;
  int rc = bbs_batch_create(jobs.data(), (uint32_t)jobs.size(), &c, &b);
  if (rc) return rc;
  float ms = 0.f;
  t = 1e30f;
  for (int r = 0; r < 2 && rc == BBS_OK; r++) {
    rc = bbs_batch_run(b, &ms);
    if (r > 0 && ms < t) t = ms;
  }
  bbs_batch_destroy(b);
  return rc;
}

int bbs_batch_autotune(const bbs_job *jobs, uint32_t n_jobs, uint32_t probe_steps,
                       bbs_batch_config *best, float *best_ms) {
  if (!jobs || !best) return bfail(nullptr, BBS_ERR_ARG, "NULL argument");
  std::vector<bbs_job> probe(jobs, jobs + n_jobs);
  for (auto &j : probe)
    if (j.params.n_steps > probe_steps) j.params.n_steps = probe_steps;
  /* 1. each layout's register budget, its jobs alone on one stream */
  std::map<uint32_t, std::vector<bbs_job>> by_layout;
  for (const auto &j : probe) {
    bbs::BatchJob J;
    std::string why;
    const int rc = fill_job(j, J, why);
    if (rc) return bfail(nullptr, rc, "job rejected: %s", why.c_str());
    by_layout[J.shape].push_back(j);
  }
  bbs_batch_config base;
  bbs_batch_config_default(&base);
  base.n_streams = 1u;
  base.waves_per_group = 1u;
  uint32_t mask = 0u;
  for (const auto &kv : by_layout) {
    const uint32_t shape = kv.first;
    const uint32_t bit = BBS_LAYOUT_BIT(shape / 100u, shape / 10u % 10u, shape % 10u);
    /* probe only where the budgets differ: the occupancy the registers
     * allow (a minimal slab) of the one-wave build against the two-wave
     * build; where they are equal, the one-wave build (no spills) is kept */
    uint32_t occ1 = 0u, occ2 = 0u;
    const bool q1 = bbs::batch_items_occupancy(shape, 1u, 4u, 1u, &occ1) == hipSuccess;
    const bool q2 = bbs::batch_items_occupancy(shape, 1u, 4u, 2u, &occ2) == hipSuccess;
    (void)hipGetLastError();
    if (q1 && q2 && occ2 <= occ1) continue;
    float t1 = 0.f, t2 = 0.f;
    bbs_batch_config c = base;
    c.two_wave_layouts = 0u;
    int rc = time_batch(kv.second, c, t1);
    if (!rc) {
      c.two_wave_layouts = bit;
      rc = time_batch(kv.second, c, t2);
    }
    if (rc) return rc;
    if (t2 < t1) mask |= bit;
  }
  /* 2. waves per workgroup x streams x item order with those budgets (LDS
   * is allocated per workgroup, so one or two waves per group pack CUs
   * best; 16 streams overlap the layouts on as many hardware queues as the
   * process has) */
  const uint32_t wpgs[2] = {1u, 2u};
  const uint32_t streams[2] = {2u, 16u};
  float best_t = 1e30f;
  bool found = false;
  for (uint32_t order = 0; order < 2; order++) {
    for (uint32_t wpg : wpgs) {
      for (uint32_t nst : streams) {
        bbs_batch_config c = base;
        c.waves_per_group = wpg;
        c.n_streams = nst;
        c.order = order;
        c.two_wave_layouts = mask;
        float t = 0.f;
        const int rc = time_batch(probe, c, t);
        if (rc) return rc;
        if (t < best_t) {
          best_t = t;
          *best = c;
          found = true;
        }
      }
    }
  }
  if (!found) return bfail(nullptr, BBS_ERR_STATE, "no configuration ran");
  if (best_ms) *best_ms = best_t;
  return BBS_OK;
}

/* ---------------------------------------------------------------- host */

int bbs_parse_csv(const char *text, uint64_t length, int8_t *matrix,
                  uint32_t *n_genes_out, uint32_t *n_cells_out) {
  if (!text || !matrix || !n_genes_out || !n_cells_out)
    return fail(nullptr, BBS_ERR_ARG, "NULL argument");
  /* pass 1: count and validate (src/main.cpp:52-101) */
  uint32_t n_cells = 0, n_genes = 0, in_line = 0;
  for (uint64_t i = 0; i < length; i++) {
    char c = text[i];
    if (c == '0' || c == '1' || c == '2' || c == '3') {
      in_line++;
    } else if (c == '\n') {
      if (n_cells == 0u)
        n_cells = in_line;
      else if (n_cells != in_line)
        return fail(nullptr, BBS_ERR_ARG, "Inconsistent rows in the input file!");
      n_genes++;
      in_line = 0;
    } else if (c == ' ') {
      continue;
    } else {
      return fail(nullptr, BBS_ERR_ARG, "Illegal character %c in the input file!", c);
    }
  }
  if (in_line == n_cells)
    n_genes++;
  else if (in_line != 0u)
    return fail(nullptr, BBS_ERR_ARG, "Inconsistent rows in the input file!");
  if (n_cells == 0u || n_genes == 0u)
    return fail(nullptr, BBS_ERR_ARG, "Empty input file!");
  if (n_cells > BBS_MAX_CELLS)
    return fail(nullptr, BBS_ERR_SIZE,
                "This build of bbSCITE only supports up %u cells. The input data "
                "contains %u cells.",
                BBS_MAX_CELLS, n_cells);
  if (n_genes > BBS_MAX_GENES)
    return fail(nullptr, BBS_ERR_SIZE,
                "This build of bbSCITE only supports up %u genes. The input data "
                "contains %u genes.",
                BBS_MAX_GENES, n_genes);
  /* pass 2: whitespace-separated integers (src/main.cpp:120-143) */
  uint64_t i = 0;
  for (uint32_t g = 0; g < n_genes; g++) {
    for (uint32_t c = 0; c < n_cells; c++) {
      while (i < length && (text[i] == ' ' || text[i] == '\n')) i++;
      long v = 0;
      uint64_t start = i;
      while (i < length && text[i] >= '0' && text[i] <= '9') {
        v = v * 10 + (text[i] - '0');
        if (v > 1000) v = 1000;
        i++;
      }
      if (i == start || v > 3)
        return fail(nullptr, BBS_ERR_ARG,
                    "The input file contains the invalid entry %ld", v);
      matrix[(size_t)g * n_cells + c] = (int8_t)v;
    }
  }
  *n_genes_out = n_genes;
  *n_cells_out = n_cells;
  return BBS_OK;
}

int bbs_encode_matrix(const int8_t *matrix, uint32_t n_genes, uint32_t n_cells,
                      uint64_t *is_mutated, uint64_t *is_known) {
  if (!matrix || !is_mutated || !is_known)
    return fail(nullptr, BBS_ERR_ARG, "NULL argument");
  if (n_genes > BBS_MAX_GENES || n_cells > BBS_MAX_CELLS)
    return fail(nullptr, BBS_ERR_SIZE, "matrix too large");
  std::memset(is_mutated, 0, sizeof(uint64_t) * 2u * n_cells);
  std::memset(is_known, 0, sizeof(uint64_t) * 2u * n_cells);
  for (uint32_t g = 0; g < n_genes; g++) {
    for (uint32_t c = 0; c < n_cells; c++) {
      const uint64_t bit = 1ull << (g & 63u);
      switch (matrix[(size_t)g * n_cells + c]) {
      case 0:
        is_known[2u * c + (g >> 6)] |= bit;
        break;
      case 1:
      case 2:
        is_known[2u * c + (g >> 6)] |= bit;
        is_mutated[2u * c + (g >> 6)] |= bit;
        break;
      case 3:
        break;
      default:
        return fail(nullptr, BBS_ERR_ARG,
                    "The input file contains the invalid entry %d",
                    (int)matrix[(size_t)g * n_cells + c]);
      }
    }
  }
  return BBS_OK;
}

int bbs_generate_dataset(uint64_t seed, uint32_t n_genes, uint32_t n_cells,
                         float prob_false_positive, float prob_false_negative,
                         float prob_missing, int8_t *matrix,
                         uint32_t *true_parent, uint32_t *attachment) {
  if (!matrix) return fail(nullptr, BBS_ERR_ARG, "matrix is NULL");
  if (n_genes < 2u || n_genes > BBS_MAX_GENES || n_cells < 1u ||
      n_cells > BBS_MAX_CELLS)
    return fail(nullptr, BBS_ERR_SIZE, "unsupported size %u x %u", n_genes, n_cells);
  const uint32_t n = n_genes + 1u;
  uint32_t parent[BBS_MAX_NODES];
  gen_tree_parent(seed, 0u, 0u, BBS_PURPOSE_DATA_TREE, n, parent);
  std::vector<uint64_t> am(2u * n);
  int rc = parent_to_rows(parent, n, am.data(), nullptr);
  if (rc) return rc;
  std::vector<uint32_t> att(n_cells);
  for (uint32_t c = 0; c < n_cells; c++) {
    bbs_u32x4 b = bbs_philox_block(seed, 0u, 0u, c >> 2, BBS_PURPOSE_DATA_CELL);
    att[c] = bbs_uint_below(b.x[c & 3u], n);
  }
  for (uint32_t g = 0; g < n_genes; g++) {
    for (uint32_t c = 0; c < n_cells; c++) {
      const bool truth = rbit(&am[2u * g], att[c]);
      bbs_u32x4 b = bbs_philox_block(seed, 0u, 0u, g * n_cells + c,
                                     BBS_PURPOSE_DATA_NOISE);
      int state;
      if (!truth)
        state = bbs_u01f(b.x[0]) < prob_false_positive ? 1 : 0;
      else
        state = bbs_u01f(b.x[0]) < prob_false_negative ? 0 : 1;
      if (bbs_u01f(b.x[1]) < prob_missing) state = 3;
      matrix[(size_t)g * n_cells + c] = (int8_t)state;
    }
  }
  if (true_parent) std::memcpy(true_parent, parent, n * sizeof(uint32_t));
  if (attachment) std::memcpy(attachment, att.data(), n_cells * sizeof(uint32_t));
  return BBS_OK;
}

int bbs_random_trees(uint64_t seed, uint32_t n_genes, uint32_t n_trees,
                     uint64_t *dm_rows) {
  if (!dm_rows) return fail(nullptr, BBS_ERR_ARG, "dm_rows is NULL");
  if (n_genes < 2u || n_genes > BBS_MAX_GENES)
    return fail(nullptr, BBS_ERR_SIZE, "unsupported gene count %u", n_genes);
  const uint32_t n = n_genes + 1u;
  for (uint32_t t = 0; t < n_trees; t++) {
    uint32_t parent[BBS_MAX_NODES];
    gen_tree_parent(seed, 0u, t, BBS_PURPOSE_TREE_SAMPLE, n, parent);
    int rc = parent_to_rows(parent, n, nullptr, dm_rows + (size_t)t * n * 2u);
    if (rc) return rc;
  }
  return BBS_OK;
}

int bbs_initial_tree(uint64_t seed, uint32_t job, uint32_t chain,
                     uint32_t n_genes, uint64_t *am, uint64_t *dm) {
  if (n_genes < 2u || n_genes > BBS_MAX_GENES)
    return fail(nullptr, BBS_ERR_SIZE, "unsupported gene count %u", n_genes);
  uint32_t parent[BBS_MAX_NODES];
  gen_tree_parent(seed, job, chain, BBS_PURPOSE_INIT_TREE, n_genes + 1u, parent);
  return parent_to_rows(parent, n_genes + 1u, am, dm);
}

int bbs_parent_vector_to_matrix(const uint32_t *parent, uint32_t n_nodes,
                                uint64_t *am, uint64_t *dm) {
  if (!parent) return fail(nullptr, BBS_ERR_ARG, "parent is NULL");
  int rc = parent_to_rows(parent, n_nodes, am, dm);
  if (rc) return fail(nullptr, rc, "not a rooted tree over %u nodes", n_nodes);
  return BBS_OK;
}

int bbs_pruefer_to_parent_vector(const uint32_t *code, uint32_t code_len,
                                 uint32_t *parent) {
  if (!code || !parent) return fail(nullptr, BBS_ERR_ARG, "NULL argument");
  int rc = pruefer_to_parent(code, code_len, parent);
  if (rc) return fail(nullptr, rc, "invalid Pruefer code");
  return BBS_OK;
}

int64_t bbs_tree_to_graphviz(const uint64_t *am, const uint64_t *dm,
                             uint32_t n_nodes, char *buf, uint64_t cap) {
  (void)am;
  if (!dm || n_nodes < 1u || n_nodes > BBS_MAX_NODES) return BBS_ERR_ARG;
  TextOut out{buf, cap, 0};
  out.put("digraph G {\n");
  out.put("node [color=deeppink4, style=filled, fontcolor=white];\n");
  for (uint32_t i = 0; i + 1u < n_nodes; i++)
    out.put(std::to_string(rows_parent(dm, n_nodes, i)) + " -> " +
            std::to_string(i) + ";\n");
  out.put("}\n");
  out.finish();
  return (int64_t)out.len;
}

int64_t bbs_tree_to_newick(const uint64_t *am, const uint64_t *dm,
                           uint32_t n_nodes, char *buf, uint64_t cap) {
  (void)am;
  if (!dm || n_nodes < 1u || n_nodes > BBS_MAX_NODES) return BBS_ERR_ARG;
  std::vector<std::vector<uint32_t>> children(n_nodes);
  for (uint32_t i = 0; i + 1u < n_nodes; i++)
    children[rows_parent(dm, n_nodes, i)].push_back(i);
  TextOut out{buf, cap, 0};
  newick_rec(children, n_nodes - 1u, out, 0u);
  out.put("\n");
  out.finish();
  return (int64_t)out.len;
}

} /* extern "C" */

/* ------------------------------------------------------------------ serving */

namespace {

/* First-fit allocator over [0, cap) elements of one device array. */
struct Arena {
  uint64_t cap = 0;
  std::map<uint64_t, uint64_t> free_; /* offset -> length */
  void init(uint64_t c) {
    cap = c;
    free_.clear();
    if (c) free_[0] = c;
  }
  bool alloc(uint64_t n, uint64_t &off) {
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        off = it->first;
        const uint64_t len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        return true;
      }
    return false;
  }
  void release(uint64_t off, uint64_t n) {
    if (n == 0) return;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

struct SrvJob {
  uint64_t id = 0;
  bbs::BatchJob J{};     /* offsets set at admission */
  uint32_t n_genes = 0, n_steps = 0, steps_done = 0;
  std::vector<uint64_t> cells; /* [n_cells][4] */
  double cost = 0.0;     /* step_cost(J) */
  uint64_t lds = 0;      /* LDS bytes of one chain wave */
  double t_submit = 0.0, t_admit = 0.0;
  bool started = false;
};

/* The resident jobs of one register layout and the slice they run: a lane
 * relaunches as soon as its own slice ends, on its own stream, so layouts
 * never wait for each other. */
struct Lane {
  uint32_t shape = 0;
  hipStream_t stream = nullptr;      /* one of the server's streams       */
  std::vector<SrvJob *> jobs;        /* resident, in admission order      */
  /* the slice in flight (valid while busy) */
  bool busy = false;
  std::vector<SrvJob *> slice_jobs;
  std::vector<bbs::BatchJob> h_jobs; /* host sources of the uploads stay  */
  std::vector<uint32_t> h_items;     /* alive until the slice completes   */
  double slice_work = 0.0;           /* heaviest job's cost x steps       */
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  DevBuf d_jobs, d_items;
};

} // namespace

struct bbs_server {
  int device = -1;
  bbs_server_config cfg{};
  hipStream_t stream[BBS_MAX_STREAMS] = {};
  hipStream_t retire_stream = nullptr; /* result copies of finished jobs only */
  std::vector<Lane *> lanes;
  std::map<uint32_t, double> ms_per_cost; /* layout -> ms per (step x cost unit) */
  uint32_t n_cus = 0, lds_per_cu = 0;
  DevBuf am, dm, best_am, best_dm, beta, score, best_beta, best_score, best_step, cells;
  Arena rows, chains, cellsA;
  std::deque<SrvJob *> pending;
  std::deque<bbs_job_result> done;
  uint64_t next_id = 1;
  uint64_t lds_cap = 0, lds_used = 0;
  uint32_t chain_cap = 0, chains_used = 0, n_resident = 0, n_busy = 0;
  uint64_t steps_total = 0, slices = 0;
  double busy_s = 0.0, busy_since = 0.0;
  std::chrono::steady_clock::time_point t0;
  std::string err;
};

namespace {

int sfail(bbs_server *s, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (s)
    s->err = buf;
  else
    g_create_error = buf;
  return code;
}

#define SHIP_TRY(s, expr)                                                      \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return sfail((s), BBS_ERR_HIP, "%s failed: %s", #expr,                   \
                   hipGetErrorString(e_));                                     \
  } while (0)

double srv_now(const bbs_server *s) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

int srv_alloc(bbs_server *s, DevBuf &d, size_t bytes) {
  d.release();
  if (bytes == 0) return BBS_OK;
  hipError_t e = hipMalloc(&d.p, bytes);
  if (e != hipSuccess) {
    d.p = nullptr;
    return sfail(s, e == hipErrorOutOfMemory ? BBS_ERR_BUDGET : BBS_ERR_HIP,
                 "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  d.bytes = bytes;
  return BBS_OK;
}

/* Grows a staging buffer (job table, item list) to at least `bytes`. */
int srv_reserve(bbs_server *s, DevBuf &d, size_t bytes) {
  if (d.bytes >= bytes) return BBS_OK;
  return srv_alloc(s, d, std::max(bytes, 2 * d.bytes));
}

void srv_free_job(bbs_server *s, SrvJob *j) {
  const bbs::BatchJob &J = j->J;
  s->rows.release(J.row_offset, (uint64_t)J.n_chains * J.n_nodes);
  s->chains.release(J.chain_offset, J.n_chains);
  s->cellsA.release(J.cell_offset, J.n_cells);
  s->lds_used -= (uint64_t)J.n_chains * j->lds;
  s->chains_used -= J.n_chains;
  delete j;
}

/* The lane of a layout, created on first use: streams are dealt out in
 * creation order, and the job and item tables are sized for the most a
 * lane can hold (chain_cap chains), so they never grow under a slice. */
int srv_lane(bbs_server *s, uint32_t shape, Lane *&out) {
  for (Lane *l : s->lanes)
    if (l->shape == shape) {
      out = l;
      return BBS_OK;
    }
  Lane *l = new Lane();
  l->shape = shape;
  l->stream = s->stream[s->lanes.size() % s->cfg.n_streams];
  s->lanes.push_back(l);
  SHIP_TRY(s, hipEventCreate(&l->ev_start));
  SHIP_TRY(s, hipEventCreate(&l->ev_end));
  int rc;
  if ((rc = srv_alloc(s, l->d_jobs, (size_t)s->chain_cap * sizeof(bbs::BatchJob))) ||
      (rc = srv_alloc(s, l->d_items, (size_t)s->chain_cap * 4u)))
    return rc;
  out = l;
  return BBS_OK;
}

/* The pending job that admission considers next: the oldest, or the one of
 * least predicted work (chains x steps x step cost) aged by its time
 * pending: work x exp(-(now - submitted) / aging_s), the oldest of equals.
 * The aged order does not change while jobs wait (the factor exp(-now /
 * aging_s) is common to all), so it is the order of ln(work) + submitted /
 * aging_s. */
size_t srv_next(const bbs_server *s) {
  if (s->cfg.admission != BBS_ADMIT_LEAST_WORK) return 0u;
  const double inv_aging = s->cfg.aging_s > 0.0f ? 1.0 / s->cfg.aging_s : 0.0;
  size_t best = 0u;
  double best_key = 0.0;
  for (size_t i = 0; i < s->pending.size(); i++) {
    const SrvJob *j = s->pending[i];
    const double work = (double)j->J.n_chains * j->n_steps * j->cost;
    const double key = std::log(work) + j->t_submit * inv_aging;
    if (i == 0u || key < best_key) {
      best = i;
      best_key = key;
    }
  }
  return best;
}

/* Admit the next pending job (srv_next) if memory and concurrent-wave
 * capacity allow (an empty server admits any job, so none waits forever). */
int srv_admit(bbs_server *s, bool &admitted) {
  admitted = false;
  const size_t at = srv_next(s);
  SrvJob *j = s->pending[at];
  bbs::BatchJob &J = j->J;
  const bool empty = s->n_resident == 0u;
  if (!empty && (s->chains_used + J.n_chains > s->chain_cap ||
                 s->lds_used + (uint64_t)J.n_chains * j->lds > s->lds_cap))
    return BBS_OK;
  uint64_t r = 0, c = 0, x = 0;
  const uint64_t n_rows = (uint64_t)J.n_chains * J.n_nodes;
  if (!s->rows.alloc(n_rows, r)) return BBS_OK;
  if (!s->chains.alloc(J.n_chains, c)) {
    s->rows.release(r, n_rows);
    return BBS_OK;
  }
  if (!s->cellsA.alloc(J.n_cells, x)) {
    s->rows.release(r, n_rows);
    s->chains.release(c, J.n_chains);
    return BBS_OK;
  }
  J.row_offset = r;
  J.chain_offset = (uint32_t)c;
  J.cell_offset = (uint32_t)x;
  s->lds_used += (uint64_t)J.n_chains * j->lds;
  s->chains_used += J.n_chains;
  Lane *l = nullptr;
  int rc = srv_lane(s, J.shape, l);
  if (rc) return rc;
  /* queued behind the lane's slice in flight, which does not touch it */
  SHIP_TRY(s, hipMemcpyAsync((uint64_t *)s->cells.p + 4ull * x, j->cells.data(),
                             j->cells.size() * 8u, hipMemcpyHostToDevice, l->stream));
  j->t_admit = srv_now(s);
  s->pending.erase(s->pending.begin() + (std::ptrdiff_t)at);
  l->jobs.push_back(j);
  s->n_resident++;
  admitted = true;
  return BBS_OK;
}

/* The finished job's result: the chain-major first maximum of its chains
 * (src/Application.hpp:456), as bbs_get_best. */
int srv_retire(bbs_server *s, SrvJob *j, double t_done) {
  const bbs::BatchJob &J = j->J;
  std::vector<float> sc(J.n_chains), bb(J.n_chains);
  std::vector<uint32_t> st(J.n_chains);
  /* The job's last slice has ended (its lane's end event completed). The
   * copies go on a stream that carries nothing else, so waiting for them
   * never waits for a slice of another lane sharing the lane's stream. */
  hipStream_t st_ = s->retire_stream;
  SHIP_TRY(s, hipMemcpyAsync(sc.data(), (float *)s->best_score.p + J.chain_offset,
                             J.n_chains * 4u, hipMemcpyDeviceToHost, st_));
  SHIP_TRY(s, hipMemcpyAsync(st.data(), (uint32_t *)s->best_step.p + J.chain_offset,
                             J.n_chains * 4u, hipMemcpyDeviceToHost, st_));
  SHIP_TRY(s, hipMemcpyAsync(bb.data(), (float *)s->best_beta.p + J.chain_offset,
                             J.n_chains * 4u, hipMemcpyDeviceToHost, st_));
  SHIP_TRY(s, hipStreamSynchronize(st_));
  uint32_t best = 0u;
  for (uint32_t c = 1; c < J.n_chains; c++)
    if (st[c] != 0xffffffffu && sc[c] > sc[best]) best = c;
  bbs_job_result r;
  std::memset(&r, 0, sizeof r);
  r.id = j->id;
  r.n_genes = j->n_genes;
  r.n_cells = J.n_cells;
  r.n_chains = J.n_chains;
  r.n_steps = j->n_steps;
  r.best_score = sc[best];
  r.best_beta = bb[best];
  r.best_chain = best;
  r.best_step = st[best];
  r.submit_s = j->t_submit;
  r.admit_s = j->t_admit;
  r.done_s = t_done;
  const size_t off = (size_t)(J.row_offset + (uint64_t)best * J.n_nodes) * 16u;
  SHIP_TRY(s, hipMemcpyAsync(r.best_am, (char *)s->best_am.p + off, (size_t)J.n_nodes * 16u,
                             hipMemcpyDeviceToHost, st_));
  SHIP_TRY(s, hipMemcpyAsync(r.best_dm, (char *)s->best_dm.p + off, (size_t)J.n_nodes * 16u,
                             hipMemcpyDeviceToHost, st_));
  SHIP_TRY(s, hipStreamSynchronize(st_));
  s->done.push_back(r);
  s->n_resident--;
  srv_free_job(s, j);
  return BBS_OK;
}

/* Launches the next slice of every resident job of an idle lane: new jobs'
 * initial trees and cleared maxima first, then one items launch. Each job
 * advances about round_ms (from the layout's measured ms per cost unit), or
 * round_cost cost units before the layout has been measured. */
int srv_lane_launch(bbs_server *s, Lane *l) {
  const size_t n = l->jobs.size();
  l->slice_jobs = l->jobs;
  l->h_jobs.resize(n);
  l->h_items.clear();
  l->slice_work = 0.0;
  auto f = s->ms_per_cost.find(l->shape);
  std::vector<uint32_t> order(n);
  for (uint32_t r = 0; r < n; r++) {
    SrvJob *j = l->jobs[r];
    bbs::BatchJob J = j->J;
    const double want = f == s->ms_per_cost.end()
                            ? (double)s->cfg.round_cost / j->cost
                            : (double)s->cfg.round_ms / (f->second * j->cost);
    const uint64_t adv = (uint64_t)std::max(1.0, std::min(want, 4.0e9));
    J.step_begin = j->steps_done;
    J.step_end = (uint32_t)std::min<uint64_t>(j->n_steps, j->steps_done + adv);
    J.score_init = j->started ? 0u : 1u;
    l->h_jobs[r] = J;
    l->slice_work = std::max(l->slice_work, j->cost * (double)(J.step_end - J.step_begin));
    order[r] = r;
    if (!j->started) {
      /* initial trees as bbs_begin draws them, no maxima yet */
      bbs::TreeGenArgs g{};
      g.am = (uint64_t *)s->am.p + 2ull * J.row_offset;
      g.dm = (uint64_t *)s->dm.p + 2ull * J.row_offset;
      g.beta = (float *)s->beta.p + J.chain_offset;
      g.beta_value = J.beta_mean;
      g.seed = J.seed;
      g.job = J.job_tag;
      g.chain_offset = 0u;
      g.purpose = BBS_PURPOSE_INIT_TREE;
      g.n_trees = J.n_chains;
      g.n_nodes = J.n_nodes;
      SHIP_TRY(s, bbs::launch_random_trees(g, l->stream));
      SHIP_TRY(s, bbs::launch_fill_u32((uint32_t *)s->best_step.p + J.chain_offset, 0xffffffffu,
                                       J.n_chains, l->stream));
      SHIP_TRY(s, bbs::launch_fill_u32((uint32_t *)s->best_score.p + J.chain_offset,
                                       0xff800000u, J.n_chains, l->stream));
      SHIP_TRY(s, bbs::launch_fill_u32((uint32_t *)s->best_beta.p + J.chain_offset,
                                       0xff800000u, J.n_chains, l->stream));
    }
  }
  /* two launches when some jobs' own slabs fit more chain waves per CU
   * than the lane's largest slab (slab_classes): those jobs first; inside
   * each class the largest work first */
  uint32_t lane_items = 0u;
  std::vector<const bbs::BatchJob *> js(n);
  for (uint32_t r = 0; r < n; r++) {
    js[r] = &l->h_jobs[r];
    lane_items += l->h_jobs[r].n_chains;
  }
  std::vector<char> cls;
  const bool split = slab_classes(l->shape, js, lane_items, s->cfg.waves_per_group, false,
                                  s->n_cus, s->cfg.two_wave_layouts, cls);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    if (cls[x] != cls[y]) return cls[x] > cls[y];
    const double wx = l->jobs[x]->cost * (double)(l->h_jobs[x].step_end - l->h_jobs[x].step_begin);
    const double wy = l->jobs[y]->cost * (double)(l->h_jobs[y].step_end - l->h_jobs[y].step_begin);
    return wx > wy;
  });
  uint32_t n_first = 0u;
  uint32_t floats_first = 0u, floats_second = 0u;
  for (uint32_t r : order) {
    for (uint32_t c = 0; c < l->h_jobs[r].n_chains; c++) l->h_items.push_back((r << 16) | c);
    if (cls[r]) {
      n_first += l->h_jobs[r].n_chains;
      floats_first = std::max(floats_first, job_floats(l->h_jobs[r]));
    } else {
      floats_second = std::max(floats_second, job_floats(l->h_jobs[r]));
    }
  }
  SHIP_TRY(s, hipMemcpyAsync(l->d_jobs.p, l->h_jobs.data(), n * sizeof(bbs::BatchJob),
                             hipMemcpyHostToDevice, l->stream));
  SHIP_TRY(s, hipMemcpyAsync(l->d_items.p, l->h_items.data(), l->h_items.size() * 4u,
                             hipMemcpyHostToDevice, l->stream));
  bbs::BatchArgs a{};
  a.jobs = (const bbs::BatchJob *)l->d_jobs.p;
  a.items = (const uint32_t *)l->d_items.p;
  a.n_items = (uint32_t)l->h_items.size();
  a.cells = (const uint64_t *)s->cells.p;
  a.am = (uint64_t *)s->am.p;
  a.dm = (uint64_t *)s->dm.p;
  a.best_am = (uint64_t *)s->best_am.p;
  a.best_dm = (uint64_t *)s->best_dm.p;
  a.beta = (float *)s->beta.p;
  a.score = (float *)s->score.p;
  a.best_beta = (float *)s->best_beta.p;
  a.best_score = (float *)s->best_score.p;
  a.best_step = (uint32_t *)s->best_step.p;
  /* table or full scoring for each launch from its own item count: a lane
   * with few chains (a long job in the tail) takes the faster table even
   * where a saturated device would not; either gives identical chains. The
   * two launches of a split slice follow each other on the lane's stream
   * (other lanes keep the device busy meanwhile). */
  const uint32_t wps = layout_wps(l->shape, s->cfg.two_wave_layouts);
  auto launch = [&](uint32_t first, uint32_t count, uint32_t max_floats) -> int {
    uint32_t fl = 0u, wpg_r = 1u;
    range_launch(max_floats, s->cfg.waves_per_group, false, count, s->n_cus, fl, wpg_r);
    SHIP_TRY(s, bbs::launch_batch_items(a, l->shape, first, count, wpg_r, fl, wps, l->stream));
    return BBS_OK;
  };
  SHIP_TRY(s, hipEventRecord(l->ev_start, l->stream));
  int rc = split ? launch(0u, n_first, floats_first) : launch(0u, a.n_items, floats_first);
  if (!rc && split) rc = launch(n_first, a.n_items - n_first, floats_second);
  if (rc) return rc;
  SHIP_TRY(s, hipEventRecord(l->ev_end, l->stream));
  l->busy = true;
  if (s->n_busy++ == 0u) s->busy_since = srv_now(s);
  s->slices++;
  return BBS_OK;
}

/* A lane's slice has ended: its measured time updates the layout's ms per
 * cost unit, its jobs advance, and the finished ones retire. */
int srv_lane_complete(bbs_server *s, Lane *l) {
  const double t_end = srv_now(s);
  float ms = 0.f;
  SHIP_TRY(s, hipEventElapsedTime(&ms, l->ev_start, l->ev_end));
  if (l->slice_work > 0.0 && ms > 0.f) {
    const double sample = ms / l->slice_work;
    auto f = s->ms_per_cost.find(l->shape);
    if (f == s->ms_per_cost.end())
      s->ms_per_cost[l->shape] = sample;
    else
      f->second = 0.5 * f->second + 0.5 * sample;
  }
  l->busy = false;
  if (--s->n_busy == 0u) s->busy_s += t_end - s->busy_since;
  std::vector<SrvJob *> still;
  for (size_t r = 0; r < l->slice_jobs.size(); r++) {
    SrvJob *j = l->slice_jobs[r];
    const bbs::BatchJob &J = l->h_jobs[r];
    s->steps_total += (uint64_t)J.n_chains * (J.step_end - J.step_begin);
    j->steps_done = J.step_end;
    j->started = true;
  }
  for (SrvJob *j : l->jobs) {
    if (j->started && j->steps_done >= j->n_steps) {
      const int rc = srv_retire(s, j, t_end);
      if (rc) return rc;
    } else {
      still.push_back(j);
    }
  }
  l->jobs.swap(still);
  l->slice_jobs.clear();
  return BBS_OK;
}

} // namespace

extern "C" {

void bbs_server_config_default(bbs_server_config *cfg) {
  if (!cfg) return;
  cfg->hbm_budget_bytes = 1ull << 30;
  cfg->round_cost = 4000000u;
  cfg->round_ms = 20.0f;
  cfg->max_resident_chains = 0u;
  cfg->waves_per_group = 1u;
  cfg->n_streams = 4u;
  cfg->two_wave_layouts = BBS_TWO_WAVE_DEFAULT;
  cfg->admission = BBS_ADMIT_ARRIVAL;
  cfg->aging_s = 0.0f;
}

const char *bbs_server_last_error(const bbs_server *s) {
  return s ? s->err.c_str() : g_create_error.c_str();
}

void bbs_server_destroy(bbs_server *s) {
  if (!s) return;
  if (s->device >= 0) (void)hipSetDevice(s->device);
  for (int i = 0; i < BBS_MAX_STREAMS; i++)
    if (s->stream[i]) (void)hipStreamSynchronize(s->stream[i]);
  for (SrvJob *j : s->pending) delete j;
  for (Lane *l : s->lanes) {
    for (SrvJob *j : l->jobs) delete j;
    l->d_jobs.release();
    l->d_items.release();
    if (l->ev_start) (void)hipEventDestroy(l->ev_start);
    if (l->ev_end) (void)hipEventDestroy(l->ev_end);
    delete l;
  }
  DevBuf *bufs[] = {&s->am,   &s->dm,    &s->best_am,   &s->best_dm,    &s->beta,
                    &s->score, &s->best_beta, &s->best_score, &s->best_step, &s->cells};
  for (DevBuf *d : bufs) d->release();
  for (int i = 0; i < BBS_MAX_STREAMS; i++)
    if (s->stream[i]) (void)hipStreamDestroy(s->stream[i]);
  if (s->retire_stream) (void)hipStreamDestroy(s->retire_stream);
  delete s;
}

int bbs_server_create(const bbs_server_config *cfg, bbs_server **out) {
  if (!out) return sfail(nullptr, BBS_ERR_ARG, "out is NULL");
  *out = nullptr;
  bbs_server_config c;
  if (cfg)
    c = *cfg;
  else
    bbs_server_config_default(&c);
  if (c.waves_per_group < 1u || c.waves_per_group > 4u || c.n_streams < 1u ||
      c.n_streams > (uint32_t)BBS_MAX_STREAMS || c.round_cost == 0u || !(c.round_ms > 0.0f) ||
      c.hbm_budget_bytes < (1ull << 20) ||
      (c.admission != BBS_ADMIT_ARRIVAL && c.admission != BBS_ADMIT_LEAST_WORK) ||
      !(c.aging_s >= 0.0f))
    return sfail(nullptr, BBS_ERR_ARG, "invalid server configuration");
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0)
    return sfail(nullptr, BBS_ERR_NODEV, "no HIP device available");
  int dev = -1;
  for (int d = 0; d < n_dev && dev < 0; d++)
    if (is_gfx950(d)) dev = d;
  if (dev < 0) return sfail(nullptr, BBS_ERR_NODEV, "no gfx950 device found");
  bbs_server *s = new bbs_server();
  s->cfg = c;
  s->device = dev;
  s->t0 = std::chrono::steady_clock::now();
  hipDeviceProp_t prop;
  hipError_t e = hipSetDevice(dev);
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
  for (uint32_t i = 0; i < c.n_streams && e == hipSuccess; i++)
    e = hipStreamCreateWithFlags(&s->stream[i], hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->retire_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    bbs_server_destroy(s);
    return sfail(nullptr, BBS_ERR_HIP, "HIP setup failed: %s", hipGetErrorString(e));
  }
  s->n_cus = (uint32_t)prop.multiProcessorCount;
  s->lds_per_cu = (uint32_t)prop.maxSharedMemoryPerMultiProcessor;
  /* as many chain waves as run at once: LDS per CU, and two waves per SIMD
   * (a chain wave uses up to 256 VGPRs) */
  /* default: 16 times the chain waves that run at once (two per SIMD), so
   * that long jobs are admitted early and share the device rather than
   * wait behind the queue */
  s->chain_cap = c.max_resident_chains ? c.max_resident_chains : s->n_cus * 8u * 16u;
  /* a resident set above what runs at once is queued by the hardware: the
   * LDS budget grows by the same factor */
  const uint64_t at_once = (uint64_t)s->n_cus * 8u;
  const uint64_t over = std::max<uint64_t>(1u, (s->chain_cap + at_once - 1u) / at_once);
  s->lds_cap = (uint64_t)s->n_cus * s->lds_per_cu * over;
  /* device memory: 64 B per tree row (am, dm and their bests), 20 B per
   * chain, 32 B per cell; rows take 90 % of the budget */
  const uint64_t B = c.hbm_budget_bytes;
  const uint64_t n_rows = B / 10u * 9u / 64u, n_chains = B / 50u / 20u, n_cells = B / 50u / 32u;
  s->rows.init(n_rows);
  s->chains.init(n_chains);
  s->cellsA.init(n_cells);
  int rc;
  if ((rc = srv_alloc(s, s->am, n_rows * 16u)) || (rc = srv_alloc(s, s->dm, n_rows * 16u)) ||
      (rc = srv_alloc(s, s->best_am, n_rows * 16u)) ||
      (rc = srv_alloc(s, s->best_dm, n_rows * 16u)) ||
      (rc = srv_alloc(s, s->beta, n_chains * 4u)) ||
      (rc = srv_alloc(s, s->score, n_chains * 4u)) ||
      (rc = srv_alloc(s, s->best_beta, n_chains * 4u)) ||
      (rc = srv_alloc(s, s->best_score, n_chains * 4u)) ||
      (rc = srv_alloc(s, s->best_step, n_chains * 4u)) ||
      (rc = srv_alloc(s, s->cells, n_cells * 32u))) {
    std::string why = s->err;
    bbs_server_destroy(s);
    return sfail(nullptr, rc, "%s", why.c_str());
  }
  *out = s;
  return BBS_OK;
}

int bbs_server_submit(bbs_server *s, const bbs_job *job, uint64_t *id) {
  if (!s || !job) return sfail(s, BBS_ERR_ARG, "NULL argument");
  SrvJob *j = new SrvJob();
  std::string why;
  int rc = fill_job(*job, j->J, why);
  if (rc) {
    delete j;
    return sfail(s, rc, "job rejected: %s", why.c_str());
  }
  if (job->params.n_steps == 0u || job->params.n_chains == 0u) {
    delete j;
    return sfail(s, BBS_ERR_ARG, "job rejected: no chains or no steps");
  }
  const uint32_t G = job->n_genes;
  const uint64_t gmask[2] = {G >= 64u ? ~0ull : ((1ull << G) - 1ull),
                             G <= 64u ? 0ull : ((1ull << (G - 64u)) - 1ull)};
  j->cells.resize(4u * (size_t)job->n_cells);
  for (uint32_t c = 0; c < job->n_cells; c++)
    for (int wd = 0; wd < 2; wd++) {
      const uint64_t K = job->is_known[2u * c + wd] & gmask[wd];
      const uint64_t M = job->is_mutated[2u * c + wd];
      j->cells[4u * c + wd] = K & M;
      j->cells[4u * c + 2u + wd] = K & ~M;
    }
  if ((uint64_t)j->J.n_chains * j->J.n_nodes > s->rows.cap || j->J.n_chains > s->chains.cap ||
      j->J.n_cells > s->cellsA.cap) {
    delete j;
    return sfail(s, BBS_ERR_BUDGET, "job rejected: larger than the server's device memory budget");
  }
  /* a lane's job and item tables hold chain_cap entries; an empty server
   * admits any job, so a larger one would overrun them */
  if (j->J.n_chains > s->chain_cap) {
    const uint32_t n = j->J.n_chains;
    delete j;
    return sfail(s, BBS_ERR_BUDGET,
                 "job rejected: %u chains exceed the server's resident set of %u chains", n,
                 s->chain_cap);
  }
  j->n_genes = G;
  j->n_steps = job->params.n_steps;
  j->cost = step_cost(j->J);
  const uint32_t f = range_slab(job_floats(j->J), UINT32_MAX, s->n_cus);
  j->lds = f ? 4ull * f : 2560ull;
  j->t_submit = srv_now(s);
  j->id = s->next_id++;
  s->pending.push_back(j);
  if (id) *id = j->id;
  return BBS_OK;
}

int bbs_server_step(bbs_server *s, uint32_t *n_resident, uint32_t *n_pending,
                    uint32_t *n_done_ready) {
  if (!s) return sfail(nullptr, BBS_ERR_ARG, "server is NULL");
  SHIP_TRY(s, hipSetDevice(s->device));
  bool progressed = false;
  for (int pass = 0; pass < 2; pass++) {
    /* lanes whose slice has ended */
    for (Lane *l : s->lanes) {
      if (!l->busy) continue;
      const hipError_t q = hipEventQuery(l->ev_end);
      if (q == hipErrorNotReady) continue;
      SHIP_TRY(s, q);
      const int rc = srv_lane_complete(s, l);
      if (rc) return rc;
      progressed = true;
    }
    /* admission in the configured order */
    while (!s->pending.empty()) {
      bool ok = false;
      const int rc = srv_admit(s, ok);
      if (rc) return rc;
      if (!ok) break;
    }
    /* every idle lane with resident jobs runs its next slice */
    for (Lane *l : s->lanes) {
      if (l->busy || l->jobs.empty()) continue;
      const int rc = srv_lane_launch(s, l);
      if (rc) return rc;
      progressed = true;
    }
    if (progressed || s->n_busy == 0u) break;
    /* nothing to do until a slice ends: poll every busy lane (HIP has no
     * wait-for-any), so whichever ends first is relaunched at once */
    for (bool ended = false; !ended;) {
      for (Lane *l : s->lanes) {
        if (!l->busy) continue;
        const hipError_t q = hipEventQuery(l->ev_end);
        if (q == hipErrorNotReady) continue;
        SHIP_TRY(s, q);
        ended = true;
        break;
      }
      if (!ended) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
  if (n_resident) *n_resident = s->n_resident;
  if (n_pending) *n_pending = (uint32_t)s->pending.size();
  if (n_done_ready) *n_done_ready = (uint32_t)s->done.size();
  return BBS_OK;
}

int bbs_server_poll(bbs_server *s, bbs_job_result *out, uint32_t max, uint32_t *n_out) {
  if (!s || !n_out || (max && !out)) return sfail(s, BBS_ERR_ARG, "NULL argument");
  uint32_t n = 0;
  while (n < max && !s->done.empty()) {
    out[n++] = s->done.front();
    s->done.pop_front();
  }
  *n_out = n;
  return BBS_OK;
}

double bbs_server_now(const bbs_server *s) { return s ? srv_now(s) : 0.0; }

int bbs_server_stats(const bbs_server *s, uint64_t *chain_steps, double *busy_s,
                     uint64_t *slices) {
  if (!s) return sfail(nullptr, BBS_ERR_ARG, "server is NULL");
  if (chain_steps) *chain_steps = s->steps_total;
  if (busy_s) *busy_s = s->busy_s;
  if (slices) *slices = s->slices;
  return BBS_OK;
}

} /* extern "C" */

/* ------------------------------------------------------------ quality */

namespace {

struct NewickParser {
  const char *p;
  std::vector<std::pair<uint64_t, uint64_t>> edges; /* (parent, child) labels */
  std::vector<uint64_t> labels;
  bool ok = true;
  void ws() {
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
  }
  bool number(uint64_t &v) {
    ws();
    if (*p < '0' || *p > '9') return false;
    v = 0;
    while (*p >= '0' && *p <= '9') {
      v = v * 10u + (uint64_t)(*p - '0');
      if (v > 1000000000ull) return false;
      p++;
    }
    return true;
  }
  /* tree : NUMBER | "(" tree ("," tree)* ")" NUMBER */
  bool tree(uint64_t &root, int depth) {
    if (depth > 256) return false;
    ws();
    if (*p == '(') {
      p++;
      std::vector<uint64_t> kids;
      for (;;) {
        uint64_t k;
        if (!tree(k, depth + 1)) return false;
        kids.push_back(k);
        ws();
        if (*p == ',') {
          p++;
          continue;
        }
        if (*p == ')') {
          p++;
          break;
        }
        return false;
      }
      if (!number(root)) return false;
      for (uint64_t k : kids) edges.emplace_back(root, k);
    } else if (!number(root)) {
      return false;
    }
    labels.push_back(root);
    return true;
  }
};

/* regularized incomplete beta I_x(a, b) by its continued fraction
 * (Abramowitz & Stegun 26.5.8), evaluated with the modified Lentz method */
double inc_beta(double a, double b, double x) {
  if (x <= 0.0) return 0.0;
  if (x >= 1.0) return 1.0;
  if (x > (a + 1.0) / (a + b + 2.0)) return 1.0 - inc_beta(b, a, 1.0 - x);
  const double tiny = 1e-300;
  const double lbeta = std::lgamma(a) + std::lgamma(b) - std::lgamma(a + b);
  const double front = std::exp(a * std::log(x) + b * std::log1p(-x) - lbeta) / a;
  double f = 1.0, c = 1.0, d = 0.0;
  for (int i = 0; i <= 400; i++) {
    const int m = i / 2;
    double num;
    if (i == 0)
      num = 1.0;
    else if (i % 2 == 0)
      num = (m * (b - m) * x) / ((a + 2.0 * m - 1.0) * (a + 2.0 * m));
    else
      num = -((a + m) * (a + b + m) * x) / ((a + 2.0 * m) * (a + 2.0 * m + 1.0));
    d = 1.0 + num * d;
    if (std::fabs(d) < tiny) d = tiny;
    d = 1.0 / d;
    c = 1.0 + num / c;
    if (std::fabs(c) < tiny) c = tiny;
    const double cd = c * d;
    f *= cd;
    if (std::fabs(1.0 - cd) < 1e-15) break;
  }
  return front * (f - 1.0);
}

/* Student t distribution with nu degrees of freedom: P(T <= t), with the
 * tail beyond |t| computed directly (no 1 - p cancellation for small p) */
double student_cdf(double t, double nu) {
  const double x = nu / (nu + t * t);
  const double tail = 0.5 * inc_beta(0.5 * nu, 0.5, x);
  return t > 0.0 ? 1.0 - tail : tail;
}
double student_sf(double t, double nu) { return student_cdf(-t, nu); }

} // namespace

extern "C" {

int bbs_newick_to_parent_vector(const char *text, uint32_t *parent, uint32_t *n_nodes) {
  if (!text || !parent || !n_nodes) return fail(nullptr, BBS_ERR_ARG, "NULL argument");
  NewickParser P{text, {}, {}};
  uint64_t root = 0;
  if (!P.tree(root, 0)) return fail(nullptr, BBS_ERR_ARG, "malformed Newick text");
  P.ws();
  if (*P.p != '\0') return fail(nullptr, BBS_ERR_ARG, "trailing characters after the tree");
  const size_t n = P.labels.size();
  if (n < 1u || n > BBS_MAX_NODES) return fail(nullptr, BBS_ERR_SIZE, "%zu nodes", n);
  std::vector<uint64_t> sorted = P.labels;
  std::sort(sorted.begin(), sorted.end());
  for (size_t i = 1; i < n; i++)
    if (sorted[i] == sorted[i - 1])
      return fail(nullptr, BBS_ERR_ARG, "label %llu appears twice",
                  (unsigned long long)sorted[i]);
  auto idx = [&](uint64_t label) {
    return (uint32_t)(std::lower_bound(sorted.begin(), sorted.end(), label) - sorted.begin());
  };
  if (idx(root) != n - 1u)
    return fail(nullptr, BBS_ERR_ARG, "the root must carry the largest label");
  for (size_t i = 0; i < n; i++) parent[i] = (uint32_t)(n - 1u);
  for (const auto &e : P.edges) parent[idx(e.second)] = idx(e.first);
  *n_nodes = (uint32_t)n;
  return BBS_OK;
}

int bbs_score_attachments(const uint64_t *dm, uint32_t n_nodes, const uint64_t *is_mutated,
                          const uint64_t *is_known, uint32_t n_cells, double alpha,
                          double beta, uint32_t *attachment, double *log_likelihood) {
  if (!dm || !is_mutated || !is_known || !log_likelihood)
    return fail(nullptr, BBS_ERR_ARG, "NULL argument");
  if (n_nodes < 2u || n_nodes > BBS_MAX_NODES || n_cells < 1u || n_cells > BBS_MAX_CELLS)
    return fail(nullptr, BBS_ERR_SIZE, "unsupported size");
  const uint32_t G = n_nodes - 1u;
  const uint64_t gmask[2] = {G >= 64u ? ~0ull : ((1ull << G) - 1ull),
                             G <= 64u ? 0ull : ((1ull << (G - 64u)) - 1ull)};
  const double l00 = std::log(1.0 - alpha), l10 = std::log(alpha), l01 = std::log(beta),
               l11 = std::log(1.0 - beta);
  double total = 0.0;
  for (uint32_t c = 0; c < n_cells; c++) {
    double best = 0.0;
    uint32_t arg = 0u;
    for (uint32_t v = 0; v < n_nodes; v++) {
      uint32_t n[2][2] = {{0u, 0u}, {0u, 0u}}; /* [observed][true] */
      for (int w = 0; w < 2; w++) {
        const uint64_t K = is_known[2u * c + w] & gmask[w], M = is_mutated[2u * c + w];
        const uint64_t A = dm[2u * v + w] & gmask[w];
        n[1][1] += (uint32_t)__builtin_popcountll(K & M & A);
        n[1][0] += (uint32_t)__builtin_popcountll(K & M & ~A);
        n[0][1] += (uint32_t)__builtin_popcountll(K & ~M & A);
        n[0][0] += (uint32_t)__builtin_popcountll(K & ~M & ~A);
      }
      const double ll = n[0][0] * l00 + n[1][0] * l10 + n[0][1] * l01 + n[1][1] * l11;
      if (v == 0u || ll > best) {
        best = ll;
        arg = v;
      }
    }
    if (attachment) attachment[c] = arg;
    total += best;
  }
  *log_likelihood = total;
  return BBS_OK;
}

int bbs_tost(const double *diff, uint32_t n, double delta, double *mean, double *p_lower,
             double *p_upper) {
  if (!diff || !p_lower || !p_upper) return fail(nullptr, BBS_ERR_ARG, "NULL argument");
  if (n < 2u || !(delta > 0.0)) return fail(nullptr, BBS_ERR_ARG, "need n >= 2 and delta > 0");
  double m = 0.0;
  for (uint32_t i = 0; i < n; i++) m += diff[i];
  m /= n;
  double ss = 0.0;
  for (uint32_t i = 0; i < n; i++) ss += (diff[i] - m) * (diff[i] - m);
  const double sd = std::sqrt(ss / (n - 1u));
  if (mean) *mean = m;
  if (sd == 0.0) {
    *p_lower = m > -delta ? 0.0 : 1.0;
    *p_upper = m < delta ? 0.0 : 1.0;
    return BBS_OK;
  }
  const double se = sd / std::sqrt((double)n), nu = (double)n - 1.0;
  *p_lower = student_sf((m + delta) / se, nu);        /* H1: mean > -delta */
  *p_upper = student_cdf((m - delta) / se, nu);       /* H1: mean < delta  */
  return BBS_OK;
}

} /* extern "C" */
