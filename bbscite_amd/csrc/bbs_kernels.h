/*
 * Launch interface between the C-ABI layer (bbscite_capi.cpp) and the HIP
 * kernels (bbs_kernels.hip). Plain structs passed by value as kernel
 * arguments; all pointers are device memory.
 */
#ifndef BBS_KERNELS_H
#define BBS_KERNELS_H

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/bbscite.h"
#include "bbs_device.h"

namespace bbs {

/* Arguments of one chain-kernel launch: chains [0, n_chains) of one job,
 * steps [step_begin, step_end). Chain c uses Philox chain index
 * chain_offset + c; states and bests are read at the start and written
 * back at the end, so consecutive launches continue the same chains. */
struct ChainArgs {
  const uint64_t *cells;  /* [n_cells][4] K&M lo/hi, K&~M lo/hi      */
  uint64_t *am, *dm;      /* [n_chains][n_nodes][2] current state     */
  float *beta, *score;    /* [n_chains]                               */
  uint64_t *best_am, *best_dm;
  float *best_beta, *best_score;
  uint32_t *best_step;    /* UINT32_MAX: no record yet                */
  bbs_step_trace *trace;  /* [n_traced][trace_stride]                 */
  uint32_t n_traced, trace_stride;
  uint64_t seed;
  uint32_t job, chain_offset;
  uint32_t n_chains, step_begin, step_end;
  uint32_t n_cells, n_nodes;
  uint32_t score_init;    /* score the initial trees on the device    */
  float L00, L10, gamma;
  PriorArgs prior;
  bbs_move_probs probs;
  uint64_t *prop_am, *prop_dm; /* [n_prop_traced][trace_stride][n_nodes][2] */
  uint32_t n_prop_traced;
};

struct ScoreArgs {
  const uint64_t *cells;
  const uint64_t *dm;     /* [n_trees][n_nodes][2] */
  const float *betas;     /* [n_trees] or null     */
  float *ll;              /* [n_trees]             */
  uint8_t *counts;        /* [n_trees][n_cells][4] or null */
  uint32_t n_trees, n_cells, n_nodes, with_prior;
  float L00, L10, L01, L11, prior_value; /* L01/L11/prior for null betas */
  PriorArgs prior;
};

struct TreeGenArgs {
  uint64_t *am;           /* nullable */
  uint64_t *dm;
  float *beta;            /* nullable */
  float beta_value;
  uint64_t seed;
  uint32_t job, chain_offset, purpose;
  uint32_t n_trees, n_nodes;
};

/* One dataset of a batched run (device copy). Offsets index the pooled
 * buffers of BatchArgs: cells in cells (4 words each), chains in the
 * per-chain scalars, rows in the row buffers (2 words each). */
struct BatchJob {
  uint64_t seed;
  uint64_t row_offset;
  uint32_t job_tag, n_nodes, n_cells, shape;
  uint32_t cell_offset, chain_offset, n_chains;
  uint32_t step_begin, step_end, score_init;
  float L00, L10, gamma, beta_mean;
  PriorArgs prior;
  bbs_move_probs probs;
};

/* items: (job << 16 | chain) in scheduling order. */
struct BatchArgs {
  const BatchJob *jobs;
  const uint32_t *items;
  uint32_t n_items;
  const uint64_t *cells;
  uint64_t *am, *dm, *best_am, *best_dm;
  float *beta, *score, *best_beta, *best_score;
  uint32_t *best_step;
};

/* Register layout for a dataset: NS node slots, CS cell slots (3 = paired
 * half slot for 65..96 cells, used when paired is true), W gene words. */
void shape_of(uint32_t n_genes, uint32_t n_cells, int *ns, int *cs, int *w,
              bool paired);
/* table_bytes: 0 for the full scorer every step, else the dynamic LDS of an
 * incremental chain wave (chain_table_bytes, or chain_half_table_bytes with
 * half: a table of the first 64 cells, the others rescored every step). */
hipError_t launch_chain(const ChainArgs &a, bool trace, uint32_t table_bytes, bool half,
                        hipStream_t stream);
/* A chain split over `split` (2 or 4) waves by cells, for launches with
 * fewer chains than SIMDs: LDS bytes of its workgroup, 0 if the split does
 * not fit (or would leave a wave without cells). */
uint32_t chain_split_bytes(uint32_t n_nodes, uint32_t n_cells, uint32_t split);
hipError_t launch_chain_split(const ChainArgs &a, bool trace, uint32_t split,
                              hipStream_t stream);
/* A chain run speculatively on k (2, 4 or 8) waves that propose consecutive
 * steps at once: LDS bytes of its workgroup, 0 if it does not fit. */
uint32_t chain_spec_bytes(uint32_t n_nodes, uint32_t n_cells, uint32_t k);
hipError_t launch_chain_spec(const ChainArgs &a, bool trace, uint32_t k, hipStream_t stream);
/* LDS one workgroup may allocate on gfx950: the whole 160 KB of a CU. */
constexpr uint32_t kMaxGroupLds = 163840u;

/* LDS bytes of an incremental chain wave (pair-score table, tree rows,
 * sums, node penalties), 0 if it exceeds the LDS of one workgroup. */
uint32_t chain_table_bytes(uint32_t n_nodes, uint32_t n_cells);
/* LDS bytes of a chain wave with the table of cells 0..63 only, for
 * datasets of over 96 cells (two full cell slots); 0 otherwise. */
uint32_t chain_half_table_bytes(uint32_t n_nodes, uint32_t n_cells);
/* Static LDS of an incremental chain wave besides chain_table_bytes. */
uint32_t chain_static_lds();
/* cells32: [128][10] words for the lane-per-tree kernel (used when counts
 * are not requested), or null to use the lane-per-cell kernel. */
hipError_t launch_score(const ScoreArgs &a, bool counts, bool mfma, hipStream_t stream,
                        const uint32_t *cells32);
hipError_t launch_random_trees(const TreeGenArgs &a, hipStream_t stream);
/* p[0..n) = value with a kernel (capturable in a hipGraph, unlike a memset) */
hipError_t launch_fill_u32(uint32_t *p, uint32_t value, uint32_t n, hipStream_t stream);
/* items [first, first + count) all have the given shape code; wave_floats
 * 0 for the full scorer, else the LDS floats of each wave's incremental
 * slab (the largest batch_wave_floats of the range's datasets);
 * waves_per_simd 1 or 2 picks the incremental scorer's register budget */
hipError_t launch_batch_items(const BatchArgs &a, uint32_t shape, uint32_t first,
                              uint32_t count, uint32_t waves_per_group, uint32_t wave_floats,
                              uint32_t waves_per_simd, hipStream_t stream);
/* Chain waves of one launch that a CU holds at once (registers and LDS of
 * the kernel instance that launch_batch_items would pick). */
hipError_t batch_items_occupancy(uint32_t shape, uint32_t waves_per_group, uint32_t wave_floats,
                                 uint32_t waves_per_simd, uint32_t *waves_per_cu);
/* LDS floats of an incremental chain wave for datasets of up to max_nodes
 * nodes and max_cells cells in this layout */
uint32_t batch_wave_floats(uint32_t shape, uint32_t max_nodes, uint32_t max_cells);

/* Device address of this device's staging array for bbs_score_trees_noalloc
 * (BBS_NOALLOC_SCRATCH_BYTES), looked up once per device with the device
 * current and kept; later calls return the kept address. */
hipError_t noalloc_scratch(int device, void **ptr);

hipError_t launch_valu_peak(const uint32_t *in, uint32_t *out, uint32_t blocks,
                            uint32_t iters, hipStream_t stream);
hipError_t launch_libm(const float *x, uint32_t n, float *e, float *l, float *l1m,
                       hipStream_t stream);

} // namespace bbs

#endif /* BBS_KERNELS_H */
