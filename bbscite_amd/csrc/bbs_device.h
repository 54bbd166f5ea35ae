/*
 * Wave-level building blocks of the chain search on gfx950 (wave64).
 *
 * Register layout of one tree inside one wavefront:
 *   lane l, slot k holds the rows of node x = l + 64 k (k < NS);
 *   a row is NS 64-bit words (NS = 1 for up to 64 nodes, 2 for up to 128).
 *   am[k][*] = descendants of x (the reference's `ancestor` row)
 *   dm[k][*] = ancestors of x   (the reference's `descendant` row)
 * Rows of nodes >= n_nodes are zero and stay zero under every move.
 *
 * Register layout of the mutation data for scoring:
 *   lane l, slot j holds cell c = l + 64 j (j < CS) as W 32-bit words of
 *   K&M (known and mutated) and K&~M (known, not mutated), plus the two
 *   popcounts |K&M| and |K&~M| as floats.
 */
#ifndef BBS_DEVICE_H
#define BBS_DEVICE_H

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "bbs_libm.h"

namespace bbs {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t rfl(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}

__device__ __forceinline__ float rfl_f(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

__device__ __forceinline__ uint32_t readlane32(uint32_t x, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)lane);
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, uint32_t lane) {
  uint32_t lo = readlane32((uint32_t)x, lane);
  uint32_t hi = readlane32((uint32_t)(x >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t ballot(bool p) {
  return __ballot(p);
}

/* Order this wave's LDS writes before its later LDS reads (LDS executes a
 * wave's operations in order; this keeps the compiler from moving them). The
 * LDS scratch of score_tree belongs to one wave, so no workgroup barrier is
 * needed and waves of one workgroup may run different chains. */
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

/* Number of set bits of m in lanes below this lane. */
__device__ __forceinline__ uint32_t mbcnt64(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* Bits of word k that are node indices < n. */
__device__ __forceinline__ uint64_t valid_mask(uint32_t k, uint32_t n) {
  uint32_t lo = 64u * k;
  if (n >= lo + 64u) return ~0ull;
  if (n <= lo) return 0ull;
  return (1ull << (n - lo)) - 1ull;
}

/* Word i / 64 of a row, selected without dynamic register indexing. */
template <int NS>
__device__ __forceinline__ uint64_t row_word(const uint64_t (&row)[NS], uint32_t i) {
  return (NS == 1 || i < 64u) ? row[0] : row[NS - 1];
}

template <int NS>
__device__ __forceinline__ bool row_bit(const uint64_t (&row)[NS], uint32_t i) {
  return (row_word<NS>(row, i) >> (i & 63u)) & 1u;
}

/* Uniform copy of the row of node v (held by lane v % 64, slot v / 64). */
template <int NS>
__device__ __forceinline__ void fetch_row(const uint64_t (&rows)[NS][NS],
                                          uint32_t v, uint64_t (&out)[NS]) {
  uint32_t lane = v & 63u;
  if (NS == 1 || v < 64u) {
#pragma unroll
    for (int wd = 0; wd < NS; wd++) out[wd] = readlane64(rows[0][wd], lane);
  } else {
#pragma unroll
    for (int wd = 0; wd < NS; wd++) out[wd] = readlane64(rows[NS - 1][wd], lane);
  }
}

/* Index of the idx-th set bit (ascending) of a uniform row restricted to
 * nodes < n; 0 if there is none (src/MutationTree.hpp:840-878). */
template <int NS>
__device__ __forceinline__ uint32_t select_bit(const uint64_t (&row)[NS],
                                               uint32_t idx, uint32_t n) {
  const uint64_t lanebit = 1ull << lane_id();
#pragma unroll
  for (int k = 0; k < NS; k++) {
    uint64_t wd = row[k] & valid_mask((uint32_t)k, n);
    uint32_t pc = (uint32_t)__popcll(wd);
    if (idx < pc) {
      bool hit = (wd & lanebit) != 0 && mbcnt64(wd) == idx;
      uint64_t m = ballot(hit);
      return 64u * (uint32_t)k + (uint32_t)__ffsll((long long)m) - 1u;
    }
    idx -= pc;
  }
  return 0u;
}

/* Parent of a non-root node v: the node p whose ancestor row equals v's
 * without bit v; the last such node as src/MutationTree.hpp:585-596 scans. */
template <int NS>
__device__ __forceinline__ uint32_t parent_of(const uint64_t (&dm)[NS][NS],
                                              const uint64_t (&dm_v)[NS],
                                              uint32_t v, uint32_t n) {
  uint64_t t[NS];
#pragma unroll
  for (int wd = 0; wd < NS; wd++)
    t[wd] = dm_v[wd] & ~((v >> 6) == (uint32_t)wd ? 1ull << (v & 63u) : 0ull);
  uint32_t result = 0u;
#pragma unroll
  for (int k = NS - 1; k >= 0; k--) {
    bool eq = lane_id() + 64u * (uint32_t)k < n;
#pragma unroll
    for (int wd = 0; wd < NS; wd++) eq = eq && dm[k][wd] == t[wd];
    uint64_t m = ballot(eq);
    if (m != 0ull) {
      result = 64u * (uint32_t)k + 63u - (uint32_t)__clzll((long long)m);
      break;
    }
  }
  return result;
}

/* glibc's expf (bbs_expf) with its 32-entry table across the lanes of a
 * register pair, lane i < 32 holding entry i. The chain's acceptance test
 * takes expf once per step of a uniform argument, so the entry comes by two
 * readlanes instead of two dependent memory loads (the table's address,
 * then the entry). */
struct Exp2Lanes {
  uint32_t lo, hi;
};

__device__ __forceinline__ Exp2Lanes exp2_lanes() {
  const uint64_t e = bbs_exp2f_table[lane_id() & 31u];
  return Exp2Lanes{(uint32_t)e, (uint32_t)(e >> 32)};
}

/* bbs_expf(x) for x equal on every lane */
__device__ __forceinline__ float expf_lanes(float x, const Exp2Lanes &t) {
  float special;
  double r;
  uint64_t ki;
  if (bbs_expf_head(x, &special, &r, &ki)) return special;
  const uint32_t i = rfl((uint32_t)ki & 31u);
  const uint64_t e = ((uint64_t)readlane32(t.hi, i) << 32) | readlane32(t.lo, i);
  return bbs_expf_tail(r, ki, e);
}

/* popcount(x) + acc as one v_bcnt_u32_b32 (the compiler would otherwise
 * re-associate a chain of popcounts into separate counts plus adds). */
__device__ __forceinline__ uint32_t bcnt_acc(uint32_t x, uint32_t acc) {
  uint32_t r;
  asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
  return r;
}

/* An empty asm that "rewrites" v: code using v after it cannot be
 * scheduled before it (volatile asm statements keep their order). */
__device__ __forceinline__ void pin4(float4 &v) {
  asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
}
__device__ __forceinline__ void pin_row(uint4 &v) {
  asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
}

/* IEEE maximum of two arithmetic results as one v_max_f32: both operands
 * are quiet (never signalling) NaNs at worst, so no canonicalisation is
 * needed; a NaN operand yields the other operand, as fmaxf. */
__device__ __forceinline__ float vmax(float a, float b) {
  float r;
  asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

/* vmax(vmax(a, b), c) as one v_max3_f32, which gfx950 defines as exactly
 * that pair of v_max_f32 operations (same NaN and signed-zero rules). */
__device__ __forceinline__ float vmax3(float a, float b, float c) {
  float r;
  asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

/* Cell slots per lane: CS = 1 holds cells 0..63; CS = 2 cells 0..127 in two
 * slots; CS = 3 cells 0..95 with the second slot "paired": lanes l and
 * l + 32 both hold cell 64 + l % 32 and score different nodes, so the 32
 * cells above 64 cost half a slot. */
template <int CS>
struct Slots {
  static constexpr int n = CS == 1 ? 1 : 2;
};

/* Mutation data of the cells this lane scores. */
template <int CS, int W>
struct Cells {
  uint32_t km[Slots<CS>::n][W];  /* K & M       */
  uint32_t knm[Slots<CS>::n][W]; /* K & ~M      */
  float fb[Slots<CS>::n];        /* |K & M|     */
  float fa[Slots<CS>::n];        /* |K & ~M|    */
};

/* Cell held by this lane in slot j. */
template <int CS>
__device__ __forceinline__ uint32_t slot_cell(int j) {
  if (j == 0) return lane_id();
  return CS == 3 ? 64u + (lane_id() & 31u) : 64u + lane_id();
}

/* cells: [n_cells][4] words (K&M lo, K&M hi, K&~M lo, K&~M hi). */
template <int CS, int W>
__device__ __forceinline__ void load_cells(const uint64_t *cells,
                                           uint32_t n_cells, Cells<CS, W> &c) {
#pragma unroll
  for (int j = 0; j < Slots<CS>::n; j++) {
    uint32_t cell = slot_cell<CS>(j);
    uint64_t q[4] = {0ull, 0ull, 0ull, 0ull};
    if (cell < n_cells) {
#pragma unroll
      for (int i = 0; i < 4; i++) q[i] = cells[4u * cell + (uint32_t)i];
    }
    uint32_t km32[4] = {(uint32_t)q[0], (uint32_t)(q[0] >> 32), (uint32_t)q[1],
                        (uint32_t)(q[1] >> 32)};
    uint32_t knm32[4] = {(uint32_t)q[2], (uint32_t)(q[2] >> 32),
                         (uint32_t)q[3], (uint32_t)(q[3] >> 32)};
    uint32_t b = 0, a = 0;
#pragma unroll
    for (int w = 0; w < W; w++) {
      c.km[j][w] = km32[w];
      c.knm[j][w] = knm32[w];
      b += (uint32_t)__popc(km32[w]);
      a += (uint32_t)__popc(knm32[w]);
    }
    c.fb[j] = (float)b;
    c.fa[j] = (float)a;
  }
}

/* Per-wave LDS: a copy of the tree's ancestor rows (one uint4 per node,
 * read as broadcasts by the scorer) and the per-cell maxima for the ordered
 * sum. */
struct WaveLds {
  uint4 *rows;  /* [n_nodes] ancestor rows of the tree being scored       */
  float *sums;  /* [cells of the layout] per-cell maxima                 */
  float *pen;   /* [TableView::pen_len] incremental scoring: 0 for a node
                 * whose stored pair scores are current, -inf for a changed
                 * or padding node                                        */
};
/* LDS of one wave for the full scorer */
struct WaveLdsStore {
  uint4 rows[128];
  float sums[128];
  __device__ WaveLds view() { return WaveLds{rows, sums, nullptr}; }
};

/* The four log-probabilities of src/TreeScorer.hpp:99-106,136-137. */
struct LogProbs {
  float L00, L01, L10, L11;
};

/* Likelihood of one (cell, node) pair from its two popcounts, in the
 * reference's order (((n00 L00 + n01 L01) + n10 L10) + n11 L11)
 * (src/TreeScorer.hpp:163-182). The reference's leading 0.0f + only turns
 * -0.0 into +0.0 for cells without known entries, which never changes the
 * ordered cell sum. */
__device__ __forceinline__ float pair_score(uint32_t n01, uint32_t n11,
                                            float fa, float fb,
                                            const LogProbs &L) {
  float f01 = (float)n01, f11 = (float)n11;
  float f00 = fa - f01, f10 = fb - f11; /* exact small integers */
  float s = f00 * L.L00 + f01 * L.L01;
  s = s + f10 * L.L10;
  s = s + f11 * L.L11;
  return s;
}

/* The two popcounts of one (cell, node) pair: n11 = |K&M&A|,
 * n01 = |K&~M&A|. */
template <int W>
__device__ __forceinline__ void pair_counts(const uint32_t (&km)[W],
                                            const uint32_t (&knm)[W], const uint4 &A,
                                            uint32_t &n01, uint32_t &n11) {
  n11 = bcnt_acc(km[0] & A.x, 0u);
  n01 = bcnt_acc(knm[0] & A.x, 0u);
  if (W > 1) {
    n11 = bcnt_acc(km[W > 1 ? 1 : 0] & A.y, n11);
    n01 = bcnt_acc(knm[W > 1 ? 1 : 0] & A.y, n01);
  }
  if (W > 2) {
    n11 = bcnt_acc(km[W > 2 ? 2 : 0] & A.z, n11);
    n01 = bcnt_acc(knm[W > 2 ? 2 : 0] & A.z, n01);
  }
  if (W > 3) {
    n11 = bcnt_acc(km[W > 3 ? 3 : 0] & A.w, n11);
    n01 = bcnt_acc(knm[W > 3 ? 3 : 0] & A.w, n01);
  }
}

/* ((0 + x[0]) + x[1]) + ... over 4 * NQ floats of LDS (NQ a multiple of 4),
 * straight-line code: while one group of 16 is added the reads of the next
 * are in flight.  The empty asm on the running sum pins each group's
 * adds before the next group's reads, so the compiler neither issues every
 * read up front (registers) nor each read next to its use (latency). */
template <int NQ>
__device__ __forceinline__ float sum_padded(const float *x) {
  static_assert(NQ % 4 == 0, "groups of 16 floats");
  constexpr int NG = NQ / 4;
  const float4 *q = reinterpret_cast<const float4 *>(x);
  float4 g[NG][4];
  float T = 0.0f;
#pragma unroll
  for (int k = 0; k < NG + 1; k++) {
    if (k < NG) {
#pragma unroll
      for (int i = 0; i < 4; i++) g[k < NG ? k : 0][i] = q[4 * (k < NG ? k : 0) + i];
    }
    if (k >= 1) {
      const int a = k - 1;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        T = T + g[a][i].x;
        T = T + g[a][i].y;
        T = T + g[a][i].z;
        T = T + g[a][i].w;
      }
      asm volatile("" : "+v"(T));
    }
  }
  return T;
}

/* The pair-score table of the current tree in LDS, node-major per cell:
 * the scores of cell c over all nodes form row c of `stride` floats.
 * stride = 8 ceil(n_nodes / 8) + 4 is 4 times an odd number, so the 16-byte
 * reads of 16 consecutive lanes (one row each) hit distinct banks; a row also
 * covers the padding nodes the 4-node groups of the scan read. There are
 * rows only for the dataset's n_cells cells. A lane without a cell (cell >=
 * n_cells, up to the layout's slots) reads and writes a spare row that
 * starts 4 (cell - n_cells) floats past the last cell's row: its values are
 * never used, and the spare rows of different lanes start at different
 * 16-byte offsets, so their writes do not collide. */
struct TableView {
  float *t;
  uint32_t stride;
  uint32_t n_cells;
  __host__ __device__ static uint32_t stride_for(uint32_t n_nodes) {
    return ((n_nodes + 7u) & ~7u) + 4u;
  }
  __host__ __device__ static uint32_t pen_len(uint32_t n_nodes) { return stride_for(n_nodes) + 4u; }
  /* floats past a row's start or S.pen that the unchanged-node scan may
   * read: 64 NS + 8 for NS = 1 node slot up to 64 nodes, else 2 */
  __host__ __device__ static uint32_t scan_reach(uint32_t n_nodes) {
    return n_nodes <= 64u ? 72u : 136u;
  }
  /* floats of a table for a layout of `slots` cells (64, 96 or 128) */
  __host__ __device__ static uint32_t floats_for(uint32_t n_nodes, uint32_t n_cells,
                                                 uint32_t slots) {
    const uint32_t st = stride_for(n_nodes);
    return n_cells < slots ? n_cells * st + 4u * (slots - n_cells - 1u) + st : slots * st;
  }
  __device__ float *row(uint32_t cell) const {
    return cell < n_cells ? t + cell * stride : t + n_cells * stride + 4u * (cell - n_cells);
  }
};

template <int CS>
struct TableRows {
  static constexpr uint32_t n = CS == 1 ? 64u : CS == 3 ? 96u : 128u;
};

/* The tree's ancestor rows (dm, one per node) into LDS for broadcast reads. */
template <int NS>
__device__ __forceinline__ void store_rows(const uint64_t (&dm)[NS][NS], uint32_t n_nodes,
                                           WaveLds &S) {
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const uint32_t x = lane_id() + 64u * (uint32_t)k;
    const uint64_t lo = dm[k][0], hi = NS > 1 ? dm[k][NS - 1] : 0ull;
    if (x < n_nodes)
      S.rows[x] =
          make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
  }
  wave_lds_sync();
}

/* S.pen for a proposal: -inf for the nodes whose ancestor row changed
 * (this lane's `diff` per node slot) and for padding nodes, 0 otherwise. */
template <int NS>
__device__ __forceinline__ void store_pen(const bool (&diff)[NS], uint32_t n_nodes, WaveLds &S) {
  const float ninf = __int_as_float(0xff800000);
  const uint32_t len = TableView::pen_len(n_nodes);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const uint32_t x = lane_id() + 64u * (uint32_t)k;
    bool d = x >= n_nodes;
#pragma unroll
    for (int j = 0; j < NS; j++)
      if (j == k) d = d || diff[j];
    if (x < len) S.pen[x] = d ? ninf : 0.0f;
  }
}

/* The ordered cell sum of S.sums (src/TreeScorer.hpp:198-205): the layout's
 * cells, +0.0 past n_cells. */
template <int CS>
__device__ __forceinline__ float ordered_sum(uint32_t n_cells, WaveLds &S) {
  wave_lds_sync();
  float T;
  /* padded to the next multiple of 16 cells */
  if constexpr (CS == 1) {
    T = n_cells <= 16u   ? sum_padded<4>(S.sums)
        : n_cells <= 32u ? sum_padded<8>(S.sums)
        : n_cells <= 48u ? sum_padded<12>(S.sums)
                         : sum_padded<16>(S.sums);
  } else if constexpr (CS == 3) {
    T = n_cells <= 80u ? sum_padded<20>(S.sums) : sum_padded<24>(S.sums);
  } else {
    T = n_cells <= 112u ? sum_padded<28>(S.sums) : sum_padded<32>(S.sums);
  }
  wave_lds_sync();
  return rfl_f(T);
}

/* A chain split over several waves of one workgroup by cells: this wave
 * holds cells [c0, c0 + its count) (one cell slot, its own pair-score table
 * and rows in LDS). Each scoring's per-cell maxima of all waves meet in
 * `shared` (n_total floats, +0.0 after them up to 128), and every wave then
 * takes the same ordered sum, so all waves make the same decisions. Two
 * shared buffers alternate between consecutive scorings: a wave writes a
 * buffer again only after the next scoring's barrier, which the other waves
 * reach only after their reads of it. */
struct CellSplit {
  float *shared;    /* the buffer of this scoring (2 x 128 floats in all) */
  uint32_t c0, n_total;
};

/* ((0 + x[0]) + x[1]) + ... over n <= 128 floats of LDS, +0.0 past n, padded
 * to the next multiple of 16 */
__device__ __forceinline__ float ordered_sum_n(uint32_t n, const float *x) {
  if (n <= 64u) {
    return n <= 16u ? sum_padded<4>(x) : n <= 32u ? sum_padded<8>(x)
         : n <= 48u ? sum_padded<12>(x) : sum_padded<16>(x);
  }
  return n <= 80u ? sum_padded<20>(x) : n <= 96u ? sum_padded<24>(x)
       : n <= 112u ? sum_padded<28>(x) : sum_padded<32>(x);
}

/* finish_sum of a cell-split chain (one cell slot per wave): this wave's
 * maxima into the shared buffer, a workgroup barrier, the ordered sum over
 * every wave's cells. */
__device__ __forceinline__ float finish_sum_split(float best, bool nan0, uint32_t n_cells,
                                                  const CellSplit &sp) {
  const uint32_t lane = lane_id();
  if (lane < n_cells) sp.shared[sp.c0 + lane] = nan0 ? __int_as_float(0x7fc00000) : best;
  __syncthreads();
  return rfl_f(ordered_sum_n(sp.n_total, sp.shared));
}

/* Per-cell maxima -> ordered cell sum (src/TreeScorer.hpp:198-205).
 * best[j]/nan0[j] of the full slots, and for the paired slot the maxima of
 * the two lane halves. Cells past n_cells store +0.0: the running sum starts
 * at +0.0 and never becomes -0.0, so adding +0.0 leaves it bit-identical. */
template <int CS>
__device__ __forceinline__ float finish_sum(const float (&best)[2], const bool (&nan0)[2],
                                            uint32_t n_cells, WaveLds &S) {
  constexpr int SL = Slots<CS>::n;
  const float qnan = __int_as_float(0x7fc00000);
  const uint32_t lane = lane_id();
  S.sums[lane] = lane < n_cells ? (nan0[0] ? qnan : best[0]) : 0.0f;
  if constexpr (CS == 3) {
    /* cell 64 + i: the maximum over the nodes of lane i and of lane 32 + i;
     * node 0 is always scored by the low half, so only its flag can make the
     * cell NaN (a NaN of a later node never wins, as in the reference) */
    const float x = nan0[1] ? qnan : best[1];
    const float y = __shfl_xor(x, 32);
    const float m = x != x ? x : vmax(x, y);
    if (lane < 32u) S.sums[64u + lane] = 64u + lane < n_cells ? m : 0.0f;
  } else if (SL > 1) {
    S.sums[64u + lane] = 64u + lane < n_cells ? (nan0[1] ? qnan : best[1]) : 0.0f;
  }
  return ordered_sum<CS>(n_cells, S);
}

/* Node loop of the full scorer over the rows in S.rows: per cell the
 * maximum over nodes (node 0 seeds it, later nodes replace it only when
 * strictly greater, src/TreeScorer.hpp:185-194). With TAB every pair score is
 * also stored in the table. With one wave per SIMD nothing else hides the LDS
 * latency, so rows are read one node pair ahead; past the last node the
 * clamped index repeats node n - 1, which rewrites its own score and cannot
 * change a maximum (nor, being equal, the first maximal node). */
template <int NS, int CS, int W, bool COUNTS, bool TAB, int J0 = 0, bool T0 = false>
__device__ __forceinline__ void node_maxima(const Cells<CS, W> &c, uint32_t n_nodes,
                                            const LogProbs &L, WaveLds &S, TableView tab,
                                            float (&best)[2], bool (&nan0)[2],
                                            uint32_t (&arg01)[2], uint32_t (&arg11)[2]) {
  static_assert(!COUNTS || CS != 3, "counts need full cell slots");
  static_assert((J0 == 0 && !T0) || CS == 2, "slot ranges and half tables: two full slots");
  constexpr int SL = Slots<CS>::n;
  const uint32_t lane = lane_id();
  float *trow[2] = {nullptr, nullptr};
  if (TAB) {
    trow[0] = tab.row(slot_cell<CS>(0));
    if (!T0) trow[1] = tab.row(slot_cell<CS>(1));
  }
  if constexpr (J0 == 0)
    nan0[0] = nan0[1] = false;
  else
    nan0[1] = false;
  /* one node for the full slots (from slot J0 on) */
  auto full = [&](const uint4 &A, uint32_t node, bool first) {
#pragma unroll
    for (int j = J0; j < (CS == 3 ? 1 : SL); j++) {
      uint32_t n01, n11;
      pair_counts<W>(c.km[j], c.knm[j], A, n01, n11);
      const float s = pair_score(n01, n11, c.fa[j], c.fb[j], L);
      if (TAB && !(T0 && j == 1)) trow[j][node] = s;
      if (first) {
        best[j] = s;
        nan0[j] = s != s;
        if (COUNTS) {
          arg01[j] = n01;
          arg11[j] = n11;
        }
      } else if (COUNTS) {
        if (s > best[j]) {
          best[j] = s;
          arg01[j] = n01;
          arg11[j] = n11;
        }
      } else {
        best[j] = vmax(best[j], s);
      }
    }
  };
  /* the paired slot: lanes < 32 score node v, lanes >= 32 node v + 1 */
  auto half = [&](const uint4 &A, uint32_t node, bool first) {
    uint32_t n01, n11;
    constexpr int H = CS == 3 ? 1 : 0;
    pair_counts<W>(c.km[H], c.knm[H], A, n01, n11);
    const float s = pair_score(n01, n11, c.fa[H], c.fb[H], L);
    if (TAB) trow[1][node] = s;
    if (first) {
      best[1] = s;
      nan0[1] = lane < 32u && s != s;
    } else {
      best[1] = vmax(best[1], s);
    }
  };
  const uint32_t hi_half = lane >> 5; /* 0 or 1 */
  const uint32_t last = n_nodes - 1u;
  auto clampn = [&](uint32_t v) { return v < last ? v : last; };
  if constexpr (CS == 3) {
    full(S.rows[0], 0u, true);
    half(S.rows[clampn(hi_half)], clampn(hi_half), true);
    if (n_nodes > 1u) full(S.rows[1], 1u, false);
    uint4 A0 = S.rows[clampn(2u)], A1 = S.rows[clampn(3u)], Ah = S.rows[clampn(2u + hi_half)];
    for (uint32_t v = 2; v < n_nodes; v += 2) {
      uint4 B0 = A0, B1 = A1, Bh = Ah;
      A0 = S.rows[clampn(v + 2u)];
      A1 = S.rows[clampn(v + 3u)];
      Ah = S.rows[clampn(v + 2u + hi_half)];
      asm volatile("" ::: "memory");
      pin_row(B0);
      pin_row(B1);
      pin_row(Bh);
      full(B0, v, false);
      half(Bh, clampn(v + hi_half), false);
      full(B1, clampn(v + 1u), false);
    }
  } else {
    full(S.rows[0], 0u, true);
    uint4 A1 = S.rows[clampn(1u)], A2 = S.rows[clampn(2u)];
    for (uint32_t v = 1; v < n_nodes; v += 2) {
      uint4 B1 = A1, B2 = A2;
      A1 = S.rows[clampn(v + 2u)];
      A2 = S.rows[clampn(v + 3u)];
      asm volatile("" ::: "memory");
      pin_row(B1);
      pin_row(B2);
      full(B1, v, false);
      full(B2, clampn(v + 1u), false);
    }
  }
}

/* ---- pair counts on the matrix cores (v_mfma_i32_32x32x32_i8) ---------
 * n11 = (K&M) A^T and n01 = (K&~M) A^T over 32 x 32 tiles, rows = nodes,
 * columns = cells, genes as the inner dimension. Lane l (h = l / 32) holds
 * in A node 32 mt + l % 32 and in B cell 32 jt + l % 32, each with genes
 * 32 kb + 16 h + b in byte b; in D cell 32 jt + l % 32 and node
 * 32 mt + 8 (r / 4) + 4 h + r % 4 in register r. The root's bit (node
 * n_nodes - 1) meets a zero in K&M and K&~M, so W gene words cover the
 * inner dimension. */
typedef int bbs_v4i __attribute__((ext_vector_type(4)));
typedef int bbs_v16i __attribute__((ext_vector_type(16)));

/* Bits 0..15 of x as 16 bytes 0/1 (byte b = bit b): each nibble times
 * 0x00204081 places its bits 0..3 at bits 0, 8, 16, 24 without carries. */
__device__ __forceinline__ bbs_v4i bits_to_bytes16(uint32_t x) {
  bbs_v4i r;
#pragma unroll
  for (int q = 0; q < 4; q++) r[q] = (int)((((x >> (4 * q)) & 0xFu) * 0x00204081u) & 0x01010101u);
  return r;
}

/* The data operands of a dataset (cells of the layout in tiles of 32). */
template <int CS, int W>
struct MfmaCells {
  static constexpr int JT = CS == 1 ? 2 : CS == 3 ? 3 : 4;
  bbs_v4i km[JT][W], knm[JT][W];
  float fa[JT], fb[JT]; /* |K&~M|, |K&M| of the lane's column cell */
};

/* cells: [n_cells][4] words (K&M lo, K&M hi, K&~M lo, K&~M hi). */
template <int CS, int W>
__device__ __forceinline__ void load_mfma_cells(const uint64_t *cells, uint32_t n_cells,
                                                MfmaCells<CS, W> &m) {
  const uint32_t lane = lane_id(), h = lane >> 5;
#pragma unroll
  for (int jt = 0; jt < MfmaCells<CS, W>::JT; jt++) {
    const uint32_t c = 32u * (uint32_t)jt + (lane & 31u);
    uint64_t q[4] = {0ull, 0ull, 0ull, 0ull};
    if (c < n_cells) {
#pragma unroll
      for (int i = 0; i < 4; i++) q[i] = cells[4u * c + (uint32_t)i];
    }
    const uint32_t km32[4] = {(uint32_t)q[0], (uint32_t)(q[0] >> 32), (uint32_t)q[1],
                              (uint32_t)(q[1] >> 32)};
    const uint32_t knm32[4] = {(uint32_t)q[2], (uint32_t)(q[2] >> 32), (uint32_t)q[3],
                               (uint32_t)(q[3] >> 32)};
    uint32_t b = 0u, a = 0u;
#pragma unroll
    for (int kb = 0; kb < W; kb++) {
      m.km[jt][kb] = bits_to_bytes16((km32[kb] >> (16u * h)) & 0xFFFFu);
      m.knm[jt][kb] = bits_to_bytes16((knm32[kb] >> (16u * h)) & 0xFFFFu);
    }
#pragma unroll
    for (int w = 0; w < 4; w++) {
      b += (uint32_t)__popc(km32[w]);
      a += (uint32_t)__popc(knm32[w]);
    }
    m.fa[jt] = (float)a;
    m.fb[jt] = (float)b;
  }
}

/* Every pair score of the tree in S.rows into the table, the counts on the
 * matrix cores and the float part as pair_score. With SUM, the per-cell
 * maxima (node 0 seeds them: its NaN sticks, later NaNs are ignored; the
 * maximum is order-free otherwise) go to S.sums and the ordered cell sum
 * is returned, as score_tree with a table. */
template <int NS, int CS, int W, bool SUM>
__device__ __forceinline__ float mfma_maxima(const MfmaCells<CS, W> &mc, uint32_t n_nodes,
                                             uint32_t n_cells, const LogProbs &L, WaveLds &S,
                                             TableView tab) {
  constexpr int MT = 2 * NS, JT = MfmaCells<CS, W>::JT;
  const uint32_t lane = lane_id(), h = lane >> 5;
  const uint32_t last = n_nodes - 1u;
  bbs_v4i at[MT][W];
#pragma unroll
  for (int mt = 0; mt < MT; mt++) {
    uint32_t n = 32u * (uint32_t)mt + (lane & 31u);
    n = n < last ? n : last;
    const uint4 row = S.rows[n];
    const uint32_t w[4] = {row.x, row.y, row.z, row.w};
#pragma unroll
    for (int kb = 0; kb < W; kb++) at[mt][kb] = bits_to_bytes16((w[kb] >> (16u * h)) & 0xFFFFu);
  }
  const float ninf = __int_as_float(0xff800000);
  const float qnan = __int_as_float(0x7fc00000);
#pragma unroll
  for (int jt = 0; jt < JT; jt++) {
    const uint32_t cell = 32u * (uint32_t)jt + (lane & 31u);
    if (32u * (uint32_t)jt >= n_cells) {
      if (SUM && lane < 32u) S.sums[cell] = 0.0f;
      continue;
    }
    float *const trow = tab.row(cell);
    float best = ninf;
    bool nan0 = false;
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
      if (32u * (uint32_t)mt >= n_nodes) break;
      bbs_v16i c11 = {}, c01 = {};
#pragma unroll
      for (int kb = 0; kb < W; kb++) {
        c11 = __builtin_amdgcn_mfma_i32_32x32x32_i8(at[mt][kb], mc.km[jt][kb], c11, 0, 0, 0);
        c01 = __builtin_amdgcn_mfma_i32_32x32x32_i8(at[mt][kb], mc.knm[jt][kb], c01, 0, 0, 0);
      }
      auto tile_pairs = [&](bool full) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const uint32_t node = 32u * (uint32_t)mt + 8u * (uint32_t)(r / 4) + 4u * h +
                                (uint32_t)(r % 4);
          const float sc = pair_score((uint32_t)c01[r], (uint32_t)c11[r], mc.fa[jt], mc.fb[jt], L);
          if (mt == 0 && r == 0) nan0 = h == 0u && sc != sc;
          const bool valid = full || node < n_nodes;
          if (valid) trow[node] = sc;
          if (SUM) best = vmax(best, valid ? sc : ninf);
        }
      };
      if (32u * (uint32_t)(mt + 1) <= n_nodes)
        tile_pairs(true);
      else
        tile_pairs(false);
    }
    if (SUM) {
      /* the cell's other nodes of each tile are in lane l ^ 32 */
      const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(best),
                                                       __float_as_uint(best), false, false);
      const float other = __uint_as_float(sw[1]);
      if (lane < 32u) S.sums[cell] = cell < n_cells ? (nan0 ? qnan : vmax(best, other)) : 0.0f;
    }
  }
  if (!SUM) return 0.0f;
  return ordered_sum<CS>(n_cells, S);
}

/*
 * TreeScorer::logscore_tree without the beta prior (src/TreeScorer.hpp:
 * 131-205): per cell the maximum over nodes (a NaN at node 0 sticks, later
 * NaNs are ignored, as std::max(old, new) does), then the sum over cells in
 * order. Lanes hold cells; the tree's rows are copied to S.rows and read as
 * LDS broadcasts. Returns the same value in every lane. With COUNTS (CS 1
 * or 2 only), n01/n11 of the first maximal node are returned per slot.
 */
template <int NS, int CS, int W, bool COUNTS, bool TAB = false, bool SPL = false,
          bool T0 = false>
__device__ __forceinline__ float
score_tree(const uint64_t (&dm)[NS][NS], const Cells<CS, W> &c, uint32_t n_nodes,
           uint32_t n_cells, const LogProbs &L, WaveLds &S,
           uint32_t (&arg01)[2], uint32_t (&arg11)[2], TableView tab = TableView{nullptr, 0u, 0u},
           const CellSplit &sp = CellSplit{nullptr, 0u, 0u}) {
  store_rows<NS>(dm, n_nodes, S);
  float best[2];
  bool nan0[2];
  node_maxima<NS, CS, W, COUNTS, TAB, 0, T0>(c, n_nodes, L, S, tab, best, nan0, arg01, arg11);
  if constexpr (SPL) {
    static_assert(CS == 1, "a split chain holds one cell slot per wave");
    return finish_sum_split(best[0], nan0[0], n_cells, sp);
  } else {
    return finish_sum<CS>(best, nan0, n_cells, S);
  }
}

/* Lowest set bit of a node mask held uniformly by the wave (bit v = node v),
 * cleared from the mask; the mask must not be empty. */
template <int NS>
__device__ __forceinline__ uint32_t pop_node(uint64_t (&m)[NS]) {
  if (NS == 1 || m[0] != 0ull) {
    const uint32_t v = (uint32_t)__ffsll((unsigned long long)m[0]) - 1u;
    m[0] &= m[0] - 1ull;
    return v;
  }
  const uint32_t v = 64u + (uint32_t)__ffsll((unsigned long long)m[NS - 1]) - 1u;
  m[NS - 1] &= m[NS - 1] - 1ull;
  return v;
}

/* Pair scores of the nodes in `changed` (rows in S.rows), two nodes at a
 * time so that the paired half slot serves both. WRITE stores them in the
 * table; otherwise their per-cell maximum goes to R, and nan0 flags a NaN
 * score of node 0 when node 0 is changed. The rows of the next pair are read
 * while the current pair is scored, alternating between two row buffers (at
 * the end one pair of rows is read and not used). */
template <int NS, int CS, int W, bool WRITE, bool T0 = false>
__device__ __forceinline__ void changed_nodes(const uint64_t (&changed)[NS], const Cells<CS, W> &c,
                                              const LogProbs &L, WaveLds &S, TableView tab,
                                              float (&R)[2], bool (&nan0)[2]) {
  static_assert(!T0 || CS == 2, "a half table: two full slots");
  constexpr int SL = T0 ? 1 : Slots<CS>::n; /* the slots this pass scores */
  const uint32_t lane = lane_id();
  const uint32_t hi_half = lane >> 5;
  float *trow[2] = {nullptr, nullptr};
  if (WRITE) {
    trow[0] = tab.row(slot_cell<CS>(0));
    if (!T0) trow[1] = tab.row(slot_cell<CS>(1));
  }
  uint64_t m[NS];
#pragma unroll
  for (int k = 0; k < NS; k++) m[k] = changed[k];
  auto nonempty = [&]() {
    bool r = false;
#pragma unroll
    for (int k = 0; k < NS; k++) r |= m[k] != 0ull;
    return r;
  };
  if (!nonempty()) return;
  struct Pair {
    uint32_t a, b;
    uint4 Aa, Ab, Ah;
  };
  /* the next pair (or the last one again) and its rows; the empty asm
   * after the reads passes the current rows through, so the current pair's
   * arithmetic cannot be scheduled ahead of the next pair's reads */
  auto fetch = [&](Pair &cur, Pair &nx) {
    nx.a = cur.a;
    nx.b = cur.b;
    if (nonempty()) {
      nx.a = pop_node<NS>(m);
      nx.b = nonempty() ? pop_node<NS>(m) : nx.a;
    }
    nx.Aa = S.rows[nx.a];
    nx.Ab = S.rows[nx.b];
    nx.Ah = S.rows[hi_half ? nx.b : nx.a];
    asm volatile(""
                 : "+v"(cur.Aa.x), "+v"(cur.Aa.y), "+v"(cur.Aa.z), "+v"(cur.Aa.w),
                   "+v"(cur.Ab.x), "+v"(cur.Ab.y), "+v"(cur.Ab.z), "+v"(cur.Ab.w),
                   "+v"(cur.Ah.x), "+v"(cur.Ah.y), "+v"(cur.Ah.z), "+v"(cur.Ah.w)
                 :
                 : "memory");
  };
  auto score = [&](const Pair &p) {
#pragma unroll
    for (int j = 0; j < (CS == 3 ? 1 : SL); j++) {
      uint32_t n01, n11;
      pair_counts<W>(c.km[j], c.knm[j], p.Aa, n01, n11);
      const float sa = pair_score(n01, n11, c.fa[j], c.fb[j], L);
      pair_counts<W>(c.km[j], c.knm[j], p.Ab, n01, n11);
      const float sb = pair_score(n01, n11, c.fa[j], c.fb[j], L);
      if (WRITE) {
        trow[j][p.a] = sa;
        trow[j][p.b] = sb;
      } else {
        R[j] = vmax3(R[j], sa, sb);
      }
    }
    if constexpr (CS == 3) {
      uint32_t n01, n11;
      pair_counts<W>(c.km[1], c.knm[1], p.Ah, n01, n11);
      const float sh = pair_score(n01, n11, c.fa[1], c.fb[1], L);
      if (WRITE)
        trow[1][hi_half ? p.b : p.a] = sh;
      else
        R[1] = vmax(R[1], sh);
    }
  };
  Pair P, Q;
  P.a = pop_node<NS>(m);
  P.b = nonempty() ? pop_node<NS>(m) : P.a;
  P.Aa = S.rows[P.a];
  P.Ab = S.rows[P.b];
  P.Ah = S.rows[hi_half ? P.b : P.a];
  while (true) {
    bool more = nonempty();
    fetch(P, Q);
    score(P);
    if (!more) break;
    more = nonempty();
    fetch(Q, P);
    score(Q);
    if (!more) break;
  }
  if (!WRITE && (changed[0] & 1ull) != 0ull) {
    /* node 0 is changed: its new score decides the NaN flags */
    const uint4 A0 = S.rows[0];
#pragma unroll
    for (int j = 0; j < (CS == 3 ? 1 : SL); j++) {
      uint32_t n01, n11;
      pair_counts<W>(c.km[j], c.knm[j], A0, n01, n11);
      const float s0 = pair_score(n01, n11, c.fa[j], c.fb[j], L);
      nan0[j] = s0 != s0;
    }
    if constexpr (CS == 3) {
      uint32_t n01, n11;
      pair_counts<W>(c.km[1], c.knm[1], A0, n01, n11);
      const float s0 = pair_score(n01, n11, c.fa[1], c.fb[1], L);
      nan0[1] = hi_half == 0u && s0 != s0;
    }
  }
}

/* Per-cell maximum over the stored scores of the nodes not in `changed`:
 * score + S.pen[node] (0 or -inf) for every node, branch-free, four nodes
 * per 16-byte read of this lane's table row. Full slots scan nodes
 * [0, 4 ceil(n/4)); in the paired half slot lanes < 32 scan the first
 * 4 ceil(n/8) nodes and lanes >= 32 the next as many. Padding nodes have
 * pen -inf. Adding +0.0 turns a -0.0 score (a cell without known entries,
 * where every node scores -0.0) into +0.0, which the ordered cell sum does
 * not see. The next group's reads are issued before the current group's
 * maxima; the empty asm with a memory clobber keeps them there. */
template <int NS, int CS, bool T0 = false>
__device__ __forceinline__ void unchanged_nodes(bool node0_changed, uint32_t n_nodes,
                                                TableView tab, const WaveLds &S,
                                                float (&U)[2], bool (&nan0)[2]) {
  static_assert(!T0 || CS == 2, "a half table: two full slots");
  constexpr int SL = T0 ? 1 : Slots<CS>::n; /* the slots the table holds */
  const uint32_t lane = lane_id();
  const uint32_t hi_half = lane >> 5;
  const uint32_t gh = (n_nodes + 7u) >> 3; /* 4-node groups per half */
  const float *const rj[2] = {tab.row(slot_cell<CS>(0)),
                              T0 ? tab.row(slot_cell<CS>(0)) : tab.row(slot_cell<CS>(1))};
  const float *const rh = rj[1] + (hi_half ? 4u * gh : 0u);
  const float *const qh = S.pen + (hi_half ? 4u * gh : 0u);
  if (!node0_changed) {
#pragma unroll
    for (int j = 0; j < (CS == 3 ? 1 : SL); j++) {
      const float x = rj[j][0];
      nan0[j] = x != x;
    }
    if constexpr (CS == 3) {
      const float x = rj[1][0];
      if (hi_half == 0u) nan0[1] = x != x;
    }
  }
  /* one step = 8 nodes of the full slots (two groups) and one group of
   * each half of the paired slot */
  struct Grp {
    float4 x[2][2], q[2], xh, qhv;
  };
  /* reads groups g, g + 1 (g even) into b; the empty asm after them passes
   * the other buffer `o` through, so its maxima cannot be scheduled ahead of
   * these reads */
  auto rd = [&](uint32_t g, Grp &b, Grp &o) {
#pragma unroll
    for (int j = 0; j < (CS == 3 ? 1 : SL); j++) {
      b.x[j][0] = *reinterpret_cast<const float4 *>(rj[j] + 4u * g);
      b.x[j][1] = *reinterpret_cast<const float4 *>(rj[j] + 4u * g + 4u);
    }
    b.q[0] = *reinterpret_cast<const float4 *>(S.pen + 4u * g);
    b.q[1] = *reinterpret_cast<const float4 *>(S.pen + 4u * g + 4u);
    if constexpr (CS == 3) {
      b.xh = *reinterpret_cast<const float4 *>(rh + 2u * g);
      b.qhv = *reinterpret_cast<const float4 *>(qh + 2u * g);
    }
    asm volatile("" ::: "memory");
#pragma unroll
    for (int j = 0; j < (CS == 3 ? 1 : SL); j++) {
      pin4(o.x[j][0]);
      pin4(o.x[j][1]);
    }
    pin4(o.q[0]);
    pin4(o.q[1]);
    if constexpr (CS == 3) {
      pin4(o.xh);
      pin4(o.qhv);
    }
  };
  auto mx4 = [&](float &u, const float4 &x, const float4 &q) {
    u = vmax3(u, x.x + q.x, x.y + q.y);
    u = vmax3(u, x.z + q.z, x.w + q.w);
  };
  auto use = [&](const Grp &b) {
#pragma unroll
    for (int j = 0; j < (CS == 3 ? 1 : SL); j++) {
      mx4(U[j], b.x[j][0], b.q[0]);
      mx4(U[j], b.x[j][1], b.q[1]);
    }
    if constexpr (CS == 3) mx4(U[1], b.xh, b.qhv);
  };
  const uint32_t g4 = (n_nodes + 3u) >> 2; /* 4-node groups of the full slots */
  /* The reads of the next step are issued unconditionally and unclamped
   * (past the end they read beyond the last group and are not used): with
   * the same number of reads in flight on every path, each wait is for
   * exactly the reads it needs, and every address is a base plus an
   * immediate offset, so no per-group address stays live across steps. The
   * furthest read ends 64 NS + 8 floats past a row's start or S.pen; the
   * slab keeps that much LDS after the last row and after S.pen
   * (scan_reach). */
  Grp A, B;
#pragma unroll
  for (int j = 0; j < 2; j++) {
    A.x[j][0] = A.x[j][1] = make_float4(0.f, 0.f, 0.f, 0.f);
    A.q[j] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  A.xh = A.qhv = make_float4(0.f, 0.f, 0.f, 0.f);
  B = A;
  rd(0u, A, B);
  /* at most 64 NS nodes: straight-line code (exact waits, no loop merge) */
#pragma unroll
  for (int it = 0; it < NS * 4; it++) {
    const uint32_t g = 4u * (uint32_t)it;
    rd(g + 2u, B, A);
    use(A);
    if (g + 2u >= g4) break;
    rd(g + 4u, A, B);
    use(B);
    if (g + 4u >= g4) break;
  }
}

/* The tree score without the prior from the pair-score table of the
 * current tree, for a proposal whose ancestor rows (in S.rows) differ from
 * the current ones only at the nodes in `changed` (S.pen marks them): new
 * scores for those nodes, stored scores for the rest. The per-cell maximum over nodes does
 * not depend on the order the nodes are visited in (the first-wins rule
 * only matters for a NaN at node 0, kept by the flags), so the result is
 * bit-identical to the full scorer. */
template <int NS, int CS, int W, bool SPL = false, bool T0 = false>
__device__ __forceinline__ float score_changed(const uint64_t (&changed)[NS], const Cells<CS, W> &c,
                                               uint32_t n_nodes, uint32_t n_cells,
                                               const LogProbs &L, WaveLds &S, TableView tab,
                                               const CellSplit &sp = CellSplit{nullptr, 0u, 0u}) {
  const float ninf = __int_as_float(0xff800000);
  float R[2] = {ninf, ninf}, U[2] = {ninf, ninf};
  bool nan0[2] = {false, false};
  changed_nodes<NS, CS, W, false, T0>(changed, c, L, S, tab, R, nan0);
  unchanged_nodes<NS, CS, T0>((changed[0] & 1ull) != 0ull, n_nodes, tab, S, U, nan0);
  float best[2] = {vmax(R[0], U[0]), vmax(R[1], U[1])};
  if constexpr (T0) {
    /* the table holds slot 0 only: slot 1 over every node of the proposal
     * (its rows are in S.rows) */
    float b1[2];
    bool n1[2];
    uint32_t u01[2], u11[2];
    node_maxima<NS, CS, W, false, false, 1>(c, n_nodes, L, S, tab, b1, n1, u01, u11);
    best[1] = b1[1];
    nan0[1] = n1[1];
  }
  if constexpr (SPL) {
    static_assert(CS == 1, "a split chain holds one cell slot per wave");
    return finish_sum_split(best[0], nan0[0], n_cells, sp);
  } else {
    return finish_sum<CS>(best, nan0, n_cells, S);
  }
}

/* Constants of the beta prior (src/TreeScorer.hpp:108-111,122-126). */
struct PriorArgs {
  float am1, bm1; /* bpriora - 1, bpriorb - 1 */
  float lg_ab, lg_a, lg_b;
};

/* kb: the logf coefficients' index base (bbs_coef_base() in the chain
 * loop, else 0) */
__device__ __forceinline__ float logscore_beta(const PriorArgs &p, float beta, uint32_t kb = 0u) {
  return p.lg_ab + p.am1 * bbs_logf_at(beta, kb) + p.bm1 * bbs_logf_at(1.0f - beta, kb) -
         p.lg_a - p.lg_b;
}

} // namespace bbs

#endif /* BBS_DEVICE_H */
