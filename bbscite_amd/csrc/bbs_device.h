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

/* popcount(x) + acc as one v_bcnt_u32_b32 (the compiler would otherwise
 * re-associate a chain of popcounts into separate counts plus adds). */
__device__ __forceinline__ uint32_t bcnt_acc(uint32_t x, uint32_t acc) {
  uint32_t r;
  asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
  return r;
}

/* IEEE maximum of two arithmetic results as one v_max_f32: both operands
 * are quiet (never signalling) NaNs at worst, so no canonicalisation is
 * needed; a NaN operand yields the other operand, as fmaxf. */
__device__ __forceinline__ float vmax(float a, float b) {
  float r;
  asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
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
  uint4 rows[128];
  float sums[192];
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

/*
 * TreeScorer::logscore_tree without the beta prior (src/TreeScorer.hpp:
 * 131-205): per cell the maximum over nodes (a NaN at node 0 sticks, later
 * NaNs are ignored, as std::max(old, new) does), then the sum over cells in
 * order. Lanes hold cells; the tree's rows are copied to S.rows and read as
 * LDS broadcasts. Returns the same value in every lane. With COUNTS (CS 1
 * or 2 only), n01/n11 of the first maximal node are returned per slot.
 */
template <int NS, int CS, int W, bool COUNTS>
__device__ __forceinline__ float
score_tree(const uint64_t (&dm)[NS][NS], const Cells<CS, W> &c, uint32_t n_nodes,
           uint32_t n_cells, const LogProbs &L, WaveLds &S,
           uint32_t (&arg01)[2], uint32_t (&arg11)[2]) {
  static_assert(!COUNTS || CS != 3, "counts need full cell slots");
  constexpr int SL = Slots<CS>::n;
#pragma unroll
  for (int k = 0; k < NS; k++) {
    const uint64_t lo = dm[k][0], hi = NS > 1 ? dm[k][NS - 1] : 0ull;
    S.rows[lane_id() + 64u * (uint32_t)k] =
        make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
  }
  wave_lds_sync();
  float best[2];
  bool nan0[2] = {false, false};
  /* one node for the full slots (node 0 seeds the maximum) */
  auto full = [&](const uint4 &A, bool first) {
#pragma unroll
    for (int j = 0; j < (CS == 3 ? 1 : SL); j++) {
      uint32_t n01, n11;
      pair_counts<W>(c.km[j], c.knm[j], A, n01, n11);
      const float s = pair_score(n01, n11, c.fa[j], c.fb[j], L);
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
  auto half = [&](const uint4 &A, bool first) {
    uint32_t n01, n11;
    constexpr int H = CS == 3 ? 1 : 0;
    pair_counts<W>(c.km[H], c.knm[H], A, n01, n11);
    const float s = pair_score(n01, n11, c.fa[H], c.fb[H], L);
    if (first) {
      best[1] = s;
      nan0[1] = lane_id() < 32u && s != s;
    } else {
      best[1] = vmax(best[1], s);
    }
  };
  const uint32_t hi_half = lane_id() >> 5; /* 0 or 1 */
  const uint32_t last = n_nodes - 1u;
  auto clampn = [&](uint32_t v) { return v < last ? v : last; };
  /* rows are read one iteration ahead: with one wave per SIMD nothing else
   * hides the LDS latency */
  if constexpr (CS == 3) {
    full(S.rows[0], true);
    half(S.rows[clampn(hi_half)], true);
    if (n_nodes > 1u) full(S.rows[1], false);
    uint4 A0 = S.rows[clampn(2u)], A1 = S.rows[clampn(3u)], Ah = S.rows[clampn(2u + hi_half)];
    for (uint32_t v = 2; v < n_nodes; v += 2) {
      const uint4 B0 = A0, B1 = A1, Bh = Ah;
      A0 = S.rows[clampn(v + 2u)];
      A1 = S.rows[clampn(v + 3u)];
      Ah = S.rows[clampn(v + 2u + hi_half)];
      full(B0, false);
      half(Bh, false);
      if (v + 1u < n_nodes) full(B1, false);
    }
  } else {
    full(S.rows[0], true);
    uint4 A1 = S.rows[clampn(1u)], A2 = S.rows[clampn(2u)];
    for (uint32_t v = 1; v < n_nodes; v += 2) {
      const uint4 B1 = A1, B2 = A2;
      A1 = S.rows[clampn(v + 2u)];
      A2 = S.rows[clampn(v + 3u)];
      full(B1, false);
      if (v + 1u < n_nodes) full(B2, false);
    }
  }
  /* Cells past n_cells store +0.0: the running sum starts at +0.0 and never
   * becomes -0.0, so adding +0.0 leaves it bit-identical. */
  const float qnan = __int_as_float(0x7fc00000);
  const uint32_t lane = lane_id();
  S.sums[lane] = lane < n_cells ? (nan0[0] ? qnan : best[0]) : 0.0f;
  if constexpr (CS == 3) {
    /* cell 64 + i: the maximum over even nodes (lane i) and odd nodes (lane
     * 32 + i); a NaN from node 0 only comes from the even half */
    const float x = nan0[1] ? qnan : best[1];
    const float y = __shfl_xor(x, 32);
    if (lane < 32u) S.sums[64u + lane] = 64u + lane < n_cells ? (x != x ? x : vmax(x, y)) : 0.0f;
  } else if (SL > 1) {
    S.sums[64u + lane] = 64u + lane < n_cells ? (nan0[1] ? qnan : best[1]) : 0.0f;
  }
  wave_lds_sync();
  float T;
  if constexpr (CS == 1) {
    T = n_cells <= 16u ? sum_padded<4>(S.sums) : n_cells <= 32u ? sum_padded<8>(S.sums)
                                                                 : sum_padded<16>(S.sums);
  } else {
    T = sum_padded<CS == 3 ? 24 : 32>(S.sums);
  }
  wave_lds_sync();
  return rfl_f(T);
}

/* Constants of the beta prior (src/TreeScorer.hpp:108-111,122-126). */
struct PriorArgs {
  float am1, bm1; /* bpriora - 1, bpriorb - 1 */
  float lg_ab, lg_a, lg_b;
};

__device__ __forceinline__ float logscore_beta(const PriorArgs &p, float beta) {
  return p.lg_ab + p.am1 * bbs_logf(beta) + p.bm1 * bbs_logf(1.0f - beta) -
         p.lg_a - p.lg_b;
}

} // namespace bbs

#endif /* BBS_DEVICE_H */
