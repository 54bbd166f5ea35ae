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

/* Mutation data of the cells this lane scores. */
template <int CS, int W>
struct Cells {
  uint32_t km[CS][W];  /* K & M       */
  uint32_t knm[CS][W]; /* K & ~M      */
  float fb[CS];        /* |K & M|     */
  float fa[CS];        /* |K & ~M|    */
};

/* cells: [n_cells][4] words (K&M lo, K&M hi, K&~M lo, K&~M hi). */
template <int CS, int W>
__device__ __forceinline__ void load_cells(const uint64_t *cells,
                                           uint32_t n_cells, Cells<CS, W> &c) {
#pragma unroll
  for (int j = 0; j < CS; j++) {
    uint32_t cell = lane_id() + 64u * (uint32_t)j;
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

/*
 * TreeScorer::logscore_tree without the beta prior (src/TreeScorer.hpp:
 * 131-205): per cell the maximum over nodes (a NaN at node 0 sticks, later
 * NaNs are ignored, as std::max(old, new) does), then the sum over cells in
 * order. Returns the same value in every lane. lds: 64*CS floats of this
 * wave. With COUNTS, n01/n11 of the first maximal node are returned per slot.
 */
template <int NS, int CS, int W, bool COUNTS>
__device__ __forceinline__ float
score_tree(const uint64_t (&dm)[NS][NS], const Cells<CS, W> &c, uint32_t n_nodes,
           uint32_t n_cells, const LogProbs &L, float *lds,
           uint32_t (&arg01)[CS], uint32_t (&arg11)[CS]) {
  float best[CS];
  bool nan0[CS];
  /* node 0 seeds the maximum (src/TreeScorer.hpp:187-188) */
  auto first = [&](const uint32_t (&A)[4]) {
#pragma unroll
    for (int j = 0; j < CS; j++) {
      uint32_t n11 = 0, n01 = 0;
#pragma unroll
      for (int w = 0; w < W; w++) {
        n11 = bcnt_acc(c.km[j][w] & A[w], n11);
        n01 = bcnt_acc(c.knm[j][w] & A[w], n01);
      }
      float s = pair_score(n01, n11, c.fa[j], c.fb[j], L);
      best[j] = s;
      nan0[j] = s != s;
      if (COUNTS) {
        arg01[j] = n01;
        arg11[j] = n11;
      }
    }
  };
  /* later nodes replace it only when strictly greater (:189-190) */
  auto rest = [&](const uint32_t (&A)[4]) {
#pragma unroll
    for (int j = 0; j < CS; j++) {
      uint32_t n11 = 0, n01 = 0;
#pragma unroll
      for (int w = 0; w < W; w++) {
        n11 = bcnt_acc(c.km[j][w] & A[w], n11);
        n01 = bcnt_acc(c.knm[j][w] & A[w], n01);
      }
      float s = pair_score(n01, n11, c.fa[j], c.fb[j], L);
      if (COUNTS) {
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
  auto words = [&](uint32_t k, uint32_t lane, uint32_t (&A)[4]) {
    A[0] = readlane32((uint32_t)dm[k][0], lane);
    A[1] = W > 1 ? readlane32((uint32_t)(dm[k][0] >> 32), lane) : 0u;
    A[2] = W > 2 ? readlane32((uint32_t)dm[k][NS - 1], lane) : 0u;
    A[3] = W > 3 ? readlane32((uint32_t)(dm[k][NS - 1] >> 32), lane) : 0u;
  };
  {
    uint32_t A[4];
    words(0u, 0u, A);
    first(A);
  }
  const uint32_t n_lo = n_nodes < 64u ? n_nodes : 64u;
  for (uint32_t v = 1; v < n_lo; v++) {
    uint32_t A[4];
    words(0u, v, A);
    rest(A);
  }
  if (NS > 1) {
    for (uint32_t v = 64u; v < n_nodes; v++) {
      uint32_t A[4];
      words((uint32_t)NS - 1u, v - 64u, A);
      rest(A);
    }
  }
#pragma unroll
  for (int j = 0; j < CS; j++)
    lds[lane_id() + 64u * (uint32_t)j] = nan0[j] ? __int_as_float(0x7fc00000) : best[j];
  wave_lds_sync();
  float T = 0.0f;
  uint32_t cc = 0;
  for (; cc + 4u <= n_cells; cc += 4u) {
    float4 q = *reinterpret_cast<const float4 *>(lds + cc);
    T = T + q.x;
    T = T + q.y;
    T = T + q.z;
    T = T + q.w;
  }
  for (; cc < n_cells; cc++) T = T + lds[cc];
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
