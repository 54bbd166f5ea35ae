/*
 * Counter-based random stream of the chain search (host and device).
 *
 * Every random draw of a chain step is a pure function of
 * (seed, job, chain, step, purpose): Philox4x32-10 (Salmon, Moraes, Dror,
 * Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11) keyed by the
 * 64-bit seed, with the 128-bit counter (index, chain, job, purpose).
 * Results therefore do not depend on where or in which order chains run.
 *
 * The reference draws the same quantities from std::minstd_rand through
 * libstdc++ distributions on the host (src/RawMoveDistribution.hpp:97-155,
 * src/Application.hpp:146-178); that stream is seeded from random_device and
 * is not reproducible, so only the draw *semantics* are kept:
 *   move type  : U[0,1) compared with cumulative normalised probabilities
 *                (src/RawMoveDistribution.hpp:119-131)
 *   v, w       : v ~ U{0..N-2}, w ~ U{0..N-3}, w++ if w >= v (:146-155)
 *   u_desc, u_nondesc, u_accept : U[0,1) floats (:105-108)
 *   beta_jump  : Normal(0, sd/chi) (:107, src/Parameters.hpp:582-584)
 *
 * This header is plain C99 so that the CPU oracle and the HIP kernels
 * compile the very same code; every floating-point step uses only IEEE
 * operations (+, -, *, fma, conversions) so host and device agree bit for bit.
 */
#ifndef BBS_RNG_H
#define BBS_RNG_H

#include <stdint.h>
#if !defined(__HIPCC__)
#include <math.h>
#endif

#if defined(__HIPCC__)
#define BBS_HD __host__ __device__ __forceinline__
#elif defined(__cplusplus)
#define BBS_HD inline
#else
#define BBS_HD static inline
#endif

/* Lookup tables: constant memory in the device pass, static data on the host. */
#if defined(__HIP_DEVICE_COMPILE__)
#define BBS_TABLE static __constant__ const
#else
#define BBS_TABLE static const
#endif

/*
 * Index base of the coefficient tables below: zero, but opaque to the device
 * compiler. Double-precision coefficients written as literals cannot be
 * instruction operands, so the compiler materialises each in a VGPR pair and
 * hoists it out of the chain loop, where the coefficients of the beta-move
 * code (used on about a tenth of the steps) held some 60 VGPRs for the whole
 * launch. Read from a table at this base, they are scalar loads issued where
 * they are used. The functions taking an index base `kb` are called with
 * bbs_coef_base() inside the chain loop and with a literal 0 elsewhere (the
 * host, and kernels that use them once per item), where the compiler folds
 * the table reads back into constants. Both read the same values.
 */
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ uint32_t bbs_coef_base(void) {
  uint32_t z;
  asm volatile("s_mov_b32 %0, 0" : "=s"(z));
  return z;
}
#else
BBS_HD uint32_t bbs_coef_base(void) { return 0u; }
#endif

/* Counter word 3: what a Philox block is used for. */
enum {
  BBS_PURPOSE_MOVE_A = 0u,   /* move type, v, w, u_desc            */
  BBS_PURPOSE_MOVE_B = 1u,   /* u_nondesc, normal pair, u_accept   */
  BBS_PURPOSE_INIT_TREE = 2u,/* Pruefer code of the initial tree  */
  BBS_PURPOSE_DATA_TREE = 3u,/* synthetic data: true tree           */
  BBS_PURPOSE_DATA_CELL = 4u,/* synthetic data: cell attachment     */
  BBS_PURPOSE_DATA_NOISE = 5u,/* synthetic data: noise, 4 draws/entry */
  BBS_PURPOSE_TREE_SAMPLE = 6u,/* bulk random trees for scoring      */
  BBS_PURPOSE_DATA_SHAPE = 7u /* dataset mixes: per-dataset sizes    */
};

typedef struct {
  uint32_t x[4];
} bbs_u32x4;

BBS_HD uint32_t bbs_mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

/* Philox4x32 with 10 rounds. */
BBS_HD bbs_u32x4 bbs_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                   uint32_t c3, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; r++) {
    uint32_t hi0 = bbs_mulhi32(M0, c0), lo0 = M0 * c0;
    uint32_t hi1 = bbs_mulhi32(M1, c2), lo1 = M1 * c2;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  bbs_u32x4 out;
  out.x[0] = c0;
  out.x[1] = c1;
  out.x[2] = c2;
  out.x[3] = c3;
  return out;
}

BBS_HD bbs_u32x4 bbs_philox_block(uint64_t seed, uint32_t job, uint32_t chain,
                                  uint32_t index, uint32_t purpose) {
  return bbs_philox4x32_10(index, chain, job, purpose, (uint32_t)seed,
                           (uint32_t)(seed >> 32));
}

/* U[0,1) float with 24 random bits; exact, never 1.0. */
BBS_HD float bbs_u01f(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

/* Uniform integer in [0, n) by multiply-shift (bias <= n / 2^32). */
BBS_HD uint32_t bbs_uint_below(uint32_t x, uint32_t n) {
  return (uint32_t)(((uint64_t)x * (uint64_t)n) >> 32);
}

/* ---- deterministic double helpers for the normal draw ---------------- */

BBS_HD double bbs_f64_from_bits(uint64_t u) {
  union {
    uint64_t u;
    double d;
  } c;
  c.u = u;
  return c.d;
}

BBS_HD uint64_t bbs_f64_bits(double d) {
  union {
    uint64_t u;
    double d;
  } c;
  c.d = d;
  return c.u;
}

/* ln 2 as a double-double, sqrt(2), then the atanh series coefficients
 * 1/25, 1/23, ..., 1/3 of bbs_log_dd, highest first. */
BBS_TABLE double bbs_log_dd_k[15] = {
    0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56, 0x1.6a09e667f3bcdp+0,
    1.0 / 25.0, 1.0 / 23.0, 1.0 / 21.0, 1.0 / 19.0, 1.0 / 17.0,
    1.0 / 15.0, 1.0 / 13.0, 1.0 / 11.0, 1.0 / 9.0, 1.0 / 7.0,
    1.0 / 5.0, 1.0 / 3.0};

/*
 * Natural log of x in (0, inf) as a double-double (hi + lo), from
 * log(x) = e*ln2 + 2*atanh(s), s = (m-1)/(m+1), m in [sqrt(1/2), sqrt(2)).
 * Relative error below 2^-58. Only IEEE +,-,*,/ and fma are used.
 * Subnormal inputs are scaled by 2^54 first. kb: the coefficients' index
 * base, 0 or bbs_coef_base() (the _at functions below all take one).
 */
BBS_HD double bbs_log_dd_at(double x, double *lo_out, uint32_t kb) {
  const double LN2_HI = bbs_log_dd_k[kb];
  const double LN2_LO = bbs_log_dd_k[kb + 1u];
  uint64_t ix = bbs_f64_bits(x);
  int e = 0;
  if ((ix >> 52) == 0) { /* subnormal */
    x = x * 0x1p54;
    ix = bbs_f64_bits(x);
    e = -54;
  }
  e += (int)((ix >> 52) & 0x7ff) - 1023;
  double m = bbs_f64_from_bits((ix & 0x000fffffffffffffull) |
                               0x3ff0000000000000ull); /* [1,2) */
  if (m > bbs_log_dd_k[kb + 2u]) {                     /* > sqrt(2) */
    m = m * 0.5;
    e += 1;
  }
  double f = m - 1.0; /* exact (Sterbenz) */
  /* d = 2 + f as double-double */
  double d_hi = 2.0 + f;
  double d_lo = f - (d_hi - 2.0);
  /* s = f / d */
  double s_hi = f / d_hi;
  double rem = fma(-s_hi, d_hi, f) - s_hi * d_lo;
  double s_lo = rem / d_hi;
  /* tail: 2 s^3 (1/3 + s^2/5 + s^4/7 + ...) */
  double z = s_hi * s_hi;
  double p = bbs_log_dd_k[kb + 3u];
  for (uint32_t i = 4u; i < 15u; i++) p = fma(p, z, bbs_log_dd_k[kb + i]);
  double tail = 2.0 * (s_hi * z) * p;
  /* log(m) = 2 s_hi + (2 s_lo + tail) */
  double lm_hi = 2.0 * s_hi;
  double lm_lo = 2.0 * s_lo + tail;
  /* e * ln2 as double-double */
  double de = (double)e;
  double el_hi = de * LN2_HI;
  double el_lo = fma(de, LN2_HI, -el_hi) + de * LN2_LO;
  /* two-sum of the high parts */
  double sh = el_hi + lm_hi;
  double bb = sh - el_hi;
  double err = (el_hi - (sh - bb)) + (lm_hi - bb);
  double lo = err + el_lo + lm_lo;
  double hi = sh + lo;
  lo = lo - (hi - sh);
  *lo_out = lo;
  return hi;
}

BBS_HD double bbs_log_dd(double x, double *lo_out) { return bbs_log_dd_at(x, lo_out, 0u); }

/* sqrt(x) for x >= 0 from Newton steps on 1/sqrt(x); within 1 ulp. */
BBS_HD double bbs_sqrt_det(double x) {
  if (!(x > 0.0)) return 0.0;
  uint64_t ix = bbs_f64_bits(x);
  int e = (int)((ix >> 52) & 0x7ff) - 1023;
  /* scale x into [1,4) with an even exponent shift */
  int half = e >> 1;
  double m = x * bbs_f64_from_bits((uint64_t)(1023 - 2 * half) << 52);
  double y = 1.0 / (1.0 + 0.5 * (m - 1.0)); /* rough 1/sqrt(m) */
  for (int i = 0; i < 6; i++) {
    double t = y * y;
    y = y * fma(-0.5 * m, t, 1.5);
  }
  double s = m * y;
  s = fma(0.5 * y, fma(-s, s, m), s);
  return s * bbs_f64_from_bits((uint64_t)(1023 + half) << 52);
}

/* 2 pi, then the Taylor coefficients 1/24!, 1/22!, ..., 1/2!, 1 of cos and
 * 1/25!, 1/23!, ..., 1/3!, 1 of sin, highest first. */
BBS_TABLE double bbs_cos_sin_k[27] = {
    0x1.921fb54442d18p+2,
    1.0 / 620448401733239439360000.0, 1.0 / 1124000727777607680000.0,
    1.0 / 2432902008176640000.0, 1.0 / 6402373705728000.0,
    1.0 / 20922789888000.0, 1.0 / 87178291200.0, 1.0 / 479001600.0,
    1.0 / 3628800.0, 1.0 / 40320.0, 1.0 / 720.0, 1.0 / 24.0, 1.0 / 2.0, 1.0,
    1.0 / 15511210043330985984000000.0, 1.0 / 25852016738884976640000.0,
    1.0 / 51090942171709440000.0, 1.0 / 121645100408832000.0,
    1.0 / 355687428096000.0, 1.0 / 1307674368000.0, 1.0 / 6227020800.0,
    1.0 / 39916800.0, 1.0 / 362880.0, 1.0 / 5040.0, 1.0 / 120.0, 1.0 / 6.0, 1.0};

/* cos(2*pi*k/2^32) for a 32-bit k. */
BBS_HD double bbs_cos_2pi_u32_at(uint32_t k, uint32_t kb) {
  const double TWO_PI = bbs_cos_sin_k[kb];
  uint32_t q = k >> 30;
  double f = (double)(k & 0x3fffffffu) * 0x1p-32; /* [0, 1/4) exact */
  double t = TWO_PI * f;                          /* [0, pi/2) */
  double t2 = t * t;
  /* Taylor series of cos and sin to t^24 / t^25 (error < 1e-19) */
  double c = bbs_cos_sin_k[kb + 1u];
  for (uint32_t i = 2u; i < 14u; i++) c = fma(c, -t2, bbs_cos_sin_k[kb + i]);
  double s = bbs_cos_sin_k[kb + 14u];
  for (uint32_t i = 15u; i < 27u; i++) s = fma(s, -t2, bbs_cos_sin_k[kb + i]);
  s = s * t;
  switch (q) {
  case 0:
    return c;
  case 1:
    return -s;
  case 2:
    return -c;
  default:
    return s;
  }
}

BBS_HD double bbs_cos_2pi_u32(uint32_t k) { return bbs_cos_2pi_u32_at(k, 0u); }

/* Standard normal by Box-Muller from two 32-bit words. */
BBS_HD double bbs_normal_from_u32_at(uint32_t a, uint32_t b, uint32_t kb) {
  double u1 = ((double)a + 1.0) * 0x1p-32; /* (0, 1] */
  double lo;
  double l = bbs_log_dd_at(u1, &lo, kb);
  double r = bbs_sqrt_det(-2.0 * l);
  return r * bbs_cos_2pi_u32_at(b, kb);
}

BBS_HD double bbs_normal_from_u32(uint32_t a, uint32_t b) {
  return bbs_normal_from_u32_at(a, b, 0u);
}

/*
 * The raw move of one chain step, the counterpart of RawMoveSample
 * (src/RawMoveDistribution.hpp:44-50) with the type already resolved.
 */
typedef struct {
  uint32_t move_type; /* 0 ChangeBeta, 1 PruneReattach, 2 SwapNodes, 3 SwapSubtrees
                         (src/MoveType.hpp:23-42) */
  uint32_t raw_v, raw_w;
  float raw_descendant_of_v, raw_nondescendant_of_v;
  float beta_jump;
  float acceptance_level;
} bbs_raw_move;

/*
 * Cumulative move-type thresholds in float, as the reference compares them
 * (src/RawMoveDistribution.hpp:119-131): p_beta, p_beta+p_prune,
 * p_beta+p_prune+p_swap.
 */
typedef struct {
  float t_beta, t_prune, t_swap;
  float beta_jump_sd;
} bbs_move_probs;

BBS_HD uint32_t bbs_move_type_of(float u, bbs_move_probs p) {
  if (u <= p.t_beta) return 0u;
  if (u <= p.t_prune) return 1u;
  if (u <= p.t_swap) return 2u;
  return 3u;
}

/*
 * Draw the raw move of (job, chain, step) for a tree of n_nodes nodes.
 * Block A = (type, v, w, u_desc), block B = (u_nondesc, normal a, normal b,
 * u_accept). The normal is evaluated only for beta-change moves: the jump
 * does not influence any other move (src/MutationTree.hpp:128-130).
 */
/* The beta jump of a beta-change move whose block B is b. */
BBS_HD float bbs_beta_jump_at(bbs_u32x4 b, bbs_move_probs p, uint32_t kb) {
  return (float)(bbs_normal_from_u32_at(b.x[1], b.x[2], kb) * (double)p.beta_jump_sd);
}

BBS_HD float bbs_beta_jump(bbs_u32x4 b, bbs_move_probs p) { return bbs_beta_jump_at(b, p, 0u); }

/* Everything of bbs_move_from_blocks but the beta jump (left 0). */
BBS_HD bbs_raw_move bbs_move_without_jump(bbs_u32x4 a, bbs_u32x4 b, uint32_t n_nodes,
                                          bbs_move_probs p) {
  bbs_raw_move m;
  m.move_type = bbs_move_type_of(bbs_u01f(a.x[0]), p);
  m.raw_v = bbs_uint_below(a.x[1], n_nodes - 1u);
  uint32_t w = bbs_uint_below(a.x[2], n_nodes - 2u);
  if (w >= m.raw_v) w++;
  m.raw_w = w;
  m.raw_descendant_of_v = bbs_u01f(a.x[3]);
  m.raw_nondescendant_of_v = bbs_u01f(b.x[0]);
  m.beta_jump = 0.0f;
  m.acceptance_level = bbs_u01f(b.x[3]);
  return m;
}

BBS_HD bbs_raw_move bbs_move_from_blocks(bbs_u32x4 a, bbs_u32x4 b, uint32_t n_nodes,
                                         bbs_move_probs p) {
  bbs_raw_move m = bbs_move_without_jump(a, b, n_nodes, p);
  if (m.move_type == 0u) m.beta_jump = bbs_beta_jump(b, p);
  return m;
}

BBS_HD bbs_raw_move bbs_draw_move(uint64_t seed, uint32_t job, uint32_t chain,
                                  uint32_t step, uint32_t n_nodes,
                                  bbs_move_probs p) {
  return bbs_move_from_blocks(bbs_philox_block(seed, job, chain, step, BBS_PURPOSE_MOVE_A),
                              bbs_philox_block(seed, job, chain, step, BBS_PURPOSE_MOVE_B),
                              n_nodes, p);
}

#endif /* BBS_RNG_H */
