/*
 * Bit-exact device counterparts of the three libm calls that feed the chain's
 * decisions in the reference:
 *   expf(float)            - Metropolis acceptance (src/Application.hpp:452-454)
 *   logf(float)            - log(beta), the beta prior (src/TreeScorer.hpp:124,136)
 *   (float)log(double)     - log(1.0 - beta) (src/TreeScorer.hpp:100,106,137)
 *
 * The reference is compiled for the host, so its values are glibc's.
 * bbs_expf and bbs_logf restate the algorithms glibc 2.35 ships for expf and
 * logf (ARM Optimized Routines, MIT OR Apache-2.0 WITH LLVM-exception:
 * exp2-table of 32 entries with a cubic, log-table of 16 entries with a
 * cubic, both evaluated in double). They match glibc on every float input
 * (exhaustive comparison: oracle/libm_exhaustive.c).
 *
 * bbs_logf_of_double rounds an accurate double-double log first to double
 * and then to float, as (float)glibc_log(x) does; the two agree for every
 * x = 1 - b with b a float in [0, 1] (same exhaustive program).
 *
 * Plain C99 so that host tests and the HIP kernels compile the same code.
 */
#ifndef BBS_LIBM_H
#define BBS_LIBM_H

#include "bbs_rng.h"

BBS_HD float bbs_f32_from_bits(uint32_t u) {
  union {
    uint32_t u;
    float f;
  } c;
  c.u = u;
  return c.f;
}

BBS_HD uint32_t bbs_f32_bits(float f) {
  union {
    uint32_t u;
    float f;
  } c;
  c.f = f;
  return c.u;
}

/* ---------------------------------------------------------------- expf */

/* 2^(i/32) as double bits, minus i << 47 so that adding ki << 47 yields the
 * scaled value (the table layout of the published algorithm). */
BBS_TABLE uint64_t bbs_exp2f_table[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full,
    0x3fef9301d0125b51ull, 0x3fef72b83c7d517bull, 0x3fef54873168b9aaull,
    0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull, 0x3fef06fe0a31b715ull,
    0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull,
    0x3feea47eb03a5585ull, 0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull,
    0x3feea11473eb0187ull, 0x3feea589994cce13ull, 0x3feeace5422aa0dbull,
    0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull,
    0x3fef3720dcef9069ull, 0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full,
    0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull};

/* (1/c, log(c)) for the 16 subintervals of the published logf algorithm. */
BBS_TABLE double bbs_logf_invc[16] = {
    0x1.661ec79f8f3bep+0, 0x1.571ed4aaf883dp+0, 0x1.49539f0f010bp+0,
    0x1.3c995b0b80385p+0, 0x1.30d190c8864a5p+0, 0x1.25e227b0b8eap+0,
    0x1.1bb4a4a1a343fp+0, 0x1.12358f08ae5bap+0, 0x1.0953f419900a7p+0,
    0x1p+0,               0x1.e608cfd9a47acp-1, 0x1.ca4b31f026aap-1,
    0x1.b2036576afce6p-1, 0x1.9c2d163a1aa2dp-1, 0x1.886e6037841edp-1,
    0x1.767dcf5534862p-1};
BBS_TABLE double bbs_logf_logc[16] = {
    -0x1.57bf7808caadep-2, -0x1.2bef0a7c06ddbp-2, -0x1.01eae7f513a67p-2,
    -0x1.b31d8a68224e9p-3, -0x1.6574f0ac07758p-3, -0x1.1aa2bc79c81p-3,
    -0x1.a4e76ce8c0e5ep-4, -0x1.1973c5a611cccp-4, -0x1.252f438e10c1ep-5,
    0x0p+0,                0x1.aa5aa5df25984p-5,  0x1.c5e53aa362eb4p-4,
    0x1.526e57720db08p-3,  0x1.bc2860d22477p-3,   0x1.1058bc8a07ee1p-2,
    0x1.4043057b6ee09p-2};

/* expf in two halves around its one table read, so that a caller may keep
 * the table elsewhere (the chain kernel holds it across the lanes of a
 * register pair). bbs_expf_head returns 1 with *special = expf(x) for
 * |x| >= 88, inf and nan; otherwise 0 with the reduction (r, ki), whose
 * table entry is bbs_exp2f_table[ki % 32]. bbs_expf_tail finishes from
 * that entry. */
BBS_HD int bbs_expf_head(float x, float *special, double *r_out, uint64_t *ki_out) {
  const double INV_LN2_N = 0x1.71547652b82fep+0 * 32;
  const double SHIFT = 0x1.8p+52;
  uint32_t ix = bbs_f32_bits(x);
  uint32_t abstop = (ix >> 20) & 0x7ffu;
  if (abstop >= 0x42bu) { /* |x| >= 88 or x is inf/nan */
    if (ix == 0xff800000u) {
      *special = 0.0f;
      return 1;
    }
    if (abstop >= 0x7f8u) {
      *special = x + x;
      return 1;
    }
    if (x > 0x1.62e42ep6f) {
      *special = bbs_f32_from_bits(0x7f800000u);
      return 1;
    }
    if (x < -0x1.9fe368p6f) {
      *special = 0.0f;
      return 1;
    }
  }
  double xd = (double)x;
  double kd = fma(INV_LN2_N, xd, SHIFT);
  uint64_t ki = bbs_f64_bits(kd);
  kd -= SHIFT;
  *r_out = fma(INV_LN2_N, xd, -kd);
  *ki_out = ki;
  return 0;
}

BBS_HD float bbs_expf_tail(double r, uint64_t ki, uint64_t entry) {
  const double C0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32;
  const double C1 = 0x1.ebfce50fac4f3p-3 / 32 / 32;
  const double C2 = 0x1.62e42ff0c52d6p-1 / 32;
  uint64_t t = entry;
  t += ki << (52 - 5);
  double s = bbs_f64_from_bits(t);
  double z = fma(C0, r, C1);
  double r2 = r * r;
  double y = fma(C2, r, 1.0);
  y = fma(z, r2, y);
  y = y * s;
  return (float)y;
}

BBS_HD float bbs_expf(float x) {
  float special;
  double r;
  uint64_t ki;
  if (bbs_expf_head(x, &special, &r, &ki)) return special;
  return bbs_expf_tail(r, ki, bbs_exp2f_table[ki % 32u]);
}

/* ---------------------------------------------------------------- logf */

/* A0, A1, A2 of the cubic and ln 2 (the chain loop reads them at
 * bbs_coef_base(): there bbs_logf runs on beta moves only). */
BBS_TABLE double bbs_logf_k[4] = {-0x1.00ea348b88334p-2, 0x1.5575b0be00b6ap-2,
                                  -0x1.ffffef20a4123p-2, 0x1.62e42fefa39efp-1};

BBS_HD float bbs_logf_at(float x, uint32_t kb) {
  const double A0 = bbs_logf_k[kb], A1 = bbs_logf_k[kb + 1u], A2 = bbs_logf_k[kb + 2u],
               LN2 = bbs_logf_k[kb + 3u];
  const uint32_t OFF = 0x3f330000u;
  uint32_t ix = bbs_f32_bits(x);
  if (ix == 0x3f800000u) return 0.0f;
  if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
    /* x < 0x1p-126 or inf or nan. */
    if (ix * 2u == 0u) return bbs_f32_from_bits(0xff800000u); /* -inf */
    if (ix == 0x7f800000u) return x;                           /* +inf */
    if ((ix & 0x80000000u) || ix * 2u >= 0xff000000u)
      return bbs_f32_from_bits(0x7fc00000u); /* nan */
    /* subnormal: normalize */
    float xs = x * 0x1p23f;
    ix = bbs_f32_bits(xs);
    ix -= 23u << 23;
  }
  uint32_t tmp = ix - OFF;
  uint32_t i = (tmp >> (23 - 4)) % 16u;
  int32_t k = (int32_t)tmp >> 23;
  uint32_t iz = ix - (tmp & 0xff800000u);
  double invc = bbs_logf_invc[i], logc = bbs_logf_logc[i];
  double z = (double)bbs_f32_from_bits(iz);
  double r = fma(z, invc, -1.0);
  double y0 = fma((double)k, LN2, logc);
  double r2 = r * r;
  double y = fma(A1, r, A2);
  y = fma(A0, r2, y);
  y = fma(y, r2, y0 + r);
  return (float)y;
}

BBS_HD float bbs_logf(float x) { return bbs_logf_at(x, 0u); }

/* ------------------------------------------------- (float)log(double) */

/* (float)log(x) for double x >= 0: log rounded to double, then to float,
 * the two roundings the reference's (float)std::log(double) performs. */
BBS_HD float bbs_logf_of_double_at(double x, uint32_t kb) {
  if (x == 0.0) return bbs_f32_from_bits(0xff800000u);
  if (x == 1.0) return 0.0f;
  if (!(x > 0.0)) return bbs_f32_from_bits(0x7fc00000u);
  if (x > 0x1.fffffffffffffp+1023) return bbs_f32_from_bits(0x7f800000u);
  double lo;
  double hi = bbs_log_dd_at(x, &lo, kb);
  return (float)hi;
}

BBS_HD float bbs_logf_of_double(double x) { return bbs_logf_of_double_at(x, 0u); }

#endif /* BBS_LIBM_H */
