/*
 * Exhaustive check (test infrastructure, CPU only) that the restated libm
 * functions of bbscite_amd/csrc/bbs_libm.h equal the host glibc on every
 * input the chain can feed them:
 *   bbs_expf(x)              == expf(x)                 for all 2^32 floats
 *   bbs_logf(x)              == logf(x)                 for all 2^32 floats
 *   bbs_logf_of_double(1-b)  == (float)log(1.0 - b)     for every float b in [0,1]
 * The reference evaluates expf at src/Application.hpp:452-454, logf at
 * src/TreeScorer.hpp:124,136 and (float)log(1.0 - beta) at
 * src/TreeScorer.hpp:106,137.
 *
 * Usage: libm_exhaustive [stride]   (stride 1 = every input; larger strides
 * sample every stride-th bit pattern). Prints one line per function with the
 * number of inputs and mismatches; exit status 1 on any mismatch.
 */
#include "../bbscite_amd/csrc/bbs_libm.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int same_float(float a, float b) {
  if (a != a && b != b) return 1; /* both NaN */
  return memcmp(&a, &b, 4) == 0;
}

int main(int argc, char **argv) {
  uint64_t stride = argc > 1 ? strtoull(argv[1], NULL, 10) : 1;
  if (stride == 0) stride = 1;
  long n_exp = 0, bad_exp = 0, n_log = 0, bad_log = 0, n_dlog = 0,
       bad_dlog = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : n_exp, bad_exp, n_log, bad_log)
  for (uint64_t u = 0; u < 0x100000000ull; u += stride) {
    float x;
    uint32_t b = (uint32_t)u;
    memcpy(&x, &b, 4);
    n_exp++;
    if (!same_float(bbs_expf(x), expf(x))) {
      bad_exp++;
      if (bad_exp < 5) fprintf(stderr, "expf mismatch at %08x\n", b);
    }
    n_log++;
    if (!same_float(bbs_logf(x), logf(x))) {
      bad_log++;
      if (bad_log < 5) fprintf(stderr, "logf mismatch at %08x\n", b);
    }
  }
  /* floats in [0, 1]: bit patterns 0 .. 0x3f800000 */
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : n_dlog, bad_dlog)
  for (uint64_t u = 0; u <= 0x3f800000ull; u += stride) {
    float beta;
    uint32_t b = (uint32_t)u;
    memcpy(&beta, &b, 4);
    double x = 1.0 - (double)beta;
    n_dlog++;
    if (!same_float(bbs_logf_of_double(x), (float)log(x))) {
      bad_dlog++;
      fprintf(stderr, "log(1-b) mismatch at b=%08x\n", b);
    }
  }
  printf("expf %ld %ld\n", n_exp, bad_exp);
  printf("logf %ld %ld\n", n_log, bad_log);
  printf("log1mb %ld %ld\n", n_dlog, bad_dlog);
  return (bad_exp || bad_log || bad_dlog) ? 1 : 0;
}
