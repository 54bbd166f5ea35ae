/*
 * Independent Philox4x32-10 (test infrastructure, CPU only): prints the
 * block that ROCm's rocRAND engine (rocrand/rocrand_philox4x32_10.h, host
 * path) produces for each (counter, key) read from stdin, one per line as
 * six hex words "c0 c1 c2 c3 k0 k1". The engine seeded with key (k0, k1),
 * subsequence c2 + 2^32 c3 and offset 4 (c0 + 2^32 c1) holds exactly the
 * block Philox(counter, key) as its next four outputs.
 */
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

int main() {
  unsigned c[4], k[2];
  while (std::scanf("%x %x %x %x %x %x", &c[0], &c[1], &c[2], &c[3], &k[0], &k[1]) == 6) {
    const unsigned long long seed = (unsigned long long)k[0] | ((unsigned long long)k[1] << 32);
    const unsigned long long sub = (unsigned long long)c[2] | ((unsigned long long)c[3] << 32);
    const unsigned long long off = 4ull * ((unsigned long long)c[0] | ((unsigned long long)c[1] << 32));
    rocrand_device::philox4x32_10_engine e(seed, sub, off);
    const uint4 r = e.next4();
    std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
  }
  return 0;
}
