// sampler.cpp -- Sobol low-discrepancy points with optional Owen-type scrambling,
// the sample inputs of the Monte Carlo runs.
//
// The reference reads its sample inputs from a file of random numbers
// (src/fv.cc:502-551, readme/README.md:440-452) and has no generator; this
// sampler produces such inputs on the host.
//
// Sobol points (Bratley & Fox, ACM TOMS 14 (1988) 88-100) with the direction
// numbers of Joe & Kuo, "Constructing Sobol sequences with better
// two-dimensional projections", SIAM J. Sci. Comput. 30 (2008) 2635-2654
// (table new-joe-kuo-6.21201), in Gray-code order (Antonov & Saleev 1979):
// point i is the XOR of the direction numbers selected by the bits of
// i ^ (i >> 1).  Scrambling is the nested uniform (Owen) scramble computed with
// the hash of Burley, "Practical Hash-based Owen Scrambling", JCGT 9(4) 2020:
// bit-reverse, Laine-Karras permutation keyed by a per-dimension seed,
// bit-reverse.  Each output bit then depends only on the bits above it, which
// keeps every (t,m,s)-net property of the unscrambled points.
#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/tsm.h"

namespace {

struct Dir {
  int s, a;
  unsigned m[8];
};

// Joe-Kuo new-joe-kuo-6.21201, dimensions 2..16 (dimension 1 is van der Corput).
const Dir kDir[] = {
    {1, 0, {1}},
    {2, 1, {1, 3}},
    {3, 1, {1, 3, 1}},
    {3, 2, {1, 1, 1}},
    {4, 1, {1, 1, 3, 3}},
    {4, 4, {1, 3, 5, 13}},
    {5, 2, {1, 1, 5, 5, 17}},
    {5, 4, {1, 1, 5, 5, 5}},
    {5, 7, {1, 1, 7, 11, 19}},
    {5, 11, {1, 1, 5, 1, 1}},
    {5, 13, {1, 1, 1, 3, 11}},
    {5, 14, {1, 3, 5, 5, 31}},
    {6, 1, {1, 3, 3, 9, 7, 49}},
    {6, 13, {1, 1, 1, 15, 21, 21}},
    {6, 16, {1, 3, 1, 13, 27, 49}},
};
constexpr int kMaxDim = 1 + (int)(sizeof(kDir) / sizeof(kDir[0]));

void directions(int d, uint32_t V[32]) {
  if (d == 0) {
    for (int k = 0; k < 32; ++k) V[k] = 1u << (31 - k);
    return;
  }
  const Dir& D = kDir[d - 1];
  for (int k = 0; k < D.s; ++k) V[k] = D.m[k] << (31 - k);
  for (int k = D.s; k < 32; ++k) {
    uint32_t v = V[k - D.s] ^ (V[k - D.s] >> D.s);
    for (int i = 1; i < D.s; ++i)
      if ((D.a >> (D.s - 1 - i)) & 1) v ^= V[k - i];
    V[k] = v;
  }
}

uint32_t reverse_bits(uint32_t x) {
  x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
  x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
  x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
  x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
  return (x >> 16) | (x << 16);
}

// Laine-Karras style permutation (Burley 2020, listing 2): only carries from
// lower to higher bits, so on bit-reversed input it is a nested scramble.
uint32_t lk_permute(uint32_t x, uint32_t seed) {
  x += seed;
  x ^= x * 0x6c50b47cu;
  x ^= x * 0xb82f1e52u;
  x ^= x * 0xc7afe638u;
  x ^= x * 0x8d22f6e6u;
  return x;
}

uint32_t owen_scramble(uint32_t x, uint32_t seed) {
  return reverse_bits(lk_permute(reverse_bits(x), seed));
}

uint32_t dim_seed(unsigned long long seed, int d) {  // splitmix64 finaliser of (seed, d)
  uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(d + 1);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return (uint32_t)((z ^ (z >> 31)) >> 32);
}

}  // namespace

extern "C" int tsm_sobol(long n, int dim, long skip, unsigned long long seed, int scramble,
                         double* out) {
  if (n < 0 || skip < 0 || dim < 1 || dim > kMaxDim || !out || n + skip > (1l << 32))
    return TSM_E_ARG;
  for (int d = 0; d < dim; ++d) {
    uint32_t V[32];
    directions(d, V);
    const uint32_t sd = dim_seed(seed, d);
    for (long i = 0; i < n; ++i) {
      const uint64_t idx = (uint64_t)(i + skip);
      const uint64_t g = idx ^ (idx >> 1);
      uint32_t x = 0;
      for (int k = 0; k < 32; ++k)
        if ((g >> k) & 1) x ^= V[k];
      if (scramble) x = owen_scramble(x, sd);
      // scrambled points sit at the centre of their 2^-32 cell, never at 0 or 1
      out[(size_t)i * dim + d] = scramble ? ((double)x + 0.5) * 0x1p-32 : (double)x * 0x1p-32;
    }
  }
  return TSM_OK;
}

extern "C" int tsm_sobol_max_dim(void) { return kMaxDim; }
