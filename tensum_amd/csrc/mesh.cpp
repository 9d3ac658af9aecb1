// mesh.cpp -- seeded jittered-triangle lattice generator of libtsm.
//
// The reference ships no mesh generator (it reads gmsh-partitioned part files,
// readme/README.md:137-201).  This one writes the NPART = 1 part-file content
// directly:
//   * vertex (i, j) of the n x n lattice has gid j*n + i and base position
//     (x0 + i*hx, y0 + j*hy), the last column/row pinned to x1/y1;
//   * random stream: splitmix64 seeded with `seed`, u = (next >> 11) * 2^-53;
//   * interior vertices, in gid order, draw ux then uy and move by
//     (jitter*(2u-1))*h in each direction; boundary vertices do not move, so
//     opposite sides of a periodic box coincide;
//   * square (i, j), in row-major order, draws u: u < 0.5 splits along
//     (i,j)-(i+1,j+1) into (a,b,c),(a,c,d), otherwise along (i+1,j)-(i,j+1)
//     into (a,b,d),(b,c,d); a=(i,j) b=(i+1,j) c=(i+1,j+1) d=(i,j+1), CCW;
//   * boundary edges: bottom (tag 1, left to right), right (2, upwards),
//     top (3, right to left), left (4, downwards);
//   * periodic sets (periodic = 1): the four corners, then the left/right
//     pairs for j = 1..n-2, then the bottom/top pairs for i = 1..n-2.
#include <cstdint>

#include "tsm_host.h"

namespace {
inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
inline double u01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }
}  // namespace

extern "C" int tsm_mesh_jittered_sizes(int n, int periodic, int* nv, int* nb, int* nc,
                                       int* np, int* npv) {
  if (n < 3) return TSM_E_ARG;
  *nv = n * n;
  *nb = 4 * (n - 1);
  *nc = 2 * (n - 1) * (n - 1);
  *np = periodic ? 2 * (n - 2) + 1 : 0;
  *npv = periodic ? 4 + 4 * (n - 2) : 0;
  return TSM_OK;
}

extern "C" int tsm_mesh_jittered(int n, double x0, double x1, double y0, double y1,
                                 double jitter, unsigned long long seed, int periodic,
                                 double* x, double* y, int* btag, int* bv, int* cv,
                                 int* pptr, int* pv) {
  if (n < 3 || !(x1 > x0) || !(y1 > y0) || jitter < 0.0 || jitter >= 0.5) return TSM_E_ARG;
  const double hx = (x1 - x0) / (n - 1), hy = (y1 - y0) / (n - 1);
  uint64_t s = seed;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      int g = j * n + i;
      double px = (i == n - 1) ? x1 : x0 + i * hx;
      double py = (j == n - 1) ? y1 : y0 + j * hy;
      if (i > 0 && i < n - 1 && j > 0 && j < n - 1) {
        double ux = u01(s), uy = u01(s);
        px += (jitter * (2.0 * ux - 1.0)) * hx;
        py += (jitter * (2.0 * uy - 1.0)) * hy;
      }
      x[g] = px;
      y[g] = py;
    }
  int c = 0;
  for (int j = 0; j < n - 1; ++j)
    for (int i = 0; i < n - 1; ++i) {
      int a = j * n + i, b = a + 1, cc = a + n + 1, d = a + n;
      double u = u01(s);
      int t[6];
      if (u < 0.5) { t[0] = a; t[1] = b; t[2] = cc; t[3] = a; t[4] = cc; t[5] = d; }
      else { t[0] = a; t[1] = b; t[2] = d; t[3] = b; t[4] = cc; t[5] = d; }
      for (int k = 0; k < 6; ++k) cv[3 * c + k] = t[k];
      c += 2;
    }
  int f = 0;
  for (int i = 0; i < n - 1; ++i) { btag[f] = 1; bv[2 * f] = i; bv[2 * f + 1] = i + 1; ++f; }
  for (int j = 0; j < n - 1; ++j) {
    btag[f] = 2; bv[2 * f] = j * n + (n - 1); bv[2 * f + 1] = (j + 1) * n + (n - 1); ++f;
  }
  for (int i = n - 2; i >= 0; --i) {
    btag[f] = 3; bv[2 * f] = (n - 1) * n + i + 1; bv[2 * f + 1] = (n - 1) * n + i; ++f;
  }
  for (int j = n - 2; j >= 0; --j) { btag[f] = 4; bv[2 * f] = (j + 1) * n; bv[2 * f + 1] = j * n; ++f; }
  if (periodic) {
    int k = 0, p = 0;
    pptr[0] = 0;
    pv[k++] = 0; pv[k++] = n - 1; pv[k++] = n * n - 1; pv[k++] = (n - 1) * n;
    pptr[++p] = k;
    for (int j = 1; j < n - 1; ++j) { pv[k++] = j * n; pv[k++] = j * n + n - 1; pptr[++p] = k; }
    for (int i = 1; i < n - 1; ++i) { pv[k++] = i; pv[k++] = (n - 1) * n + i; pptr[++p] = k; }
  }
  return TSM_OK;
}
