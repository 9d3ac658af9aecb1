// tsm_tiling.hip -- the vertex tiling of the LDS-staged kernels built on the
// device (the arrays build_tiles forms on the host, tiles.cpp, bit for bit),
// for every mesh but those of the fused kernel's second halo ring.
//
// Every per-tile list is a sort of 64-bit keys (tile << 32 | id) followed by
// binary searches: periodic-set members first (whole sets in set order, in
// tiles of their own), then the other vertices in Morton order of their
// quantised coordinates (a stable radix sort), per-vertex cell, face and boundary-face lists in
// ascending id (the reference's order), the halo of each tile (every corner
// of its vertices' cells and every end of their faces that another tile
// owns, ascending), its local faces (every face with an owned end, ascending
// = reference order) and boundary faces, then per owned vertex the 16-bit
// gather entries and the packed cell entries.  The arrays stay on the device;
// the host receives the numbering and the sizes.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <functional>
#include <string>
#include <type_traits>
#include <vector>

#include "tsm_device.h"
#include "tsm_host.h"

namespace tsm {
namespace {

constexpr unsigned long long NONE = ~0ull;

template <class T>
struct Scratch {  // device array freed on scope exit
  T* p = nullptr;
  hipError_t alloc(size_t n) { return hipMalloc(&p, (n ? n : 1) * sizeof(T)); }
  ~Scratch() { if (p) (void)hipFree(p); }
};

inline int nblk(long long n) { return (int)((std::max(n, 1LL) + 255) / 256); }

__device__ __forceinline__ unsigned spread16d(unsigned v) {
  v &= 0xffff;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// Morton key of each vertex from its coordinates quantised to 16 bits, with
// the host's expressions (this file is compiled without contraction too).
// Periodic-set members get a key above every Morton code: they sort last and
// the host places them first, in set order.
__global__ void k_morton(int nv, const double* x, const double* y, double x0, double y0, double sx,
                         double sy, const int* per, unsigned long long* key, int* id) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const unsigned qx = (unsigned)((x[i] - x0) * sx + 0.5), qy = (unsigned)((y[i] - y0) * sy + 0.5);
  key[i] = ((unsigned long long)(per[i] >= 0 ? 1u : 0u) << 32) | (spread16d(qx) | (spread16d(qy) << 1));
  id[i] = i;
}

// tile of each new vertex: the last tile start at or below it
__global__ void k_owner(int nv, int nt, const int* vbeg, int* owner) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nv) return;
  int lo = 0, hi = nt;  // vbeg[lo] <= n < vbeg[hi]
  while (hi - lo > 1) {
    const int m = (lo + hi) >> 1;
    if (vbeg[m] <= n) lo = m; else hi = m;
  }
  owner[n] = lo;
}

__global__ void k_scatter_perm(int nv, const int* order, int* perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) perm[order[i]] = i;
}

// (new vertex << 32 | element) for each of the `per` vertices of every
// element (cells: 3, faces and boundary faces: 2); skip[e] drops an element
__global__ void k_incidence(long long n, int per, const int* ev, const int* perm, const int* skip,
                            unsigned long long* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long e = i / per;
  out[i] = (skip && skip[e]) ? NONE
                             : ((unsigned long long)(unsigned)perm[ev[i]] << 32) | (unsigned)e;
}

// first index of keys[0, n) that is >= k
__device__ __forceinline__ long long lower_bound64(const unsigned long long* keys, long long n,
                                                   unsigned long long k) {
  long long lo = 0, hi = n;
  while (lo < hi) {
    const long long m = (lo + hi) >> 1;
    if (keys[m] < k) lo = m + 1; else hi = m;
  }
  return lo;
}
__device__ __forceinline__ int lower_bound32(const int* a, int lo, int hi, int k) {
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (a[m] < k) lo = m + 1; else hi = m;
  }
  return lo;
}

// ptr[i] = first of the sorted keys whose high word is >= i, i in [0, n]
__global__ void k_segments(int n, const unsigned long long* keys, long long nk, int* ptr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= n) ptr[i] = (int)lower_bound64(keys, nk, (unsigned long long)(unsigned)i << 32);
}

// number of sorted keys before the first NONE
__global__ void k_count_valid(const unsigned long long* keys, long long nk, int* out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = (int)lower_bound64(keys, nk, NONE);
}

__global__ void k_words(long long n, const unsigned long long* keys, int* lo, int* hi) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (lo) lo[i] = (int)(unsigned)(keys[i] & 0xffffffffull);
  if (hi) hi[i] = (int)(unsigned)(keys[i] >> 32);
}

__global__ void k_halo_count(int nv, const int* vc_ptr, const int* vf_ptr, int* cnt) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < nv) cnt[n] = 3 * (vc_ptr[n + 1] - vc_ptr[n]) + 2 * (vf_ptr[n + 1] - vf_ptr[n]);
  if (n == nv) cnt[n] = 0;
}

// halo candidates of each vertex: the corners of its cells, then the ends of
// its faces, as (tile << 32 | vertex) where another tile owns them
__global__ void k_halo_candidates(int nv, const int* owner, const int* vc_ptr, const int* vc,
                                  const int* cell, const int* vf_ptr, const int* vf, const int* fv,
                                  const int* perm, const int* off, unsigned long long* out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nv) return;
  const int t = owner[n];
  int o = off[n];
  for (int k = vc_ptr[n]; k < vc_ptr[n + 1]; ++k)
    for (int j = 0; j < 3; ++j) {
      const int u = perm[cell[3 * (size_t)vc[k] + j]];
      out[o++] = owner[u] == t ? NONE : (((unsigned long long)(unsigned)t << 32) | (unsigned)u);
    }
  for (int k = vf_ptr[n]; k < vf_ptr[n + 1]; ++k)
    for (int j = 0; j < 2; ++j) {
      const int u = perm[fv[2 * (size_t)vf[k] + j]];
      out[o++] = owner[u] == t ? NONE : (((unsigned long long)(unsigned)t << 32) | (unsigned)u);
    }
}

// the tiles an element (face or boundary face) is local to: those owning an end
__global__ void k_local(int ne, const int* owner, const int* ev, const int* perm, const int* skip,
                        unsigned long long* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  if (skip && skip[e]) { out[2 * (size_t)e] = out[2 * (size_t)e + 1] = NONE; return; }
  const unsigned t0 = (unsigned)owner[perm[ev[2 * (size_t)e]]];
  const unsigned t1 = (unsigned)owner[perm[ev[2 * (size_t)e + 1]]];
  out[2 * (size_t)e] = ((unsigned long long)t0 << 32) | (unsigned)e;
  out[2 * (size_t)e + 1] = t1 == t0 ? NONE : (((unsigned long long)t1 << 32) | (unsigned)e);
}

struct TileArrays {
  int nv;
  const int *perm, *owner, *vbeg, *hbeg, *halo;
};

// slot of new vertex u in tile t: owned slots first, then the tile's halo
// (a vertex in neither gets a slot past every limit, which the callers flag)
__device__ __forceinline__ int slot_of(const TileArrays& A, int t, int u) {
  const int b0 = A.vbeg[t], b1 = A.vbeg[t + 1];
  if (u >= b0 && u < b1) return u - b0;
  const int h0 = A.hbeg[t], h1 = A.hbeg[t + 1];
  const int k = lower_bound32(A.halo, h0, h1, u);
  return (k < h1 && A.halo[k] == u) ? (b1 - b0) + (k - h0) : 0x7fffffff;
}

// local faces: (left slot | right slot << 16) and the normal
__global__ void k_face_slots(int n, TileArrays A, const int* lf_tile, const int* lf, const int* fv,
                             const double* fnx, const double* fny, unsigned* f_lr, double* f_n,
                             int* bad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int t = lf_tile[k], f = lf[k];
  const int l = slot_of(A, t, A.perm[fv[2 * (size_t)f]]);
  const int r = slot_of(A, t, A.perm[fv[2 * (size_t)f + 1]]);
  if (l > 1023 || r > 1023) atomicOr(bad, 1);
  f_lr[k] = (unsigned)(l & 0xffff) | ((unsigned)(r & 0xffff) << 16);
  f_n[2 * (size_t)k] = fnx[f];
  f_n[2 * (size_t)k + 1] = fny[f];
}

__global__ void k_mark_corr(int nb, const int* lcell, const int* skip, unsigned char* ccorr) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb && !skip[b]) ccorr[lcell[b]] = 1;
}

__global__ void k_add(int n, const int* a, const int* b, int* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}

// per owned vertex: its faces as local-face indices (bit 15: right end), then
// its boundary faces (bit 14), in ascending id; its cells as packed slots
// with the boundary-correction flag, and whether any of them has one
__global__ void k_vertex_entries(TileArrays A, const int* vf_ptr, const int* vf, const int* fv,
                                 const int* vb_ptr, const int* vb, const int* fbeg, const int* lf,
                                 const int* bbeg, const int* tbf, const int* gptr,
                                 const int* vc_ptr, const int* vc, const int* cell,
                                 const unsigned char* ccorr, int n_periodic_tiles,
                                 unsigned short* gent, unsigned* cent, int* corr, int* bad) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= A.nv) return;
  const int t = A.owner[n];
  int o = gptr[n];
  for (int k = vf_ptr[n]; k < vf_ptr[n + 1]; ++k) {
    const int f = vf[k];
    const int j = lower_bound32(lf, fbeg[t], fbeg[t + 1], f) - fbeg[t];
    if (j > 0x3fff) atomicOr(bad, 2);
    unsigned e = (unsigned)j & 0x3fffu;
    if (A.perm[fv[2 * (size_t)f + 1]] == n) e |= 0x8000u;
    gent[o++] = (unsigned short)e;
  }
  for (int k = vb_ptr[n]; k < vb_ptr[n + 1]; ++k) {
    const int j = lower_bound32(tbf, bbeg[t], bbeg[t + 1], vb[k]) - bbeg[t];
    if (j > 0x3fff) atomicOr(bad, 2);
    gent[o++] = (unsigned short)(0x4000u | ((unsigned)j & 0x3fffu));
  }
  int any = 0;
  for (int k = vc_ptr[n]; k < vc_ptr[n + 1]; ++k) {
    const int c = vc[k];
    unsigned s[3];
    for (int j = 0; j < 3; ++j) {
      const int q = slot_of(A, t, A.perm[cell[3 * (size_t)c + j]]);
      if (q > 1023) atomicOr(bad, 1);
      s[j] = (unsigned)q & 1023u;
    }
    cent[k] = s[0] | (s[1] << 10) | (s[2] << 20) | ((unsigned)ccorr[c] << 30);
    any |= ccorr[c];
  }
  // periodic tiles form their boundary corrections themselves
  corr[n] = t >= n_periodic_tiles ? any : 0;
}

__global__ void k_tile_sizes(int nt, const int* vbeg, const int* hbeg, const int* fbeg, int* ns,
                             int* nfl) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  ns[t] = vbeg[t + 1] - vbeg[t] + hbeg[t + 1] - hbeg[t];
  nfl[t] = fbeg[t + 1] - fbeg[t];
}

__global__ void k_iota(int n, int* a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

__global__ void k_fill(int n, int v, int* a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

}  // namespace

int build_tiles_device(const Geometry& g, int tb, bool ring2, Tiling& T, TileGeo& TG,
                       const std::function<int(void**, size_t)>& dalloc, std::string& msg) {
  const int nv = g.nv, nc = g.nc, nf = (int)g.face.size(), nb = (int)g.bface.size();
  if (ring2 || tb <= 0 || nv <= 0) return 1;
  // the halo candidates (at most 9 per cell and 4 per face) index with int
  if (9LL * nc + 4LL * nf >= (1LL << 31)) return 1;
#define TT(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { msg = std::string("device tiling: ") + hipGetErrorString(e_); return TSM_E_HIP; } } while (0)
#define TA(ptr, n) do { void* q_ = nullptr; int r_ = dalloc(&q_, std::max<size_t>((size_t)(n), 1) * sizeof(*(ptr))); if (r_) { msg = "device tiling: allocation"; return r_; } (ptr) = static_cast<std::remove_reference_t<decltype(ptr)>>(q_); } while (0)
  hipStream_t st;
  TT(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sg{st};
  // ---- tiles: periodic sets first, whole and in set order, a quarter tile
  // of members per tile (their members lie on opposite sides of the domain,
  // so each brings its own halo), then every tb vertices of the rest ----------
  std::vector<int> prefix, tile_start;
  {
    const int cap = std::max(4, tb / 4);
    int cnt = 0;
    for (const auto& P : g.periodic) {
      if ((int)P.size() > cap) { msg = "periodic set larger than a tile"; return TSM_E_MESH; }
      if (cnt == 0 || cnt + (int)P.size() > cap) { tile_start.push_back((int)prefix.size()); cnt = 0; }
      for (int v : P) prefix.push_back(v);
      cnt += (int)P.size();
    }
  }
  const int npt = (int)tile_start.size(), np = (int)prefix.size();
  int n_member = 0;
  for (int i = 0; i < nv; ++i) n_member += g.periodic_of[i] >= 0;
  if (n_member != np) { msg = "tiling lost vertices"; return TSM_E_MESH; }
  for (int k = np; k < nv; k += tb) tile_start.push_back(k);
  const int nt = (int)tile_start.size();
  std::vector<int> h_vbeg(tile_start);
  h_vbeg.push_back(nv);
  // ---- inputs -----------------------------------------------------------------
  const int nbx = std::max(nb, 1);
  std::vector<int> h_fv(2 * (size_t)nf), h_bv(2 * (size_t)nbx, 0), h_lc(nbx, 0), h_bskip(nbx, 1);
  std::vector<double> h_fnx(nf), h_fny(nf);
  for (int f = 0; f < nf; ++f) {
    h_fv[2 * f] = g.face[f].v[0]; h_fv[2 * f + 1] = g.face[f].v[1];
    h_fnx[f] = g.face[f].nx; h_fny[f] = g.face[f].ny;
  }
  for (int b = 0; b < nb; ++b) {
    h_bv[2 * b] = g.bface[b].v[0]; h_bv[2 * b + 1] = g.bface[b].v[1];
    h_lc[b] = g.bface[b].lcell;
    h_bskip[b] = g.bface[b].periodic ? 1 : 0;
  }
  Scratch<double> x, y, fnx, fny;
  Scratch<int> cell, fv, bv, lc, bskip;
  TT(x.alloc(nv)); TT(y.alloc(nv)); TT(fnx.alloc(nf)); TT(fny.alloc(nf));
  TT(cell.alloc(3 * (size_t)nc)); TT(fv.alloc(2 * (size_t)nf)); TT(bv.alloc(h_bv.size()));
  TT(lc.alloc(nbx)); TT(bskip.alloc(nbx));
  TT(hipMemcpyAsync(x.p, g.x.data(), nv * sizeof(double), hipMemcpyHostToDevice, st));
  TT(hipMemcpyAsync(y.p, g.y.data(), nv * sizeof(double), hipMemcpyHostToDevice, st));
  TT(hipMemcpyAsync(fnx.p, h_fnx.data(), nf * sizeof(double), hipMemcpyHostToDevice, st));
  TT(hipMemcpyAsync(fny.p, h_fny.data(), nf * sizeof(double), hipMemcpyHostToDevice, st));
  TT(hipMemcpyAsync(cell.p, g.cell.data(), 3 * (size_t)nc * sizeof(int), hipMemcpyHostToDevice, st));
  TT(hipMemcpyAsync(fv.p, h_fv.data(), h_fv.size() * sizeof(int), hipMemcpyHostToDevice, st));
  TT(hipMemcpyAsync(bv.p, h_bv.data(), h_bv.size() * sizeof(int), hipMemcpyHostToDevice, st));
  TT(hipMemcpyAsync(lc.p, h_lc.data(), nbx * sizeof(int), hipMemcpyHostToDevice, st));
  TT(hipMemcpyAsync(bskip.p, h_bskip.data(), nbx * sizeof(int), hipMemcpyHostToDevice, st));
  // ---- one temporary-storage block for the largest hipCUB call ----------------
  const long long n_inc = std::max<long long>({3LL * nc, 2LL * nf, 2LL * nbx});
  const long long n_cap = 9LL * nc + 4LL * nf;  // halo candidates, at most
  size_t tmp_bytes = 0;
  {
    size_t a = 0;
    TT(hipcub::DeviceRadixSort::SortPairs(nullptr, a, (unsigned long long*)nullptr,
                                          (unsigned long long*)nullptr, (int*)nullptr, (int*)nullptr,
                                          nv, 0, 33, st));
    tmp_bytes = std::max(tmp_bytes, a);
    TT(hipcub::DeviceRadixSort::SortKeys(nullptr, a, (unsigned long long*)nullptr,
                                         (unsigned long long*)nullptr, (int)n_cap, 0, 64, st));
    tmp_bytes = std::max(tmp_bytes, a);
    TT(hipcub::DeviceSelect::Unique(nullptr, a, (unsigned long long*)nullptr,
                                    (unsigned long long*)nullptr, (int*)nullptr, (int)n_cap, st));
    tmp_bytes = std::max(tmp_bytes, a);
    TT(hipcub::DeviceScan::ExclusiveSum(nullptr, a, (int*)nullptr, (int*)nullptr, nv + 1, st));
    tmp_bytes = std::max(tmp_bytes, a);
    TT(hipcub::DeviceReduce::Max(nullptr, a, (int*)nullptr, (int*)nullptr, nt, st));
    tmp_bytes = std::max(tmp_bytes, a);
    TT(hipcub::DeviceSelect::Flagged(nullptr, a, (int*)nullptr, (int*)nullptr, (int*)nullptr,
                                     (int*)nullptr, nv, st));
    tmp_bytes = std::max(tmp_bytes, a);
  }
  Scratch<char> tmp;
  TT(tmp.alloc(tmp_bytes));
  size_t ts = tmp_bytes;
  // ---- numbering: vertices in Morton order (stable) ----------------------------
  double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
  for (int i = 0; i < nv; ++i) {
    x0 = std::min(x0, g.x[i]); x1 = std::max(x1, g.x[i]);
    y0 = std::min(y0, g.y[i]); y1 = std::max(y1, g.y[i]);
  }
  const double sx = (x1 > x0) ? 65535.0 / (x1 - x0) : 0.0, sy = (y1 > y0) ? 65535.0 / (y1 - y0) : 0.0;
  Scratch<unsigned long long> mk, mk_s;
  Scratch<int> mid, msorted, order, perm, per, owner;
  TT(mk.alloc(nv)); TT(mk_s.alloc(nv)); TT(mid.alloc(nv)); TT(msorted.alloc(nv));
  TT(order.alloc(nv)); TT(perm.alloc(nv)); TT(per.alloc(nv)); TT(owner.alloc(nv));
  TT(hipMemcpyAsync(per.p, g.periodic_of.data(), nv * sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_morton, dim3(nblk(nv)), dim3(256), 0, st, nv, x.p, y.p, x0, y0, sx, sy, per.p,
                     mk.p, mid.p);
  ts = tmp_bytes;
  TT(hipcub::DeviceRadixSort::SortPairs(tmp.p, ts, mk.p, mk_s.p, mid.p, msorted.p, nv, 0, 33, st));
  // new numbering: the periodic members in set order, then the others by key
  if (np > 0)
    TT(hipMemcpyAsync(order.p, prefix.data(), np * sizeof(int), hipMemcpyHostToDevice, st));
  if (nv > np)
    TT(hipMemcpyAsync(order.p + np, msorted.p, (nv - np) * sizeof(int), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_scatter_perm, dim3(nblk(nv)), dim3(256), 0, st, nv, order.p, perm.p);
  int* d_vbeg = nullptr;
  TA(d_vbeg, nt + 1);
  TT(hipMemcpyAsync(d_vbeg, h_vbeg.data(), (nt + 1) * sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_owner, dim3(nblk(nv)), dim3(256), 0, st, nv, nt, d_vbeg, owner.p);
  // ---- vertex incidences: cells, faces, boundary faces, ascending -------------
  Scratch<unsigned long long> inc, inc_s;
  TT(inc.alloc(n_inc)); TT(inc_s.alloc(n_inc));
  // per new vertex: a sorted incidence list (low words) and its pointers
  auto incidence = [&](long long n, int per, const int* ev, const int* skip, int* ptr,
                       int* list) -> int {
    hipLaunchKernelGGL(k_incidence, dim3(nblk(n)), dim3(256), 0, st, n, per, ev, perm.p, skip, inc.p);
    size_t t2 = tmp_bytes;
    TT(hipcub::DeviceRadixSort::SortKeys(tmp.p, t2, inc.p, inc_s.p, (int)n, 0, 64, st));
    hipLaunchKernelGGL(k_segments, dim3(nblk(nv + 1)), dim3(256), 0, st, nv, inc_s.p, n, ptr);
    hipLaunchKernelGGL(k_words, dim3(nblk(n)), dim3(256), 0, st, n, inc_s.p, list, (int*)nullptr);
    return 0;
  };
  int *d_cptr = nullptr, *d_cgid = nullptr;   // cells per vertex: the tiling's own arrays
  TA(d_cptr, nv + 1);
  TA(d_cgid, 3 * (size_t)nc);
  Scratch<int> vf_ptr, vf, vb_ptr, vb;
  TT(vf_ptr.alloc(nv + 1)); TT(vf.alloc(2 * (size_t)nf));
  TT(vb_ptr.alloc(nv + 1)); TT(vb.alloc(2 * (size_t)nbx));
  if (int r = incidence(3LL * nc, 3, cell.p, nullptr, d_cptr, d_cgid)) return r;
  if (int r = incidence(2LL * nf, 2, fv.p, nullptr, vf_ptr.p, vf.p)) return r;
  // boundary faces: periodic ones sort last as NONE, past every pointer
  if (int r = incidence(2LL * nbx, 2, bv.p, bskip.p, vb_ptr.p, vb.p)) return r;
  // ---- halos -------------------------------------------------------------------
  Scratch<int> hcnt, hoff, small;
  TT(hcnt.alloc(nv + 1)); TT(hoff.alloc(nv + 1)); TT(small.alloc(4));
  hipLaunchKernelGGL(k_halo_count, dim3(nblk(nv + 1)), dim3(256), 0, st, nv, d_cptr, vf_ptr.p, hcnt.p);
  ts = tmp_bytes;
  TT(hipcub::DeviceScan::ExclusiveSum(tmp.p, ts, hcnt.p, hoff.p, nv + 1, st));
  int n_cand = 0;
  TT(hipMemcpyAsync(&n_cand, hoff.p + nv, sizeof(int), hipMemcpyDeviceToHost, st));
  TT(hipStreamSynchronize(st));
  if (n_cand < 0 || n_cand > n_cap) { msg = "device tiling: halo candidate count"; return TSM_E_MESH; }
  Scratch<unsigned long long> hk, hk_s, hk_u;
  TT(hk.alloc(n_cand)); TT(hk_s.alloc(n_cand)); TT(hk_u.alloc(n_cand));
  hipLaunchKernelGGL(k_halo_candidates, dim3(nblk(nv)), dim3(256), 0, st, nv, owner.p, d_cptr, d_cgid,
                     cell.p, vf_ptr.p, vf.p, fv.p, perm.p, hoff.p, hk.p);
  ts = tmp_bytes;
  TT(hipcub::DeviceRadixSort::SortKeys(tmp.p, ts, hk.p, hk_s.p, n_cand, 0, 64, st));
  ts = tmp_bytes;
  TT(hipcub::DeviceSelect::Unique(tmp.p, ts, hk_s.p, hk_u.p, small.p, n_cand, st));
  int nu = 0;
  TT(hipMemcpyAsync(&nu, small.p, sizeof(int), hipMemcpyDeviceToHost, st));
  TT(hipStreamSynchronize(st));
  hipLaunchKernelGGL(k_count_valid, dim3(1), dim3(64), 0, st, hk_u.p, (long long)nu, small.p + 1);
  int nhalo = 0;
  TT(hipMemcpyAsync(&nhalo, small.p + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  TT(hipStreamSynchronize(st));
  int *d_hbeg = nullptr, *d_halo = nullptr;
  TA(d_hbeg, nt + 1);
  TA(d_halo, nhalo);
  hipLaunchKernelGGL(k_segments, dim3(nblk(nt + 1)), dim3(256), 0, st, nt, hk_u.p, (long long)nhalo,
                     d_hbeg);
  hipLaunchKernelGGL(k_words, dim3(nblk(nhalo)), dim3(256), 0, st, (long long)nhalo, hk_u.p, d_halo,
                     (int*)nullptr);
  // ---- local faces and boundary faces of each tile ----------------------------
  Scratch<unsigned long long> lk, lk_s;
  TT(lk.alloc(2 * (size_t)std::max(nf, nbx))); TT(lk_s.alloc(2 * (size_t)std::max(nf, nbx)));
  Scratch<int> lf_tile;
  TT(lf_tile.alloc(2 * (size_t)nf));
  auto locals = [&](int ne, const int* ev, const int* skip, int** beg, int** list, int* tile_of,
                    int& count) -> int {
    hipLaunchKernelGGL(k_local, dim3(nblk(ne)), dim3(256), 0, st, ne, owner.p, ev, perm.p, skip, lk.p);
    size_t t2 = tmp_bytes;
    TT(hipcub::DeviceRadixSort::SortKeys(tmp.p, t2, lk.p, lk_s.p, 2 * ne, 0, 64, st));
    hipLaunchKernelGGL(k_count_valid, dim3(1), dim3(64), 0, st, lk_s.p, 2LL * ne, small.p + 2);
    TT(hipMemcpyAsync(&count, small.p + 2, sizeof(int), hipMemcpyDeviceToHost, st));
    TT(hipStreamSynchronize(st));
    TA(*beg, nt + 1);
    TA(*list, count);
    hipLaunchKernelGGL(k_segments, dim3(nblk(nt + 1)), dim3(256), 0, st, nt, lk_s.p, (long long)count,
                       *beg);
    hipLaunchKernelGGL(k_words, dim3(nblk(count)), dim3(256), 0, st, (long long)count, lk_s.p, *list,
                       tile_of);
    return 0;
  };
  int *d_fbeg = nullptr, *d_lf = nullptr, *d_bbeg = nullptr, *d_tbf = nullptr;
  int n_lf = 0, n_tbf = 0;
  if (int r = locals(nf, fv.p, nullptr, &d_fbeg, &d_lf, lf_tile.p, n_lf)) return r;
  if (int r = locals(nbx, bv.p, bskip.p, &d_bbeg, &d_tbf, nullptr, n_tbf)) return r;
  // ---- slots of the local faces, gather and cell entries ----------------------
  const TileArrays A{nv, perm.p, owner.p, d_vbeg, d_hbeg, d_halo};
  unsigned* d_flr = nullptr;
  double* d_fn = nullptr;
  TA(d_flr, n_lf);
  TA(d_fn, 2 * (size_t)n_lf);
  Scratch<int> bad;
  TT(bad.alloc(1));
  TT(hipMemsetAsync(bad.p, 0, sizeof(int), st));
  hipLaunchKernelGGL(k_face_slots, dim3(nblk(n_lf)), dim3(256), 0, st, n_lf, A, lf_tile.p, d_lf,
                     fv.p, fnx.p, fny.p, d_flr, d_fn, bad.p);
  Scratch<unsigned char> ccorr;
  TT(ccorr.alloc(nc));
  TT(hipMemsetAsync(ccorr.p, 0, nc, st));
  if (nb > 0)
    hipLaunchKernelGGL(k_mark_corr, dim3(nblk(nb)), dim3(256), 0, st, nb, lc.p, bskip.p, ccorr.p);
  int* d_gptr = nullptr;
  TA(d_gptr, nv + 1);
  hipLaunchKernelGGL(k_add, dim3(nblk(nv + 1)), dim3(256), 0, st, nv + 1, vf_ptr.p, vb_ptr.p, d_gptr);
  int n_gent = 0;
  TT(hipMemcpyAsync(&n_gent, d_gptr + nv, sizeof(int), hipMemcpyDeviceToHost, st));
  TT(hipStreamSynchronize(st));
  unsigned short* d_gent = nullptr;
  unsigned* d_cent = nullptr;
  TA(d_gent, n_gent);
  TA(d_cent, 3 * (size_t)nc);
  Scratch<int> corr, iota, gsel;
  TT(corr.alloc(nv)); TT(iota.alloc(nv)); TT(gsel.alloc(nv));
  hipLaunchKernelGGL(k_vertex_entries, dim3(nblk(nv)), dim3(256), 0, st, A, vf_ptr.p, vf.p, fv.p,
                     vb_ptr.p, vb.p, d_fbeg, d_lf, d_bbeg, d_tbf, d_gptr, d_cptr, d_cgid, cell.p,
                     ccorr.p, npt, d_gent, d_cent, corr.p, bad.p);
  // vertices with a boundary-corrected cell (outside the periodic tiles), ascending
  hipLaunchKernelGGL(k_iota, dim3(nblk(nv)), dim3(256), 0, st, nv, iota.p);
  ts = tmp_bytes;
  TT(hipcub::DeviceSelect::Flagged(tmp.p, ts, iota.p, corr.p, gsel.p, small.p + 3, nv, st));
  // largest slot and local-face counts of a tile
  Scratch<int> tns, tnf, mx;
  TT(tns.alloc(nt)); TT(tnf.alloc(nt)); TT(mx.alloc(2));
  hipLaunchKernelGGL(k_tile_sizes, dim3(nblk(nt)), dim3(256), 0, st, nt, d_vbeg, d_hbeg, d_fbeg,
                     tns.p, tnf.p);
  ts = tmp_bytes;
  TT(hipcub::DeviceReduce::Max(tmp.p, ts, tns.p, mx.p, nt, st));
  ts = tmp_bytes;
  TT(hipcub::DeviceReduce::Max(tmp.p, ts, tnf.p, mx.p + 1, nt, st));
  int h_small[4] = {0, 0, 0, 0}, h_mx[2] = {0, 0}, h_bad = 0;
  TT(hipMemcpyAsync(h_small, small.p, sizeof(h_small), hipMemcpyDeviceToHost, st));
  TT(hipMemcpyAsync(h_mx, mx.p, sizeof(h_mx), hipMemcpyDeviceToHost, st));
  TT(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  TT(hipStreamSynchronize(st));
  if (h_mx[0] > 1023 || (h_bad & 1)) { msg = "tile exceeds 1023 slots"; return TSM_E_MESH; }
  if (h_mx[1] >= (1 << 14) || (h_bad & 2)) { msg = "tile exceeds 16383 local faces"; return TSM_E_MESH; }
  const int n_gbv = h_small[3];
  int* d_gbv = nullptr;
  TA(d_gbv, n_gbv);
  if (n_gbv > 0)
    TT(hipMemcpyAsync(d_gbv, gsel.p, n_gbv * sizeof(int), hipMemcpyDeviceToDevice, st));
  // periodic tiles and, per member, its set's first slot and size (the
  // members are the first np new ids, in set order)
  std::vector<int> h_tper(nt, 0), h_pinfo(nv, -1);
  {
    int k = 0, t = 0;
    for (const auto& P : g.periodic) {
      while (t + 1 < npt && h_vbeg[t + 1] <= k) ++t;
      for (size_t m = 0; m < P.size(); ++m) h_pinfo[k + m] = ((k - h_vbeg[t]) << 8) | (int)P.size();
      h_tper[t] = 1;
      k += (int)P.size();
    }
  }
  int *d_tper = nullptr, *d_pinfo = nullptr;
  TA(d_tper, nt);
  TA(d_pinfo, nv);
  TT(hipMemcpyAsync(d_tper, h_tper.data(), nt * sizeof(int), hipMemcpyHostToDevice, st));
  TT(hipMemcpyAsync(d_pinfo, h_pinfo.data(), nv * sizeof(int), hipMemcpyHostToDevice, st));
  // ---- the host's part: numbering and sizes -------------------------------------
  T = Tiling{};
  T.tb = tb;
  T.ntiles = nt;
  T.n_periodic_tiles = npt;
  T.max_slots = h_mx[0];
  T.max_faces = h_mx[1];
  T.inv.resize(nv);
  T.perm.resize(nv);
  TT(hipMemcpyAsync(T.inv.data(), order.p, nv * sizeof(int), hipMemcpyDeviceToHost, st));
  TT(hipMemcpyAsync(T.perm.data(), perm.p, nv * sizeof(int), hipMemcpyDeviceToHost, st));
  T.vbeg = h_vbeg;
  TT(hipStreamSynchronize(st));
  TG.vbeg = d_vbeg; TG.hbeg = d_hbeg; TG.halo = d_halo; TG.fbeg = d_fbeg; TG.bbeg = d_bbeg;
  TG.tbf = d_tbf; TG.gptr = d_gptr; TG.cptr = d_cptr; TG.cgid = d_cgid; TG.tper = d_tper;
  TG.pinfo = d_pinfo; TG.f_lr = d_flr; TG.cent = d_cent; TG.f_n = d_fn; TG.gent = d_gent;
  TG.gbv = d_gbv;
  TG.n_gbv = n_gbv;
#undef TT
#undef TA
  return 0;
}

}  // namespace tsm

namespace {
template <class T>
long long count_diff(const T* dev, const std::vector<T>& host, size_t n, int& err) {
  std::vector<T> h(std::max<size_t>(n, 1));
  if (n && hipMemcpy(h.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) {
    err = 1;
    return 0;
  }
  long long d = (long long)std::max(host.size(), n) - (long long)std::min(host.size(), n);
  for (size_t i = 0; i < std::min(host.size(), n); ++i) d += !(h[i] == host[i]);
  return d;
}
}  // namespace

extern "C" int tsm_device_tiling_check(const tsm_mesh* mesh, int tile_vertices, long long* n_diff,
                                       char* which) {
  if (!mesh || !n_diff || !which || tile_vertices <= 0) return TSM_E_ARG;
  tsm::Geometry g;
  std::string msg;
  if (!tsm::build_geometry(*mesh, 0, g, msg)) return TSM_E_MESH;
  tsm::Tiling H, D;
  if (!tsm::build_tiles(g, tile_vertices, H, msg, false) || !tsm::validate_tiles(g, H, msg))
    return TSM_E_MESH;
  std::vector<void*> owned;
  struct Free { std::vector<void*>& v; ~Free() { for (void* p : v) (void)hipFree(p); } } fr{owned};
  auto dalloc = [&owned](void** p, size_t bytes) -> int {
    if (hipMalloc(p, bytes) != hipSuccess) return TSM_E_HIP;
    owned.push_back(*p);
    return TSM_OK;
  };
  tsm::TileGeo TG{};
  const int rc = tsm::build_tiles_device(g, tile_vertices, false, D, TG, dalloc, msg);
  if (rc == 1) return TSM_E_ARG;
  if (rc != 0) return rc;
  int err = 0;
  long long total = 0;
  which[0] = 0;
  auto note = [&](const char* name, long long d) {
    if (d && !which[0]) { std::snprintf(which, 16, "%s", name); }
    total += d;
  };
  const int nv = g.nv;
  note("scalars", (D.ntiles != H.ntiles) + (D.max_slots != H.max_slots) +
                      (D.max_faces != H.max_faces) + (D.n_periodic_tiles != H.n_periodic_tiles) +
                      (TG.n_gbv != (int)H.gbv.size()));
  note("perm", (long long)(D.perm != H.perm) + (D.inv != H.inv) + (D.vbeg != H.vbeg));
  note("vbeg", count_diff(TG.vbeg, H.vbeg, H.vbeg.size(), err));
  note("hbeg", count_diff(TG.hbeg, H.hbeg, H.hbeg.size(), err));
  note("halo", count_diff(TG.halo, H.halo, H.halo.size(), err));
  note("fbeg", count_diff(TG.fbeg, H.fbeg, H.fbeg.size(), err));
  note("f_lr", count_diff(TG.f_lr, H.f_lr, H.f_lr.size(), err));
  note("f_n", count_diff(TG.f_n, H.f_n, H.f_n.size(), err));
  note("bbeg", count_diff(TG.bbeg, H.bbeg, H.bbeg.size(), err));
  note("tbf", count_diff(TG.tbf, H.tbf, H.tbf.size(), err));
  note("gptr", count_diff(TG.gptr, H.gptr, (size_t)nv + 1, err));
  note("gent", count_diff(TG.gent, H.gent, H.gent.size(), err));
  note("cptr", count_diff(TG.cptr, H.cptr, (size_t)nv + 1, err));
  note("cent", count_diff(TG.cent, H.cent, H.cent.size(), err));
  note("cgid", count_diff(TG.cgid, H.cgid, H.cgid.size(), err));
  note("tper", count_diff(TG.tper, H.tper, H.tper.size(), err));
  note("pinfo", count_diff(TG.pinfo, H.pinfo, H.pinfo.size(), err));
  if (TG.n_gbv > 0 && TG.n_gbv == (int)H.gbv.size()) note("gbv", count_diff(TG.gbv, H.gbv, H.gbv.size(), err));
  if (err) return TSM_E_HIP;
  *n_diff = total;
  return TSM_OK;
}
