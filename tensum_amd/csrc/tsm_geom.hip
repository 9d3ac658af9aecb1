// tsm_geom.hip -- dual-mesh geometry of median duals on the device (the part
// of Grid::preproc, src/grid_preproc.cc:1147-1194, that the time loop needs,
// for NPART = 1).  Faces are found by sorting edge keys and numbered in the
// order the reference creates them (the boundary faces of the part file first,
// then each cell's edges in cell order as they first appear,
// src/grid_preproc.cc:461-590); faces are oriented and get their normals, and
// cells and vertices their areas.  Every value is formed with the arithmetic
// of the host builder (geometry.cpp, compiled without contraction like this
// file), so the two agree bit for bit; the host builder remains for Voronoi
// duals and for meshes whose edge lists it has to repair (edges of a single
// cell that no boundary face names).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <string>
#include <vector>

#include "tsm_host.h"

namespace tsm {
namespace {

enum { G_OK = 0, G_FALLBACK = 1, G_BAD_EDGE = 2, G_DUP_BFACE = 3, G_LOOSE_BFACE = 4, G_AREA = 5 };

template <class T>
struct Dev {  // device array freed on scope exit
  T* p = nullptr;
  hipError_t alloc(size_t n) { return hipMalloc(&p, (n ? n : 1) * sizeof(T)); }
  ~Dev() { if (p) (void)hipFree(p); }
};

// Edge keys: boundary faces at positions [0, nb), then cell c's edges
// (v0,v1), (v1,v2), (v2,v0) at nb + 3c + k.  key = min * nv + max.
__global__ void k_edge_keys(int nb, int nc, long long nv, const int* bv, const int* cv,
                            unsigned long long* key, int* pos) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long ne = nb + 3LL * nc;
  if (i >= ne) return;
  int a, b;
  if (i < nb) {
    a = bv[2 * i]; b = bv[2 * i + 1];
  } else {
    const long long q = i - nb, c = q / 3;
    const int k = (int)(q % 3);
    a = cv[3 * c + k]; b = cv[3 * c + (k + 1) % 3];
  }
  const long long lo = a < b ? a : b, hi = a < b ? b : a;
  key[i] = (unsigned long long)(lo * nv + hi);
  pos[i] = (int)i;
}

// One thread per sorted entry; the first entry of each equal-key run (the
// sort is stable, so it holds the smallest position: the edge's first
// appearance) records the run: its cells in appearance order, whether it
// starts a new interior face, and any mesh error.
__global__ void k_runs(long long ne, int nb, const unsigned long long* key, const int* pos,
                       int* first, int* newflag, int* lcell, int* rcell, int* bad) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ne) return;
  if (i > 0 && key[i] == key[i - 1]) return;
  long long L = 1;
  while (i + L < ne && L < 4 && key[i + L] == key[i]) ++L;
  const int p0 = pos[i];
  for (long long j = 0; j < L; ++j) first[pos[i + j]] = p0;
  int code = G_OK, lc = -1, rc = -1;
  if (p0 < nb) {  // a boundary face of the part file
    for (long long j = 1; j < L; ++j) {
      const int p = pos[i + j];
      if (p < nb) { code = G_DUP_BFACE; break; }
      const int c = (p - nb) / 3;
      if (lc < 0) lc = c; else if (rc < 0) rc = c; else code = G_BAD_EDGE;
    }
    if (code == G_OK && lc < 0) code = G_LOOSE_BFACE;
  } else {
    newflag[p0 - nb] = 1;
    lc = (p0 - nb) / 3;
    if (L == 1) code = G_FALLBACK;       // a one-cell edge no boundary face names
    else if (L > 2) code = G_BAD_EDGE;
    else rc = (pos[i + 1] - nb) / 3;
  }
  lcell[p0] = lc;
  rcell[p0] = rc;
  if (code != G_OK) bad[p0] = code;
}

__device__ __forceinline__ int face_of(int h, int nb, const int* rank) {
  return h < nb ? h : nb + rank[h - nb];
}

// Face records in face order, and each cell edge's face.
__global__ void k_faces(long long ne, int nb, const int* bv, const int* btag, const int* bper,
                        const int* cv, const int* first, const int* rank, const int* lcell,
                        const int* rcell, int* fv, int* ftype, int* fper, int* flc, int* frc,
                        int* edge_face) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ne) return;
  const int h = first[p];
  const int f = face_of(h, nb, rank);
  if (p >= nb) edge_face[p - nb] = f;
  if (h != p) return;
  int a, b;
  if (p < nb) {
    a = bv[2 * p]; b = bv[2 * p + 1];
  } else {
    const long long q = p - nb, c = q / 3;
    const int k = (int)(q % 3);
    a = cv[3 * c + k]; b = cv[3 * c + (k + 1) % 3];
  }
  fv[2 * f] = a; fv[2 * f + 1] = b;
  ftype[f] = p < nb ? btag[p] : -1;
  fper[f] = p < nb ? bper[p] : 0;
  flc[f] = lcell[p];
  frc[f] = rcell[p];
}

// Per cell: face slots (its faces in face order, src/grid_preproc.cc:600-623),
// boundary flag, centroid and area (src/grid_preproc.cc:18-254).
__global__ void k_cells(int nc, int nb, const double* x, const double* y, const int* cv,
                        const int* edge_face, const int* fper, int* slot, int* nonper,
                        double* ccx, double* ccy, double* area, int* bad) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int f0 = edge_face[3 * c], f1 = edge_face[3 * c + 1], f2 = edge_face[3 * c + 2];
  if (f0 > f1) { int t = f0; f0 = f1; f1 = t; }
  if (f1 > f2) { int t = f1; f1 = f2; f2 = t; }
  if (f0 > f1) { int t = f0; f0 = f1; f1 = t; }
  slot[3 * c] = f0; slot[3 * c + 1] = f1; slot[3 * c + 2] = f2;
  int np = 0;
  if (f0 < nb && !fper[f0]) np = 1;
  if (f1 < nb && !fper[f1]) np = 1;
  if (f2 < nb && !fper[f2]) np = 1;
  nonper[c] = np;
  const int v0 = cv[3 * c], v1 = cv[3 * c + 1], v2 = cv[3 * c + 2];
  const double third = 1.0 / 3.0;
  ccx[c] = x[v0] * third + x[v1] * third + x[v2] * third;
  ccy[c] = y[v0] * third + y[v1] * third + y[v2] * third;
  const double l1x = x[v1] * 1.0 + x[v0] * -1.0, l1y = y[v1] * 1.0 + y[v0] * -1.0;
  const double l2x = x[v2] * 1.0 + x[v0] * -1.0, l2y = y[v2] * 1.0 + y[v0] * -1.0;
  const double a = 0.5 * (l1x * l2y - l1y * l2x);
  area[c] = a;
  if (!(a > 0.0)) bad[c] = G_AREA;
}

// Vertex sums of a third of each cell's area, cells in cell order (the order
// of the host loop): one thread per run of the (vertex, cell) pairs sorted by
// vertex (stable, so the cells of a vertex stay in cell order).
__global__ void k_vertex_area(int n, const int* vs, const int* cs, const double* area,
                              double* mcarea) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i > 0 && vs[i] == vs[i - 1]) return;
  double m = 0.0;
  for (int j = i; j < n && vs[j] == vs[i]; ++j) m += area[cs[j]] / 3.0;
  mcarea[vs[i]] = m;
}

// Face centroids, orientation and dual-face normals (src/grid_preproc.cc:
// 259-365), and the boundary faces' edge normals with their own orientation.
__global__ void k_normals(int nf, int nb, const double* x, const double* y, const double* ccx,
                          const double* ccy, int* fv, const int* ftype, const int* flc,
                          const int* frc, double* fc, double* fn, double* fmeas, int* bvo,
                          double* bn, double* bmeas) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  int v0 = fv[2 * f], v1 = fv[2 * f + 1];
  const double cx = x[v0] * 0.5 + x[v1] * 0.5, cy = y[v0] * 0.5 + y[v1] * 0.5;
  fc[2 * f] = cx; fc[2 * f + 1] = cy;
  if (f < nb) {  // boundary face: the part file's vertex order
    const double drx = x[v1] * 1.0 + x[v0] * -1.0, dry = y[v1] * 1.0 + y[v0] * -1.0;
    double nx = dry, ny = -drx;
    const double fx = x[v0] * 0.5 + x[v1] * 0.5, fy = y[v0] * 0.5 + y[v1] * 0.5;
    const double dcx = fx * 1.0 + ccx[flc[f]] * -1.0, dcy = fy * 1.0 + ccy[flc[f]] * -1.0;
    const double m = sqrt(nx * nx + ny * ny);
    int b0 = v0, b1 = v1;
    if (dcx * nx + dcy * ny < 0.0) { b0 = v1; b1 = v0; nx *= -1.0; ny *= -1.0; }
    bvo[2 * f] = b0; bvo[2 * f + 1] = b1;
    bn[2 * f] = nx; bn[2 * f + 1] = ny;
    bmeas[f] = m;
  }
  const int lc = flc[f];
  const double drx = x[v1] * 1.0 + x[v0] * -1.0, dry = y[v1] * 1.0 + y[v0] * -1.0;
  const double dcx = cx * 1.0 + ccx[lc] * -1.0, dcy = cy * 1.0 + ccy[lc] * -1.0;
  if (dcx * dry + dcy * -drx < 0.0) { fv[2 * f] = v1; fv[2 * f + 1] = v0; }
  double nx = cx * 1.0 + ccx[lc] * -1.0, ny = cy * 1.0 + ccy[lc] * -1.0;
  if (ftype[f] == -1) {
    nx += ccx[frc[f]] * 1.0 + cx * -1.0;
    ny += ccy[frc[f]] * 1.0 + cy * -1.0;
  }
  fn[2 * f] = -ny;
  fn[2 * f + 1] = nx;
  fmeas[f] = sqrt(ny * ny + nx * nx);
}

inline int blocks(long long n) { return (int)((n + 255) / 256); }

}  // namespace

// 0: built; 1: not handled here (the caller uses the host builder); TSM_E_*
// with `msg` on a mesh error or a HIP failure.
int build_geometry_device(const tsm_mesh& m, Geometry& g, std::string& msg) {
  const int nv = m.n_vertex, nc = m.n_cell, nb = m.n_bface;
  if (nv <= 0 || nc <= 0) { msg = "empty mesh"; return TSM_E_MESH; }
  const long long ne = nb + 3LL * nc;
  if (ne >= (1LL << 31) - 1) return 1;
  // host-side checks of the input (cheap, and the host builder's messages)
  for (long long k = 0; k < 3LL * nc; ++k)
    if (m.cell_v[k] < 0 || m.cell_v[k] >= nv) { msg = "cell vertex out of range"; return TSM_E_MESH; }
  std::vector<int> bper(std::max(nb, 1), 0), btype(std::max(nb, 1), 0);
  for (int i = 0; i < nb; ++i) {
    if (m.bface_v[2 * i] < 0 || m.bface_v[2 * i] >= nv || m.bface_v[2 * i + 1] < 0 ||
        m.bface_v[2 * i + 1] >= nv) {
      msg = "boundary face vertex out of range"; return TSM_E_MESH;
    }
    int bt = -1;
    for (int k = 0; k < m.n_bc; ++k)
      if (m.bc_tag[k] == m.bface_tag[i]) { bt = m.bc_type[k]; break; }
    if (bt < 0) { msg = "boundary tag " + std::to_string(m.bface_tag[i]) + " has no bc"; return TSM_E_MESH; }
    btype[i] = bt;
    bper[i] = bt == TSM_BC_PERIODIC;
  }
#define GT(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { msg = std::string("device geometry: ") + hipGetErrorString(e_); return TSM_E_HIP; } } while (0)
  hipStream_t st;
  GT(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sg{st};
  Dev<double> x, y;
  Dev<int> cv, bv, btag, bp;
  GT(x.alloc(nv)); GT(y.alloc(nv)); GT(cv.alloc(3 * (size_t)nc)); GT(bv.alloc(2 * (size_t)nb));
  GT(btag.alloc(nb)); GT(bp.alloc(nb));
  GT(hipMemcpyAsync(x.p, m.x, nv * sizeof(double), hipMemcpyHostToDevice, st));
  GT(hipMemcpyAsync(y.p, m.y, nv * sizeof(double), hipMemcpyHostToDevice, st));
  GT(hipMemcpyAsync(cv.p, m.cell_v, 3 * (size_t)nc * sizeof(int), hipMemcpyHostToDevice, st));
  if (nb > 0) {
    GT(hipMemcpyAsync(bv.p, m.bface_v, 2 * (size_t)nb * sizeof(int), hipMemcpyHostToDevice, st));
    GT(hipMemcpyAsync(btag.p, m.bface_tag, nb * sizeof(int), hipMemcpyHostToDevice, st));
    GT(hipMemcpyAsync(bp.p, bper.data(), nb * sizeof(int), hipMemcpyHostToDevice, st));
  }
  // ---- faces: sort the edge keys, number the runs ---------------------------
  Dev<unsigned long long> key, key_s;
  Dev<int> pos, pos_s, first, newflag, rank, lcell, rcell, bad;
  GT(key.alloc(ne)); GT(key_s.alloc(ne)); GT(pos.alloc(ne)); GT(pos_s.alloc(ne));
  GT(first.alloc(ne)); GT(newflag.alloc(3 * (size_t)nc)); GT(rank.alloc(3 * (size_t)nc));
  GT(lcell.alloc(ne)); GT(rcell.alloc(ne)); GT(bad.alloc(std::max<long long>(ne, nc)));
  GT(hipMemsetAsync(newflag.p, 0, 3 * (size_t)nc * sizeof(int), st));
  GT(hipMemsetAsync(bad.p, 0, std::max<long long>(ne, nc) * sizeof(int), st));
  hipLaunchKernelGGL(k_edge_keys, dim3(blocks(ne)), dim3(256), 0, st, nb, nc, (long long)nv, bv.p,
                     cv.p, key.p, pos.p);
  int key_bits = 1;
  while (key_bits < 64 && ((unsigned long long)nv * nv) >> key_bits) ++key_bits;
  size_t tmp_bytes = 0, tmp2 = 0, tmp3 = 0, tmp4 = 0;
  GT(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key.p, key_s.p, pos.p, pos_s.p, (int)ne,
                                        0, key_bits, st));
  GT(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, newflag.p, rank.p, 3 * nc, st));
  GT(hipcub::DeviceReduce::Max(nullptr, tmp3, bad.p, rank.p, (int)std::max<long long>(ne, nc), st));
  int vbits = 1;
  while (vbits < 32 && ((unsigned)nv >> vbits)) ++vbits;
  GT(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp4, cv.p, pos.p, pos.p, pos.p, 3 * nc, 0, vbits, st));
  Dev<char> tmp;
  GT(tmp.alloc(std::max(std::max(tmp_bytes, tmp2), std::max(tmp3, tmp4))));
  const size_t tmp_size = std::max(std::max(tmp_bytes, tmp2), std::max(tmp3, tmp4));
  size_t ts = tmp_size;
  GT(hipcub::DeviceRadixSort::SortPairs(tmp.p, ts, key.p, key_s.p, pos.p, pos_s.p, (int)ne, 0,
                                        key_bits, st));
  hipLaunchKernelGGL(k_runs, dim3(blocks(ne)), dim3(256), 0, st, ne, nb, key_s.p, pos_s.p, first.p,
                     newflag.p, lcell.p, rcell.p, bad.p);
  ts = tmp_size;
  GT(hipcub::DeviceScan::ExclusiveSum(tmp.p, ts, newflag.p, rank.p, 3 * nc, st));
  int n_new = 0, last_flag = 0;
  GT(hipMemcpyAsync(&n_new, rank.p + 3 * (size_t)nc - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  GT(hipMemcpyAsync(&last_flag, newflag.p + 3 * (size_t)nc - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  GT(hipStreamSynchronize(st));
  const int nf = nb + n_new + last_flag;
  Dev<int> fv, ftype, fper, flc, frc, edge_face, slot, nonper, bvo, worst;
  Dev<double> fc, fn, fmeas, bn, bmeas, ccx, ccy, area, mcarea;
  GT(fv.alloc(2 * (size_t)nf)); GT(ftype.alloc(nf)); GT(fper.alloc(nf)); GT(flc.alloc(nf));
  GT(frc.alloc(nf)); GT(edge_face.alloc(3 * (size_t)nc)); GT(slot.alloc(3 * (size_t)nc));
  GT(nonper.alloc(nc)); GT(bvo.alloc(2 * (size_t)nb)); GT(worst.alloc(1));
  GT(fc.alloc(2 * (size_t)nf)); GT(fn.alloc(2 * (size_t)nf)); GT(fmeas.alloc(nf));
  GT(bn.alloc(2 * (size_t)nb)); GT(bmeas.alloc(nb)); GT(ccx.alloc(nc)); GT(ccy.alloc(nc));
  GT(area.alloc(nc)); GT(mcarea.alloc(nv));
  GT(hipMemsetAsync(mcarea.p, 0, nv * sizeof(double), st));
  hipLaunchKernelGGL(k_faces, dim3(blocks(ne)), dim3(256), 0, st, ne, nb, bv.p, btag.p, bp.p, cv.p,
                     first.p, rank.p, lcell.p, rcell.p, fv.p, ftype.p, fper.p, flc.p, frc.p,
                     edge_face.p);
  hipLaunchKernelGGL(k_cells, dim3(blocks(nc)), dim3(256), 0, st, nc, nb, x.p, y.p, cv.p,
                     edge_face.p, fper.p, slot.p, nonper.p, ccx.p, ccy.p, area.p, bad.p);
  ts = tmp_size;
  GT(hipcub::DeviceReduce::Max(tmp.p, ts, bad.p, worst.p, (int)std::max<long long>(ne, nc), st));
  // vertex areas: (vertex, cell) pairs sorted by vertex, cells kept in order
  {
    Dev<int> cid, vs, cs;
    GT(cid.alloc(3 * (size_t)nc)); GT(vs.alloc(3 * (size_t)nc)); GT(cs.alloc(3 * (size_t)nc));
    std::vector<int> h(3 * (size_t)nc);
    for (long long k = 0; k < 3LL * nc; ++k) h[k] = (int)(k / 3);
    GT(hipMemcpyAsync(cid.p, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, st));
    ts = tmp_size;
    GT(hipcub::DeviceRadixSort::SortPairs(tmp.p, ts, cv.p, vs.p, cid.p, cs.p, 3 * nc, 0, vbits, st));
    hipLaunchKernelGGL(k_vertex_area, dim3(blocks(3LL * nc)), dim3(256), 0, st, 3 * nc, vs.p, cs.p,
                       area.p, mcarea.p);
    GT(hipStreamSynchronize(st));
  }
  hipLaunchKernelGGL(k_normals, dim3(blocks(nf)), dim3(256), 0, st, nf, nb, x.p, y.p, ccx.p, ccy.p,
                     fv.p, ftype.p, flc.p, frc.p, fc.p, fn.p, fmeas.p, bvo.p, bn.p, bmeas.p);
  GT(hipGetLastError());
  int code = 0;
  GT(hipMemcpyAsync(&code, worst.p, sizeof(int), hipMemcpyDeviceToHost, st));
  GT(hipStreamSynchronize(st));
  switch (code) {
    case G_OK: break;
    case G_FALLBACK: return 1;
    case G_BAD_EDGE: msg = "edge shared by more than two cells"; return TSM_E_MESH;
    case G_DUP_BFACE: case G_LOOSE_BFACE: msg = "boundary face not attached to any cell"; return TSM_E_MESH;
    default: msg = "a cell has non-positive area"; return TSM_E_MESH;
  }
  // ---- back to the host Geometry -------------------------------------------
  auto get = [&](auto* dst, const auto* src, size_t n) {
    return hipMemcpyAsync(dst, src, n * sizeof(*src), hipMemcpyDeviceToHost, st);
  };
  std::vector<int> hfv(2 * (size_t)nf), hft(nf), hfp(nf), hfl(nf), hfr(nf), hbv(2 * (size_t)nb);
  std::vector<double> hfc(2 * (size_t)nf), hfn(2 * (size_t)nf), hfm(nf), hbn(2 * (size_t)nb), hbm(nb);
  std::vector<int> hslot(3 * (size_t)nc), hnp(nc);
  g.nv = nv; g.nc = nc;
  g.x.assign(m.x, m.x + nv); g.y.assign(m.y, m.y + nv);
  g.cell.assign(m.cell_v, m.cell_v + 3 * (size_t)nc);
  g.ccx.resize(nc); g.ccy.resize(nc); g.cell_area.resize(nc); g.mcarea.resize(nv);
  GT(get(hfv.data(), fv.p, hfv.size())); GT(get(hft.data(), ftype.p, nf)); GT(get(hfp.data(), fper.p, nf));
  GT(get(hfl.data(), flc.p, nf)); GT(get(hfr.data(), frc.p, nf)); GT(get(hfc.data(), fc.p, hfc.size()));
  GT(get(hfn.data(), fn.p, hfn.size())); GT(get(hfm.data(), fmeas.p, nf));
  if (nb > 0) {
    GT(get(hbv.data(), bvo.p, hbv.size())); GT(get(hbn.data(), bn.p, hbn.size())); GT(get(hbm.data(), bmeas.p, nb));
  }
  GT(get(hslot.data(), slot.p, hslot.size())); GT(get(hnp.data(), nonper.p, nc));
  GT(get(g.ccx.data(), ccx.p, nc)); GT(get(g.ccy.data(), ccy.p, nc));
  GT(get(g.cell_area.data(), area.p, nc)); GT(get(g.mcarea.data(), mcarea.p, nv));
  GT(hipStreamSynchronize(st));
#undef GT
  g.face.resize(nf);
  for (int f = 0; f < nf; ++f) {
    Face& F = g.face[f];
    F.v[0] = hfv[2 * f]; F.v[1] = hfv[2 * f + 1];
    F.type = hft[f]; F.periodic = hfp[f] != 0; F.lcell = hfl[f]; F.rcell = hfr[f];
    F.cx = hfc[2 * f]; F.cy = hfc[2 * f + 1]; F.nx = hfn[2 * f]; F.ny = hfn[2 * f + 1];
    F.measure = hfm[f];
  }
  g.bface.resize(nb);
  for (int b = 0; b < nb; ++b) {
    BFace& B = g.bface[b];
    B.v[0] = hbv[2 * b]; B.v[1] = hbv[2 * b + 1];
    B.tag = m.bface_tag[b]; B.bc = btype[b]; B.periodic = bper[b] != 0; B.lcell = hfl[b];
    B.nx = hbn[2 * b]; B.ny = hbn[2 * b + 1]; B.measure = hbm[b];
  }
  g.cell_slot_face = std::move(hslot);
  g.cell_has_nonper_bface.assign(hnp.begin(), hnp.end());
  g.dcarea = g.mcarea;
  g.removed_faces = 0;
  // periodic sets: sums in member order (host loop, as in build_geometry)
  g.periodic.assign(m.n_periodic, {});
  g.periodic_of.assign(nv, -1);
  for (int i = 0; i < m.n_periodic; ++i) {
    for (int k = m.periodic_ptr[i]; k < m.periodic_ptr[i + 1]; ++k) {
      int v = m.periodic_v[k];
      if (v < 0 || v >= nv) { msg = "periodic vertex out of range"; return TSM_E_MESH; }
      if (g.periodic_of[v] != -1) { msg = "vertex in two periodic sets"; return TSM_E_MESH; }
      g.periodic_of[v] = i;
      g.periodic[i].push_back(v);
    }
    if (g.periodic[i].empty()) { msg = "empty periodic set"; return TSM_E_MESH; }
  }
  for (auto& P : g.periodic) {
    double mc = 0, dc = 0;
    for (int v : P) { mc += g.mcarea[v] * 1.0; dc += g.dcarea[v] * 1.0; }
    for (int v : P) { g.mcarea[v] = mc; g.dcarea[v] = dc; }
  }
  g.weight.assign(nv, 1.0);
  for (auto& P : g.periodic) {
    double w = 1.0 / ((double)P.size() * 1.0);
    for (int v : P) g.weight[v] = w;
  }
  return 0;
}

}  // namespace tsm

extern "C" int tsm_device_geometry(const tsm_mesh* mesh, int cell_type, int* fv, double* fn,
                                   int* bv, double* bn, double* dc, double* mc, int* nf, int* nb) {
  if (!mesh || !nf || !nb || cell_type != 0) return TSM_E_ARG;
  tsm::Geometry g;
  std::string msg;
  const int rc = tsm::build_geometry_device(*mesh, g, msg);
  if (rc == 1) return TSM_E_ARG;
  if (rc != 0) return rc;
  *nf = (int)g.face.size();
  *nb = (int)g.bface.size();
  for (size_t f = 0; f < g.face.size(); ++f) {
    if (fv) { fv[2 * f] = g.face[f].v[0]; fv[2 * f + 1] = g.face[f].v[1]; }
    if (fn) { fn[2 * f] = g.face[f].nx; fn[2 * f + 1] = g.face[f].ny; }
  }
  for (size_t b = 0; b < g.bface.size(); ++b) {
    if (bv) { bv[2 * b] = g.bface[b].v[0]; bv[2 * b + 1] = g.bface[b].v[1]; }
    if (bn) { bn[2 * b] = g.bface[b].nx; bn[2 * b + 1] = g.bface[b].ny; }
  }
  for (int i = 0; i < g.nv; ++i) {
    if (dc) dc[i] = g.dcarea[i];
    if (mc) mc[i] = g.mcarea[i];
  }
  return TSM_OK;
}
