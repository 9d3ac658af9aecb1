// tsm_tiled.hip -- LDS-staged stage kernels (one workgroup per vertex tile).
//
// Per SSP-RK3 stage two kernels run over all tiles (grid = tiles x samples):
//   k_grad_tile: W of the tile's owned + halo slots -> entropy variables in
//     LDS (prim2ent once per slot) -> Green-Gauss vertex gradients of the
//     owned vertices -> dE (src/fv.cc:699-1059);
//   k_fu_tile: W, dE, coordinates of the slots -> per-slot density, beta,
//     entropy (and sound speed) in LDS -> every local face flux into LDS ->
//     per owned vertex gather in reference face order, periodic set sums
//     through LDS, boundary fluxes, SSP-RK3 update (src/fv.cc:1066-1192,
//     :1635-1707).  The last stage also forms the residual-norm partials
//     (src/fv.cc:1801-1879) and the next step's local time step
//     (src/fv.cc:1499-1621) from the new state, so a step needs no separate
//     time-step pass.
// Every sum is taken in the same order as in the reference; no atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "tsm_device.h"
#include "tsm_physics.h"

namespace tsm {

namespace {

__device__ __forceinline__ Prim ld_prim(const double* W, size_t i) {
  const double4 q = reinterpret_cast<const double4*>(W)[i];
  return Prim{q.x, q.y, q.z, q.w};
}

template <int NV>
__device__ __forceinline__ void bsum(double* v, double* red) {
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    v[k] = x;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[wid * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = red[k];
      for (int w = 1; w < nw; ++w) s += red[w * NV + k];
      v[k] = s;
    }
  }
}
// XCD-aware tile order.  Workgroups go to the chip's 8 XCDs in turn (block b
// to XCD b mod 8), each XCD with its own L2.  Block b of a row of n tiles
// takes tile start(x) + b/8 of the contiguous range of XCD x = b mod 8, so
// every XCD walks a compact run of neighbouring tiles and the halo records
// it reads are mostly the owned records of tiles it has just staged (a
// bijection of [0, n): XCD x gets n/8 tiles, one more for the first n mod 8).
#ifndef TSM_XCDS
#define TSM_XCDS 8
#endif
__device__ __forceinline__ int xcd_tile(int b, int n) {
  if (TSM_XCDS <= 1) return b;
  const int x = b % TSM_XCDS, j = b / TSM_XCDS;
  const int q = n / TSM_XCDS, rem = n % TSM_XCDS;
  return x * q + min(x, rem) + j;
}

// Face entries of a vertex's gather held in registers (see k_fu_tile).
constexpr int GE = 12;

// One face entry of a vertex's gather: +F at the left end, -F at the right.
__device__ __forceinline__ void gather_face(unsigned e, const double* SF, int MF, double* acc) {
  const int j = e & 0x3fffu;
  const double q0 = SF[j], q1 = SF[MF + j], q2 = SF[2 * MF + j], q3 = SF[3 * MF + j];
  if (e & 0x8000u) { acc[0] -= q0; acc[1] -= q1; acc[2] -= q2; acc[3] -= q3; }
  else { acc[0] += q0; acc[1] += q1; acc[2] += q2; acc[3] += q3; }
}

// A face's share of a vertex's spectral radius, |u.n| + c|n|, with the
// vertex's own new state (src/fv.cc:1511-1531).
__device__ __forceinline__ double face_lambda(unsigned e, const double* SN, int MF, const Prim& w1,
                                              double cm) {
  const int j = e & 0x3fffu;
  const double nx = SN[j], ny = SN[MF + j], ar = SN[2 * MF + j];
  const double vn = w1.u * nx + w1.v * ny;
  return (fabs(vn) + cm * ar) * 1.0;
}

// Per-block partials of the last stage without a block barrier: every wave
// reduces its residual-norm sums and smallest local dt by shuffles, its lane 0
// leaves them in LDS and counts itself in with an LDS integer atomic; the
// last wave to arrive adds the waves' values in wave order and writes the
// block's slot.  Waves that finish early leave at once.
__device__ __forceinline__ void block_partials(double* part, double dtv, double* wred, int* arrive,
                                               double* partial, double* dtpart, size_t slot) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double x = part[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    part[k] = x;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) dtv = fmin(dtv, __shfl_down(dtv, o, 64));
  if ((threadIdx.x & 63) != 0) return;
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) wred[5 * w + k] = part[k];
  wred[5 * w + 4] = dtv;
  __threadfence_block();
  if (atomicAdd(arrive, 1) != nw - 1) return;
  __threadfence_block();
  double sum[4] = {wred[0], wred[1], wred[2], wred[3]};
  double m = wred[4];
  for (int u = 1; u < nw; ++u) {
#pragma unroll
    for (int k = 0; k < 4; ++k) sum[k] += wred[5 * u + k];
    m = fmin(m, wred[5 * u + 4]);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) partial[slot * 4 + k] = sum[k];
  dtpart[slot] = m;
}

__device__ __forceinline__ double bmin(double x, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmin(x, __shfl_down(x, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int w = 1; w < nw; ++w) x = fmin(x, red[w]);
  }
  return x;
}

// ssw: the face's Liou sensor, ssw[l] + ssw[r] (src/fv.cc:1081; KEPES-TeCNO Roe only)
template <int FLUX, int REC>
__device__ __forceinline__ Cons pair_flux(const Gas& gas, const FluxOpts& fo, double rx, double ry,
                                          const Side& L, const Side& R, const double* gdl,
                                          const double* gdr, double unx, double uny,
                                          double ssw = 0.0) {
  if (FLUX == F_ROE || FLUX == F_RUS)
    return tecno_flux<FLUX == F_ROE, REC>(gas, fo, rx, ry, L, R, gdl, gdr, unx, uny, ssw);
  if (FLUX == F_KEPEC) return kepec_flux_s(gas, L, R, unx, uny);
  if (FLUX == F_KEP) return kep_flux(gas, L.w, R.w, unx, uny);
  if (FLUX == F_ROEEC) return roe_ec_flux(gas, L.w, R.w, unx, uny);
  return avg_flux(gas, L.w, R.w, unx, uny);
}

// Boundary flux of bface b at a vertex with state wm (src/fv.cc:1100-1189)
template <int FLUX, int REC>
__device__ __forceinline__ Cons bflux(const DevGeo& G, const FluxOpts& fo, const Gas& gas, int b,
                                      int s, int v, const Prim& wm) {
  int type = G.btype[b];
  double nx = G.bn[2 * b], ny = G.bn[2 * b + 1];
  Prim s0 = wm, s1 = wm;
  bc_apply2(type, bc_values(G, s, b, bc_end(G, b, v)), nx, ny, s0, s1);
  if (type == TSM_BC_FARFIELD) {
    double area = fsqrt(nx * nx + ny * ny);
    double ia = fdiv(1.0, area);
    const double z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    constexpr bool TEC = FLUX == F_ROE || FLUX == F_RUS;
    Side L = make_side<TEC, FLUX == F_ROE>(gas, s0), R = make_side<TEC, FLUX == F_ROE>(gas, s1);
    Cons F = pair_flux<FLUX, REC>(gas, fo, 0.0, 0.0, L, R, z, z, nx * ia, ny * ia);
    F.m *= area; F.mx *= area; F.my *= area; F.e *= area;
    return F;
  }
  return inviscid_bflux(gas, s0, s1, nx, ny);
}

}  // namespace

// ---------------------------------------------------------------------------
// gradient kernel
// ---------------------------------------------------------------------------
// Boundary correction of one cell's gradient (src/fv.cc:926-966): for each
// non-periodic boundary face of the cell, n_b times the mean of the boundary
// states' entropy variables minus the mean of the vertex ones.
__device__ __forceinline__ void grad_bnd_correction(const DevGeo& G, const StageArgs& A,
                                                    const Gas& gas, int s, size_t off, int c,
                                                    double* g) {
  for (int z = 0; z < 2; ++z) {
    const int bf = G.cbf[2 * c + z];
    if (bf < 0) break;
    const int v0 = G.bv[2 * bf], v1 = G.bv[2 * bf + 1];
    const double nx = G.bn[2 * bf], ny = G.bn[2 * bf + 1];
    const int type = G.btype[bf];
    const Prim w0 = ld_prim(A.W_in, off + v0), w1 = ld_prim(A.W_in, off + v1);
    const Ent E0 = prim2ent(gas, w0), E1 = prim2ent(gas, w1);
    Ent q0 = E0, q1 = E1;
    if (bc_sets_state(type)) {
      Prim s0 = w0, s1 = w1;
      bc_apply1(type, bc_values(G, s, bf, 0), nx, ny, s0);
      bc_apply1(type, bc_values(G, s, bf, 1), nx, ny, s1);
      q0 = prim2ent(gas, s0);
      q1 = prim2ent(gas, s1);
    }
    const double dq[4] = {(q0.f + q1.f - E0.f - E1.f) / 2.0, (q0.vx + q1.vx - E0.vx - E1.vx) / 2.0,
                          (q0.vy + q1.vy - E0.vy - E1.vy) / 2.0, (q0.l + q1.l - E0.l - E1.l) / 2.0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      g[2 * q] += nx * dq[q];
      g[2 * q + 1] += ny * dq[q];
    }
  }
}

// Gradient of one cell from its corners' entropy variables and coordinates:
// N0 = dr12, N1 = dr20, N2 = dr01 rotated (src/grid_preproc.cc:336-362), then
// g_k = sum_j N_j V_k(j) (src/fv.cc:900-922).
__device__ __forceinline__ void cell_grad(double xa, double ya, double xb, double yb, double xd,
                                          double yd, const double* ea, const double* eb,
                                          const double* ed, double* g) {
  const double n0x = (yb - yd), n0y = -(xb - xd);
  const double n1x = (yd - ya), n1y = -(xd - xa);
  const double n2x = (ya - yb), n2y = -(xa - xb);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    g[2 * q] = n0x * ea[q] + n1x * eb[q] + n2x * ed[q];
    g[2 * q + 1] = n0y * ea[q] + n1y * eb[q] + n2y * ed[q];
  }
}

// Tiles [t0, t0 + gridDim.x).  PER: periodic tiles, with boundary corrections
// and the in-tile periodic set sums.  Otherwise vertices with a
// boundary-corrected cell are left to k_grad_bnd, which keeps this kernel's
// register count low enough for seven waves per SIMD.
template <bool PER>
__device__ __forceinline__ void grad_tile_body(const DevGeo& G, const TileGeo& TG,
                                               const StageArgs& A, int t, int s) {
  extern __shared__ double lds[];
  if (A.ctl[s].done) return;
  const int MS = TG.max_slots, TB = TG.tb;
  double* SE = lds;            // 4 x MS entropy variables
  double* SX = lds + 4 * MS;   // 2 x MS coordinates
  double* SP = lds;            // periodic tiles: 8 x TB exchange, after the cell loop
  const int b0 = TG.vbeg[t], nown = TG.vbeg[t + 1] - b0;
  const int h0 = TG.hbeg[t], ns = nown + TG.hbeg[t + 1] - h0;
  const size_t off = (size_t)s * G.nv;
  const Gas gas{A.gamma[s], A.gasR[s]};
  const int tid = threadIdx.x;
  const bool own = tid < nown;
  const int v = b0 + tid;
  // the owned vertex's cell range and scale, requested before the slot phase
  // so that their latency overlaps it
  int kc0 = 0, kc1 = 0;
  double fact = 0.0;
  if (own) {
    kc0 = TG.cptr[v];
    kc1 = TG.cptr[v + 1];
    fact = G.gfact[v];
  }
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    const int gid = (i < nown) ? b0 + i : TG.halo[h0 + i - nown];
    const Ent e = prim2ent(gas, ld_prim(A.W_in, off + gid));
    SE[i] = e.f; SE[MS + i] = e.vx; SE[2 * MS + i] = e.vy; SE[3 * MS + i] = e.l;
    SX[i] = G.x[gid]; SX[MS + i] = G.y[gid];
  }
  // the vertex's first CEP cell entries, requested before the barrier so that
  // the cell loop does not wait on them one by one
  constexpr int CEP = 8;
  unsigned ce[CEP];
#pragma unroll
  for (int u = 0; u < CEP; ++u) ce[u] = (kc0 + u < kc1) ? TG.cent[kc0 + u] : 0u;
  __syncthreads();
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  bool skip = false;
  if (own) {
    for (int k = kc0; k < kc1; ++k) {
      unsigned e = 0;
      if (k - kc0 < CEP) {
#pragma unroll
        for (int u = 0; u < CEP; ++u) if (u == k - kc0) e = ce[u];
      } else {
        e = TG.cent[k];
      }
      if (!PER && (e >> 30)) { skip = true; break; }
      const int a = e & 1023, b = (e >> 10) & 1023, d = (e >> 20) & 1023;
      const double ea[4] = {SE[a], SE[MS + a], SE[2 * MS + a], SE[3 * MS + a]};
      const double eb[4] = {SE[b], SE[MS + b], SE[2 * MS + b], SE[3 * MS + b]};
      const double ed[4] = {SE[d], SE[MS + d], SE[2 * MS + d], SE[3 * MS + d]};
      double g[8];
      cell_grad(SX[a], SX[MS + a], SX[b], SX[MS + b], SX[d], SX[MS + d], ea, eb, ed, g);
      if (PER && (e >> 30)) grad_bnd_correction(G, A, gas, s, off, TG.cgid[k], g);
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] += g[q];
    }
  }
  if (PER && TG.tper[t]) {  // periodic sets of this tile: sum members in set order
    __syncthreads();        // SE no longer read: SP reuses it
    if (own)
#pragma unroll
      for (int q = 0; q < 8; ++q) SP[q * TB + tid] = acc[q];
    __syncthreads();
    if (own) {
      const int pi = TG.pinfo[v];
      if (pi >= 0) {
        const int first = pi >> 8, cnt = pi & 255;
        double tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int m = 0; m < cnt; ++m)
#pragma unroll
          for (int q = 0; q < 8; ++q) tot[q] += SP[q * TB + first + m] * 1.0;
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = tot[q];
      }
    }
  }
  if (own && !skip) {
    double* dEv = A.dE + 8 * (off + v);
#pragma unroll
    for (int q = 0; q < 8; ++q) dEv[q] = acc[q] * fact;
  }
}
// Seven waves per SIMD for the tiles without periodic sets: the LDS of a
// 128-vertex tile (six slot rows) allows 14 tiles of two waves per CU, so the
// eighth wave is never resident and its registers go to the prefetched cell
// entries (measured at 1025^2: 62.1 us per stage against 65.8).
#ifndef TSM_GRAD_WAVES
#define TSM_GRAD_WAVES 7
#endif
template <bool PER>
__global__ __launch_bounds__(512, PER ? 4 : TSM_GRAD_WAVES) void k_grad_tile(DevGeo G, TileGeo TG, StageArgs A,
                                                                int t0) {
  grad_tile_body<PER>(G, TG, A, xcd_tile(blockIdx.x, gridDim.x) + t0, blockIdx.y);
}
// The same over the tiles of several contexts (one block per tile and sample
// of every level): PER selects the periodic tiles [0, np) of each level, or
// the others.
template <bool PER>
__global__ __launch_bounds__(512, PER ? 4 : TSM_GRAD_WAVES) void k_grad_tile_packed(const PackLevel* __restrict__ pk,
                                                                       int nl) {
  const int b = blockIdx.x;
  int l = 0;
  while (l + 1 < nl && b >= (PER ? pk[l + 1].base_gper : pk[l + 1].base_grest)) ++l;
  const PackLevel& L = pk[l];
  const int np = L.TG.n_periodic_tiles, n = PER ? np : L.TG.ntiles - np;
  const int local = b - (PER ? L.base_gper : L.base_grest);
  grad_tile_body<PER>(L.G, L.TG, L.A, local % n + (PER ? 0 : np), local / n);
}

// Gradients of the vertices k_grad_tile<false> leaves out (a cell with a
// boundary correction, outside periodic tiles): the same cell order and
// arithmetic from global memory.  8 lanes per vertex, one cell per lane, then
// the lane group's first lane adds the cell gradients in cell order (the
// order of the tile kernel), so the per-cell loads run side by side; each
// lane reaches its cell's operands through one GbEntry record.
constexpr int GB_LANES = 8;
__device__ __forceinline__ void grad_bnd_body(const DevGeo& G, const TileGeo& TG,
                                              const StageArgs& A, int nbv, int bx, int s) {
  const int i = (bx * blockDim.x + threadIdx.x) / GB_LANES, j = threadIdx.x % GB_LANES;
  if (A.ctl[s].done) return;
  const bool live = i < nbv;
  const int k0 = live ? TG.gb_ptr[i] : 0, nc = live ? TG.gb_ptr[i + 1] - k0 : 0;
  const size_t off = (size_t)s * G.nv;
  const Gas gas{A.gamma[s], A.gasR[s]};
  const int base = (threadIdx.x & 63) & ~(GB_LANES - 1);
  // the longest cell list of the lane group sets the chunk count for all its lanes
  int ncm = nc;
#pragma unroll
  for (int o = GB_LANES / 2; o > 0; o >>= 1) ncm = max(ncm, __shfl_xor(ncm, o, 64));
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int c0 = 0; c0 < ncm; c0 += GB_LANES) {  // chunks of GB_LANES cells, in cell order
    double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (c0 + j < nc) {
      const GbEntry E = TG.gb_ent[k0 + c0 + j];
      const Prim wa = ld_prim(A.W_in, off + E.a), wb = ld_prim(A.W_in, off + E.b),
                 wd = ld_prim(A.W_in, off + E.d);
      const Ent Ea = prim2ent(gas, wa), Eb = prim2ent(gas, wb), Ed = prim2ent(gas, wd);
      const double ea[4] = {Ea.f, Ea.vx, Ea.vy, Ea.l}, eb[4] = {Eb.f, Eb.vx, Eb.vy, Eb.l},
                   ed[4] = {Ed.f, Ed.vx, Ed.vy, Ed.l};
      cell_grad(G.x[E.a], G.y[E.a], G.x[E.b], G.y[E.b], G.x[E.d], G.y[E.d], ea, eb, ed, g);
      // boundary correction (src/fv.cc:926-966), as grad_bnd_correction with
      // the end states taken from the corners (the same loads and prim2ent)
      for (int z = 0; z < 2; ++z) {
        const int bf = z ? E.bf1 : E.bf0;
        if (bf < 0) break;
        const int i0 = (E.ends >> (4 * z)) & 3, i1 = (E.ends >> (4 * z + 2)) & 3;
        const Prim w0 = i0 == 0 ? wa : (i0 == 1 ? wb : wd);
        const Prim w1 = i1 == 0 ? wa : (i1 == 1 ? wb : wd);
        const Ent E0 = i0 == 0 ? Ea : (i0 == 1 ? Eb : Ed);
        const Ent E1 = i1 == 0 ? Ea : (i1 == 1 ? Eb : Ed);
        const double nx = G.bn[2 * bf], ny = G.bn[2 * bf + 1];
        const int type = G.btype[bf];
        Ent q0 = E0, q1 = E1;
        if (bc_sets_state(type)) {
          Prim s0 = w0, s1 = w1;
          bc_apply1(type, bc_values(G, s, bf, 0), nx, ny, s0);
          bc_apply1(type, bc_values(G, s, bf, 1), nx, ny, s1);
          q0 = prim2ent(gas, s0);
          q1 = prim2ent(gas, s1);
        }
        const double dq[4] = {(q0.f + q1.f - E0.f - E1.f) / 2.0,
                              (q0.vx + q1.vx - E0.vx - E1.vx) / 2.0,
                              (q0.vy + q1.vy - E0.vy - E1.vy) / 2.0,
                              (q0.l + q1.l - E0.l - E1.l) / 2.0};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          g[2 * q] += nx * dq[q];
          g[2 * q + 1] += ny * dq[q];
        }
      }
    }
    for (int m = 0; m < GB_LANES; ++m) {  // every lane takes part in the shuffles
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const double x = __shfl(g[q], base + m, 64);
        if (c0 + m < nc) acc[q] += x;
      }
    }
  }
  if (live && j == 0) {
    const int v = TG.gbv[i];
    const double fact = G.gfact[v];
    double* dEv = A.dE + 8 * (off + v);
#pragma unroll
    for (int q = 0; q < 8; ++q) dEv[q] = acc[q] * fact;
  }
}
__global__ __launch_bounds__(256) void k_grad_bnd(DevGeo G, TileGeo TG, StageArgs A, int nbv) {
  grad_bnd_body(G, TG, A, nbv, blockIdx.x, blockIdx.y);
}
constexpr int GB_BLOCK = 256;
__host__ __device__ inline int grad_bnd_blocks(int nbv) { return (nbv * GB_LANES + GB_BLOCK - 1) / GB_BLOCK; }
__global__ __launch_bounds__(256) void k_grad_bnd_packed(const PackLevel* __restrict__ pk, int nl) {
  const int b = blockIdx.x;
  int l = 0;
  while (l + 1 < nl && b >= pk[l + 1].base_gbnd) ++l;
  const PackLevel& L = pk[l];
  const int nbx = grad_bnd_blocks(L.TG.n_gbv), local = b - L.base_gbnd;
  grad_bnd_body(L.G, L.TG, L.A, L.TG.n_gbv, local % nbx, local / nbx);
}

// ---------------------------------------------------------------------------
// flux + gather + update kernel
// ---------------------------------------------------------------------------
template <int FLUX, int REC, bool NS = false>
struct SlotLayout {
  static constexpr bool TEC = FLUX == F_ROE || FLUX == F_RUS;
  static constexpr bool GRAD = TEC && REC != R_FIRST;
  // rows: T u v p | (viscous: x y E0..E3) | rho beta | V0..V3 (TeCNO) | c (Roe) |
  //       dE x8 (gradients) | x y (gradients without viscous terms).
  // The first KEEP rows stay valid for the gather phase; the rows after them
  // take the face fluxes once the flux phase is done.
  static constexpr int KEEP = 4 + (NS ? 6 : 0);
  static constexpr int QE = 6;  // viscous: entropy variables
  static constexpr int QR = KEEP, QB = KEEP + 1;
  static constexpr int QS = KEEP + 2, QC = QS + (TEC ? 4 : 0), QD = QC + (FLUX == F_ROE ? 1 : 0);
  static constexpr int QX = NS ? 4 : QD + 8;
  static constexpr int NQ = QD + (GRAD ? 8 : 0) + ((GRAD && !NS) ? 2 : 0);
};
// Faces per thread of the flux phase: the host launches at least
// ceil(max_faces / FPT) threads per tile.
constexpr int FPT = 2;

// Waves per SIMD the Euler flux kernels are compiled for: 4 caps them at 128
// VGPRs (the last stage's instance spills a few values); the Navier-Stokes
// instances keep 3 (at most 168 VGPRs), which holds their viscous terms in
// registers.
#ifndef TSM_FU_WAVES
#define TSM_FU_WAVES 4
#endif

// NS: 0 Euler; 1 Navier-Stokes; 2 Navier-Stokes on a mesh with wall cells
// (their terms need registers the other meshes should not pay for)
template <int FLUX, int REC, int LAST, int NS>
__device__ __forceinline__ void fu_tile_body(const DevGeo& G, const TileGeo& TG,
                                             const StageArgs& A, const StepParams& P,
                                             const double* Wn, double* dtpart, int r, int t,
                                             int s, int nbx) {
  using L = SlotLayout<FLUX, REC, (NS != 0)>;
  extern __shared__ double lds[];
  __shared__ double wred[8 * 5];
  __shared__ int arrive;
  if (A.ctl[s].done) return;
  const int MS = TG.max_slots, MF = TG.max_faces, TB = TG.tb;
  // LDS: slot rows Q (NQ x MS).  Once the face fluxes are in registers the
  // rows past the primitive state are dead and hold the face fluxes SF
  // (4 x MF) and, in the last stage, the face normals and areas SN (3 x MF); the
  // periodic exchange SR (4 x TB) reuses the primitive rows once every owned
  // state has been read.
  double* Q = lds;
  double* SF = lds + L::KEEP * MS;
  double* SN = SF + 4 * MF;
  double* SR = lds;
  const int b0 = TG.vbeg[t], nown = TG.vbeg[t + 1] - b0;
  const int h0 = TG.hbeg[t], ns = nown + TG.hbeg[t + 1] - h0;
  const int f0 = TG.fbeg[t], nfl = TG.fbeg[t + 1] - f0;
  const size_t off = (size_t)s * G.nv;
  const Gas gas{A.gamma[s], A.gasR[s]};

  if (LAST && threadIdx.x == 0) arrive = 0;  // read after the phase barriers
  // this thread's faces and its owned vertex's gather range, loaded before
  // the slot phase so that their latency overlaps it
  // (measured on the Rusanov TeCNO flux with the homogeneous limiters; the
  // other instances load them where they are used, as those run short of
  // registers or lose occupancy with them)
  constexpr bool PREF = FLUX == F_RUS && !NS && REC != R_ALBADA && REC != R_TVB;
  int gk0 = 0, gne = 0, gnf = 0;
  if (PREF && threadIdx.x < nown) {
    gk0 = TG.gptr[b0 + threadIdx.x];
    gne = TG.gptr[b0 + threadIdx.x + 1] - gk0;
  }
  unsigned flr[FPT];
  double fnx[FPT], fny[FPT];
#pragma unroll
  for (int q = 0; q < FPT; ++q) {
    const int j = threadIdx.x + q * blockDim.x;
    flr[q] = 0; fnx[q] = 1.0; fny[q] = 0.0;
    if (PREF && j < nfl) {
      flr[q] = TG.f_lr[f0 + j];
      fnx[q] = TG.f_n[2 * (size_t)(f0 + j)];
      fny[q] = TG.f_n[2 * (size_t)(f0 + j) + 1];
    }
  }
  // ---- phase 1: slot quantities ------------------------------------------
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    const int gid = (i < nown) ? b0 + i : TG.halo[h0 + i - nown];
    const Prim w = ld_prim(A.W_in, off + gid);
    const Side q = make_side<L::TEC, FLUX == F_ROE>(gas, w);
    Q[i] = w.T; Q[MS + i] = w.u; Q[2 * MS + i] = w.v; Q[3 * MS + i] = w.p;
    Q[L::QR * MS + i] = q.rho; Q[L::QB * MS + i] = q.beta;
    if (L::TEC)
#pragma unroll
      for (int k = 0; k < 4; ++k) Q[(L::QS + k) * MS + i] = q.V[k];
    // Roe: the sound speed, negated where the vertex's Liou sensor is set
    // (liou_fix; c > 0, so the sign carries the flag without an LDS row)
    if (FLUX == F_ROE) Q[L::QC * MS + i] = (A.ssw && A.ssw[off + gid] > 0.0) ? -q.c : q.c;
    if (L::GRAD)
#pragma unroll
      for (int k = 0; k < 8; ++k) Q[(L::QD + k) * MS + i] = A.dE[8 * (off + gid) + k];
    if (L::GRAD || NS) {
      Q[L::QX * MS + i] = G.x[gid];
      Q[(L::QX + 1) * MS + i] = G.y[gid];
    }
    if (NS) {  // entropy variables of the viscous cell fluxes (prim2ent)
      const Ent e = prim2ent(gas, w);
      Q[L::QE * MS + i] = e.f; Q[(L::QE + 1) * MS + i] = e.vx;
      Q[(L::QE + 2) * MS + i] = e.vy; Q[(L::QE + 3) * MS + i] = e.l;
    }
  }
  __syncthreads();

  // ---- phase 2: local face fluxes (registers, then LDS) ------------------
  double fl[FPT][4], fn[FPT][3];
#pragma unroll
  for (int q = 0; q < FPT; ++q) {
    const int j = threadIdx.x + q * blockDim.x;
    if (j >= nfl) continue;
    const unsigned lr = PREF ? flr[q] : TG.f_lr[f0 + j];
    const int l = lr & 0xffff, rr = lr >> 16;
    const double nx = PREF ? fnx[q] : TG.f_n[2 * (size_t)(f0 + j)];
    const double ny = PREF ? fny[q] : TG.f_n[2 * (size_t)(f0 + j) + 1];
    const double area = fsqrt(nx * nx + ny * ny);
    const double ia = fdiv(1.0, area);
    Side Ls, Rs;
    Ls.w = Prim{Q[l], Q[MS + l], Q[2 * MS + l], Q[3 * MS + l]};
    Rs.w = Prim{Q[rr], Q[MS + rr], Q[2 * MS + rr], Q[3 * MS + rr]};
    Ls.rho = Q[L::QR * MS + l]; Rs.rho = Q[L::QR * MS + rr];
    Ls.beta = Q[L::QB * MS + l]; Rs.beta = Q[L::QB * MS + rr];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      Ls.V[k] = L::TEC ? Q[(L::QS + k) * MS + l] : 0.0;
      Rs.V[k] = L::TEC ? Q[(L::QS + k) * MS + rr] : 0.0;
    }
    const double cl = (FLUX == F_ROE) ? Q[L::QC * MS + l] : 0.0;
    const double cr = (FLUX == F_ROE) ? Q[L::QC * MS + rr] : 0.0;
    Ls.c = fabs(cl);
    Rs.c = fabs(cr);
    const double ssw = (cl < 0.0 || cr < 0.0) ? 1.0 : 0.0;
    double gdl[4] = {0, 0, 0, 0}, gdr[4] = {0, 0, 0, 0}, rx = 0.0, ry = 0.0;
    if (L::GRAD) {
      rx = Q[L::QX * MS + rr] - Q[L::QX * MS + l];
      ry = Q[(L::QX + 1) * MS + rr] - Q[(L::QX + 1) * MS + l];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        gdl[k] = Q[(L::QD + 2 * k) * MS + l] * rx + Q[(L::QD + 2 * k + 1) * MS + l] * ry;
        gdr[k] = Q[(L::QD + 2 * k) * MS + rr] * rx + Q[(L::QD + 2 * k + 1) * MS + rr] * ry;
      }
    }
    const Cons F = pair_flux<FLUX, REC>(gas, A.fo, rx, ry, Ls, Rs, gdl, gdr, nx * ia, ny * ia, ssw);
    fl[q][0] = F.m * area; fl[q][1] = F.mx * area; fl[q][2] = F.my * area; fl[q][3] = F.e * area;
    fn[q][0] = nx; fn[q][1] = ny; fn[q][2] = area;
  }
  __syncthreads();  // every slot row read: rows 4.. become SF / SN
#pragma unroll
  for (int q = 0; q < FPT; ++q) {
    const int j = threadIdx.x + q * blockDim.x;
    if (j >= nfl) continue;
    SF[j] = fl[q][0]; SF[MF + j] = fl[q][1]; SF[2 * MF + j] = fl[q][2]; SF[3 * MF + j] = fl[q][3];
    if (LAST) { SN[j] = fn[q][0]; SN[MF + j] = fn[q][1]; SN[2 * MF + j] = fn[q][2]; }
  }
  __syncthreads();

  // ---- phase 3: gather, periodic sums, update ----------------------------
  const int tid = threadIdx.x;
  const bool own = tid < nown;
  const int v = b0 + tid;
  const int bb0 = TG.bbeg[t];
  double acc[4] = {0, 0, 0, 0};
  Prim w0{1, 0, 0, 1};
  if (own) {
    w0 = Prim{Q[tid], Q[MS + tid], Q[2 * MS + tid], Q[3 * MS + tid]};
    if (LAST && A.W_keep)  // the last stage's input state, kept for the caller
      reinterpret_cast<double4*>(A.W_keep)[off + v] = make_double4(w0.T, w0.u, w0.v, w0.p);
  }
  const bool per = TG.tper[t] != 0;
  if (per) __syncthreads();  // owned states read: SR may overwrite the primitive rows
  // A vertex's gather entries: its faces, then its boundary faces.  The first
  // GE face entries are loaded together into registers (and reused by the
  // last stage's time-step loop); boundary faces and longer lists follow from
  // memory, so every sum keeps the reference order.
  unsigned ge[GE];
  // the update's operands, requested before the gather so that their latency
  // overlaps it: W^n (U_old of every stage but the first SSP one) and the area
  // (not in the last stage, whose time-step loop needs the registers)
  double4 wn4 = make_double4(1.0, 0.0, 0.0, 1.0);
  double area_v = 1.0;
  if (own) {
    if (!PREF) {
      gk0 = TG.gptr[v];
      gne = TG.gptr[v + 1] - gk0;
    }
    if (!LAST && PREF) {
      if (r != 0 || A.scheme == 1) wn4 = reinterpret_cast<const double4*>(Wn)[off + v];
      area_v = G.dcarea[v];
    }
#pragma unroll
    for (int u = 0; u < GE; ++u) ge[u] = (u < gne) ? (unsigned)TG.gent[gk0 + u] : 0x4000u;
#pragma unroll
    for (int u = 0; u < GE; ++u)
      if (u < gne && !(ge[u] & 0x4000u)) { gather_face(ge[u], SF, MF, acc); gnf = u + 1; }
    for (int k = gnf; k < gne; ++k) {
      const unsigned e = TG.gent[gk0 + k];
      if (e & 0x4000u) {
        const int b = TG.tbf[bb0 + (e & 0x3fffu)];
        const Cons Fb = bflux<FLUX, REC>(G, A.fo, gas, b, s, v, w0);
        acc[0] += Fb.m * 0.5; acc[1] += Fb.mx * 0.5; acc[2] += Fb.my * 0.5; acc[3] += Fb.e * 0.5;
      } else {
        gather_face(e, SF, MF, acc);
      }
    }
    if (NS) {  // viscous cell fluxes after the faces: interior cells in cell order
               // (src/fv.cc:1197-1232), then wall cells once per bface in bface order
               // (src/fv.cc:1236-1312)
      const double fact = 1.0 / 3.0;
      for (int k = TG.cptr[v]; k < TG.cptr[v + 1]; ++k) {
        const unsigned e = TG.cent[k];
        if (NS == 2 && (e >> 30)) continue;  // wall cell
        const int sa = e & 1023, sb = (e >> 10) & 1023, sd = (e >> 20) & 1023;
        const int j = (sa == tid) ? 0 : ((sb == tid) ? 1 : 2);
        double ea[4], eb[4], ed[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          ea[q] = Q[(L::QE + q) * MS + sa]; eb[q] = Q[(L::QE + q) * MS + sb];
          ed[q] = Q[(L::QE + q) * MS + sd];
        }
        Ent sm;
        sm.f = ea[0] * fact + eb[0] * fact + ed[0] * fact;
        sm.vx = ea[1] * fact + eb[1] * fact + ed[1] * fact;
        sm.vy = ea[2] * fact + eb[2] * fact + ed[2] * fact;
        sm.l = ea[3] * fact + eb[3] * fact + ed[3] * fact;
        const double xa = Q[L::QX * MS + sa], ya = Q[(L::QX + 1) * MS + sa];
        const double xb = Q[L::QX * MS + sb], yb = Q[(L::QX + 1) * MS + sb];
        const double xd = Q[L::QX * MS + sd], yd = Q[(L::QX + 1) * MS + sd];
        double dE[8];
        cell_grad(xa, ya, xb, yb, xd, yd, ea, eb, ed, dE);
        const double cf = fdiv(1.0, 2.0 * G.carea[TG.cgid[k]]);
#pragma unroll
        for (int q = 0; q < 8; ++q) dE[q] *= cf;
        double nx, ny;
        if (j == 0) { nx = (yb - yd); ny = -(xb - xd); }
        else if (j == 1) { nx = (yd - ya); ny = -(xd - xa); }
        else { nx = (ya - yb); ny = -(xa - xb); }
        const Cons Fv = viscous_cell_flux(gas, A.vp, sm, dE, nx, ny);
        const double cm = -(0.5 * 1.0);
        acc[0] += Fv.m * cm; acc[1] += Fv.mx * cm; acc[2] += Fv.my * cm; acc[3] += Fv.e * cm;
      }
      if (NS == 2)
      for (int q = G.vbe_ptr[v]; q < G.vbe_ptr[v + 1]; ++q) {
        const unsigned ev = (unsigned)G.vbe[q];
        const int b = (int)(ev & 0x7fffffffu);
        const int cg = G.blc[b];
        int k = TG.cptr[v];  // the cell's entry among the vertex's cells (always present)
        const int kend = TG.cptr[v + 1] - 1;
        while (k < kend && TG.cgid[k] != cg) ++k;
        const unsigned e = TG.cent[k];
        const int sa = e & 1023, sb = (e >> 10) & 1023, sd = (e >> 20) & 1023;
        const int j = (sa == tid) ? 0 : ((sb == tid) ? 1 : 2);
        double ea[4], eb[4], ed[4];
#pragma unroll
        for (int z = 0; z < 4; ++z) {
          ea[z] = Q[(L::QE + z) * MS + sa]; eb[z] = Q[(L::QE + z) * MS + sb];
          ed[z] = Q[(L::QE + z) * MS + sd];
        }
        const double xa = Q[L::QX * MS + sa], ya = Q[(L::QX + 1) * MS + sa];
        const double xb = Q[L::QX * MS + sb], yb = Q[(L::QX + 1) * MS + sb];
        const double xd = Q[L::QX * MS + sd], yd = Q[(L::QX + 1) * MS + sd];
        double dE[8];
        cell_grad(xa, ya, xb, yb, xd, yd, ea, eb, ed, dE);
        grad_bnd_correction(G, A, gas, s, off, cg, dE);
        const double cf = fdiv(1.0, 2.0 * G.carea[cg]);
#pragma unroll
        for (int z = 0; z < 8; ++z) dE[z] *= cf;
        double nx, ny;
        if (j == 0) { nx = (yb - yd); ny = -(xb - xd); }
        else if (j == 1) { nx = (yd - ya); ny = -(xd - xa); }
        else { nx = (ya - yb); ny = -(xa - xb); }
        const Ent sw = wall_cell_state(G, A.W_in, off, s, gas, cg);
        const Cons Fv = viscous_cell_flux(gas, A.vp, sw, dE, nx, ny);
        const double cm = -(0.5 * 1.0);
        acc[0] += Fv.m * cm; acc[1] += Fv.mx * cm; acc[2] += Fv.my * cm; acc[3] += Fv.e * cm;
        if (ev >> 31) {  // the vertex is on bface b: the viscous wall flux
          const Cons Fw = viscous_cell_flux(gas, A.vp, sw, dE, G.bn[2 * b], G.bn[2 * b + 1],
                                            bc_adiabatic(G.btype[b]));
          const double cw = 0.5 * 1.0;
          acc[0] += Fw.m * cw; acc[1] += Fw.mx * cw; acc[2] += Fw.my * cw; acc[3] += Fw.e * cw;
        }
      }
    }
  }
  if (per) {
    if (own)
#pragma unroll
      for (int k = 0; k < 4; ++k) SR[k * TB + tid] = acc[k];
    __syncthreads();
    if (own) {
      const int pi = TG.pinfo[v];
      if (pi >= 0) {
        const int first = pi >> 8, cnt = pi & 255;
        double tot[4] = {0, 0, 0, 0};
        for (int m = 0; m < cnt; ++m)
#pragma unroll
          for (int k = 0; k < 4; ++k) tot[k] += SR[k * TB + first + m] * 1.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[k] = tot[k];
      }
    }
  }
  double part[4] = {0, 0, 0, 0};
  double lam = 0.0;
  Prim w1{1, 0, 0, 1};
  if (own) {
    const double area = (LAST || !PREF) ? G.dcarea[v] : area_v;
    if (LAST) {
      const double wt = G.weight[v];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double q = fdiv(acc[k], area);
        part[k] = (q * q) * wt;
      }
    }
    const double factor = fdiv(A.dtl ? A.dtl[off + v] : A.ctl[s].dt, 1.0 * area);
    Cons U;
    if (A.scheme == 1) {  // RK4 (src/fv.cc:1661-1673): U_old plus a multiple of R, or of
                          // the stages' weighted sum R2 in the last stage
      const double rks[4] = {1.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};
      double* R2 = A.R2 + 4 * (off + v);
      const double ir = 1.0 / rks[r];
      double q[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double r2 = (r == 0 ? 0.0 : R2[k]) + acc[k] * ir;
        R2[k] = r2;
        q[k] = (r < 3) ? acc[k] : r2;
      }
      const double c = (r < 3) ? -factor * rks[r < 3 ? r + 1 : 3] : -factor / 6.0;
      const Cons Uo = prim2con(gas, (PREF && !LAST) ? Prim{wn4.x, wn4.y, wn4.z, wn4.w}
                                                    : ld_prim(Wn, off + v));
      U.m = Uo.m * 1.0 + q[0] * c; U.mx = Uo.mx * 1.0 + q[1] * c;
      U.my = Uo.my * 1.0 + q[2] * c; U.e = Uo.e * 1.0 + q[3] * c;
    } else {  // SSP-RK3, or Heun's two-stage SSP form (src/fv.cc:1640-1651, :1675-1686)
      const double a_rk[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
      const double b_rk[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
      const bool heun = A.scheme == 2;
      const double b = heun ? (1.0 - r * 0.5) : b_rk[r], a = heun ? r * 0.5 : a_rk[r];
      const double c = -factor * b;
      U = prim2con(gas, w0);
      U.m *= b; U.mx *= b; U.my *= b; U.e *= b;
      if (r == 0) {  // a = 0: U_old * 0 contributes an exact zero
        U.m += 0.0 + acc[0] * c; U.mx += 0.0 + acc[1] * c;
        U.my += 0.0 + acc[2] * c; U.e += 0.0 + acc[3] * c;
      } else {
        const Cons Uo = prim2con(gas, (PREF && !LAST) ? Prim{wn4.x, wn4.y, wn4.z, wn4.w}
                                                      : ld_prim(Wn, off + v));
        U.m += Uo.m * a + acc[0] * c; U.mx += Uo.mx * a + acc[1] * c;
        U.my += Uo.my * a + acc[2] * c; U.e += Uo.e * a + acc[3] * c;
      }
    }
    w1 = con2prim(gas, U);
    reinterpret_cast<double4*>(A.W_out)[off + v] = make_double4(w1.T, w1.u, w1.v, w1.p);
    if (LAST) {  // spectral radius for the next step's dt, src/fv.cc:1511-1553
      const double cm = sound_speed(gas, w1);
#pragma unroll
      for (int u = 0; u < GE; ++u)
        if (u < gnf) lam += face_lambda(ge[u], SN, MF, w1, cm);
      for (int k = gnf; k < gne; ++k) {
        const unsigned e = TG.gent[gk0 + k];
        if (e & 0x4000u) {
          const int bf = TG.tbf[bb0 + (e & 0x3fffu)];
          const double nx = G.bn[2 * bf], ny = G.bn[2 * bf + 1];
          const double ar = fsqrt(nx * nx + ny * ny);
          const double vn = w1.u * nx + w1.v * ny;
          lam += 0.5 * (fabs(vn) + cm * ar);
        } else {
          lam += face_lambda(e, SN, MF, w1, cm);
        }
      }
      if (NS) {  // viscous share, after every inviscid term (src/fv.cc:1555-1580)
        double mu;
        const double T = w1.T;
        if (A.vp.mu_model == 0) mu = A.vp.mu_ref;
        else if (A.vp.mu_model == 1) mu = A.vp.mu_ref * pow(T, 1.5) / (T + A.vp.T_0);
        else mu = A.vp.mu_ref * pow(T / A.vp.T_ref, A.vp.omega);
        const double rho = density(gas, w1);
        for (int k = 0; k < gne; ++k) {
          const unsigned e = TG.gent[gk0 + k];
          if (e & 0x4000u) continue;
          const int j = e & 0x3fffu;
          const unsigned lr = TG.f_lr[f0 + j];
          const int sl = lr & 0xffff, sr = lr >> 16;
          const double dx = Q[L::QX * MS + sr] * 1.0 + Q[L::QX * MS + sl] * -1.0;
          const double dy = Q[(L::QX + 1) * MS + sr] * 1.0 + Q[(L::QX + 1) * MS + sl] * -1.0;
          const double dl = fsqrt(dx * dx + dy * dy);
          lam += fdiv(fdiv(mu * SN[2 * MF + j], dl), rho) * 1.0;
        }
      }
    }
  }
  if (LAST) {
    if (per) {
      __syncthreads();  // SR reads of the residual exchange are complete
      if (own) SR[tid] = lam;
      __syncthreads();
      if (own) {
        const int pi = TG.pinfo[v];
        if (pi >= 0) {
          const int first = pi >> 8, cnt = pi & 255;
          double tot = 0.0;
          for (int m = 0; m < cnt; ++m) tot = tot + SR[first + m] * 1.0;
          lam = tot;
        }
      }
    }
    double dtv = 1.0e20;
    if (own) {
      dtv = (P.time_step > 0.0) ? P.time_step : fdiv(P.cfl * G.dcarea[v], lam);
      if (A.dtl) A.dtl[off + v] = dtv;  // steady mode: read by this vertex only, next step
    }
    block_partials(part, dtv, wred, &arrive, A.partial, dtpart, (size_t)s * nbx + t);
  }
}
#ifndef TSM_FU_WAVES_LAST
#define TSM_FU_WAVES_LAST TSM_FU_WAVES
#endif
template <int FLUX, int REC, int LAST, int NS>
__global__ __launch_bounds__(512, NS ? 3 : (LAST ? TSM_FU_WAVES_LAST : TSM_FU_WAVES)) void k_fu_tile(DevGeo G, TileGeo TG,
                                                                        StageArgs A, StepParams P,
                                                                        const double* Wn,
                                                                        double* dtpart, int r) {
  fu_tile_body<FLUX, REC, LAST, NS>(G, TG, A, P, Wn, dtpart, r, xcd_tile(blockIdx.x, gridDim.x),
                                    blockIdx.y, gridDim.x);
}
// The same over the tiles and samples of several contexts (the levels of a
// multilevel run): block b belongs to the level whose range [base_fu, ...)
// holds it, tile (b - base) mod ntiles, sample (b - base) / ntiles.
template <int FLUX, int REC, int LAST, int NS>
__global__ __launch_bounds__(512, NS ? 3 : (LAST ? TSM_FU_WAVES_LAST : TSM_FU_WAVES)) void k_fu_tile_packed(
    const PackLevel* __restrict__ pk, int nl, int r) {
  const int b = blockIdx.x;
  int l = 0;
  while (l + 1 < nl && b >= pk[l + 1].base_fu) ++l;
  const PackLevel& L = pk[l];
  const int nt = L.TG.ntiles, local = b - L.base_fu;
  fu_tile_body<FLUX, REC, LAST, NS>(L.G, L.TG, L.A, L.P, L.Wn, L.dtpart, r, local % nt, local / nt,
                                    nt);
}

// ---------------------------------------------------------------------------
// fused stage kernel (meshes without periodic sets, schemes with gradients):
// gradient + flux + gather + update in one pass over a tile with two halo
// rings.  The owned and first-ring slots ("gradient slots", the first n1)
// get their Green-Gauss gradients from the cells around them, whose corners
// are the second ring; each first-ring gradient is the same sum, in the same
// order, as the one its owning tile forms, so the result equals the
// two-kernel path bit for bit and dE never goes through memory.
// ---------------------------------------------------------------------------
template <int FLUX, int REC>
struct FusedLayout {  // LDS rows: MS1 = gradient slots, MS2 = all slots
  static constexpr bool ROE = FLUX == F_ROE;
  // per gradient slot: T u v p | rho beta | V0..V3 | c (Roe) | dE x8
  static constexpr int R1 = 4 + 2 + 4 + (ROE ? 1 : 0) + 8;
  // per slot: entropy variables of the gradient (prim2ent) x4 | x y
  static constexpr int R2 = 6;
};

template <int FLUX, int REC, int LAST>
__global__ __launch_bounds__(512, TSM_FU_WAVES) void k_stage_fused(DevGeo G, TileGeo TG,
                                                                  StageArgs A, StepParams P,
                                                                  const double* Wn,
                                                                  double* dtpart, int r) {
  using FL = FusedLayout<FLUX, REC>;
  extern __shared__ double lds[];
  __shared__ double wred[8 * 5];
  __shared__ int arrive;
  const int t = xcd_tile(blockIdx.x, gridDim.x), s = blockIdx.y;
  if (A.ctl[s].done) return;
  const int MS1 = TG.max_slots1, MS2 = TG.max_slots, MF = TG.max_faces;
  double* Q = lds;                          // R1 x MS1
  double* QV = Q + 6 * MS1;                 // flux entropy variables
  double* QC = QV + 4 * MS1;                // Roe sound speed
  double* QD = QC + (FL::ROE ? MS1 : 0);    // dE
  double* SE = Q + FL::R1 * MS1;            // 4 x MS2 gradient entropy variables
  double* SX = SE + 4 * MS2;                // 2 x MS2 coordinates
  double* SF = lds + 4 * MS1;               // after the flux phase: 4 x MF fluxes
  double* SN = SF + 4 * MF;                 // last stage: 3 x MF normals, areas
  const int b0 = TG.vbeg[t], nown = TG.vbeg[t + 1] - b0;
  const int h0 = TG.hbeg[t], n1 = nown + TG.hbeg[t + 1] - h0;
  const int k0 = TG.h2beg[t], ns = n1 + TG.h2beg[t + 1] - k0;
  const int f0 = TG.fbeg[t], nfl = TG.fbeg[t + 1] - f0;
  const size_t off = (size_t)s * G.nv;
  const Gas gas{A.gamma[s], A.gasR[s]};

  if (LAST && threadIdx.x == 0) arrive = 0;  // read after the phase barriers
  // ---- phase 1: slot quantities (two rings) ------------------------------
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    const int gid = (i < nown) ? b0 + i : (i < n1 ? TG.halo[h0 + i - nown] : TG.halo2[k0 + i - n1]);
    const Prim w = ld_prim(A.W_in, off + gid);
    const Ent e = prim2ent(gas, w);
    SE[i] = e.f; SE[MS2 + i] = e.vx; SE[2 * MS2 + i] = e.vy; SE[3 * MS2 + i] = e.l;
    SX[i] = G.x[gid]; SX[MS2 + i] = G.y[gid];
    if (i < n1) {
      const Side q = make_side<true, FL::ROE>(gas, w);
      Q[i] = w.T; Q[MS1 + i] = w.u; Q[2 * MS1 + i] = w.v; Q[3 * MS1 + i] = w.p;
      Q[4 * MS1 + i] = q.rho; Q[5 * MS1 + i] = q.beta;
#pragma unroll
      for (int k = 0; k < 4; ++k) QV[k * MS1 + i] = q.V[k];
      if (FL::ROE) QC[i] = (A.ssw && A.ssw[off + gid] > 0.0) ? -q.c : q.c;  // sign: Liou sensor
    }
  }
  __syncthreads();

  // ---- phase 2: gradients of the gradient slots (src/fv.cc:887-996) -------
  const int gb = TG.g1beg[t];
  for (int k = threadIdx.x; k < n1; k += blockDim.x) {
    const int gid = (k < nown) ? b0 + k : TG.halo[h0 + k - nown];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = TG.g1ptr[gb + k]; c < TG.g1ptr[gb + k + 1]; ++c) {
      const unsigned e = TG.g1ent[c];
      const int a = e & 1023, b = (e >> 10) & 1023, d = (e >> 20) & 1023;
      const double xa = SX[a], ya = SX[MS2 + a], xb = SX[b], yb = SX[MS2 + b], xd = SX[d], yd = SX[MS2 + d];
      const double n0x = (yb - yd), n0y = -(xb - xd);
      const double n1x = (yd - ya), n1y = -(xd - xa);
      const double n2x = (ya - yb), n2y = -(xa - xb);
      double g[8];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double ea = SE[q * MS2 + a], eb = SE[q * MS2 + b], ed = SE[q * MS2 + d];
        g[2 * q] = n0x * ea + n1x * eb + n2x * ed;
        g[2 * q + 1] = n0y * ea + n1y * eb + n2y * ed;
      }
      if (e >> 30) {  // boundary correction, src/fv.cc:926-966
        const int cg = TG.g1cgid[c];
        for (int z = 0; z < 2; ++z) {
          const int bf = G.cbf[2 * cg + z];
          if (bf < 0) break;
          const int v0 = G.bv[2 * bf], v1 = G.bv[2 * bf + 1];
          const double nx = G.bn[2 * bf], ny = G.bn[2 * bf + 1];
          const int type = G.btype[bf];
          const Prim w0 = ld_prim(A.W_in, off + v0), w1 = ld_prim(A.W_in, off + v1);
          const Ent E0 = prim2ent(gas, w0), E1 = prim2ent(gas, w1);
          Ent q0 = E0, q1 = E1;
          if (bc_sets_state(type)) {
            Prim s0 = w0, s1 = w1;
            bc_apply1(type, bc_values(G, s, bf, 0), nx, ny, s0);
            bc_apply1(type, bc_values(G, s, bf, 1), nx, ny, s1);
            q0 = prim2ent(gas, s0);
            q1 = prim2ent(gas, s1);
          }
          const double dq[4] = {(q0.f + q1.f - E0.f - E1.f) / 2.0,
                                (q0.vx + q1.vx - E0.vx - E1.vx) / 2.0,
                                (q0.vy + q1.vy - E0.vy - E1.vy) / 2.0,
                                (q0.l + q1.l - E0.l - E1.l) / 2.0};
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            g[2 * q] += nx * dq[q];
            g[2 * q + 1] += ny * dq[q];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] += g[q];
    }
    const double fact = G.gfact[gid];
#pragma unroll
    for (int q = 0; q < 8; ++q) QD[q * MS1 + k] = acc[q] * fact;
  }
  __syncthreads();

  // ---- phase 3: local face fluxes (registers, then LDS) ------------------
  double fl[FPT][4], fn[FPT][3];
#pragma unroll
  for (int q = 0; q < FPT; ++q) {
    const int j = threadIdx.x + q * blockDim.x;
    if (j >= nfl) continue;
    const unsigned lr = TG.f_lr[f0 + j];
    const int l = lr & 0xffff, rr = lr >> 16;
    const double nx = TG.f_n[2 * (size_t)(f0 + j)], ny = TG.f_n[2 * (size_t)(f0 + j) + 1];
    const double area = fsqrt(nx * nx + ny * ny);
    const double ia = fdiv(1.0, area);
    Side Ls, Rs;
    Ls.w = Prim{Q[l], Q[MS1 + l], Q[2 * MS1 + l], Q[3 * MS1 + l]};
    Rs.w = Prim{Q[rr], Q[MS1 + rr], Q[2 * MS1 + rr], Q[3 * MS1 + rr]};
    Ls.rho = Q[4 * MS1 + l]; Rs.rho = Q[4 * MS1 + rr];
    Ls.beta = Q[5 * MS1 + l]; Rs.beta = Q[5 * MS1 + rr];
#pragma unroll
    for (int k = 0; k < 4; ++k) { Ls.V[k] = QV[k * MS1 + l]; Rs.V[k] = QV[k * MS1 + rr]; }
    const double cl = FL::ROE ? QC[l] : 0.0, cr = FL::ROE ? QC[rr] : 0.0;
    Ls.c = fabs(cl);
    Rs.c = fabs(cr);
    const double ssw = (cl < 0.0 || cr < 0.0) ? 1.0 : 0.0;
    const double rx = SX[rr] - SX[l], ry = SX[MS2 + rr] - SX[MS2 + l];
    double gdl[4], gdr[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      gdl[k] = QD[2 * k * MS1 + l] * rx + QD[(2 * k + 1) * MS1 + l] * ry;
      gdr[k] = QD[2 * k * MS1 + rr] * rx + QD[(2 * k + 1) * MS1 + rr] * ry;
    }
    const Cons F = pair_flux<FLUX, REC>(gas, A.fo, rx, ry, Ls, Rs, gdl, gdr, nx * ia, ny * ia, ssw);
    fl[q][0] = F.m * area; fl[q][1] = F.mx * area; fl[q][2] = F.my * area; fl[q][3] = F.e * area;
    fn[q][0] = nx; fn[q][1] = ny; fn[q][2] = area;
  }
  __syncthreads();  // every slot row read: rows past the primitive state become SF / SN
#pragma unroll
  for (int q = 0; q < FPT; ++q) {
    const int j = threadIdx.x + q * blockDim.x;
    if (j >= nfl) continue;
    SF[j] = fl[q][0]; SF[MF + j] = fl[q][1]; SF[2 * MF + j] = fl[q][2]; SF[3 * MF + j] = fl[q][3];
    if (LAST) { SN[j] = fn[q][0]; SN[MF + j] = fn[q][1]; SN[2 * MF + j] = fn[q][2]; }
  }
  __syncthreads();

  // ---- phase 4: gather and update (no periodic sets on these meshes) -----
  const int tid = threadIdx.x;
  const bool own = tid < nown;
  const int v = b0 + tid;
  const int bb0 = TG.bbeg[t];
  double acc[4] = {0, 0, 0, 0};
  double part[4] = {0, 0, 0, 0};
  double lam = 0.0;
  unsigned ge[GE];  // see k_fu_tile
  int gk0 = 0, gne = 0, gnf = 0;
  if (own) {
    const Prim w0 = Prim{Q[tid], Q[MS1 + tid], Q[2 * MS1 + tid], Q[3 * MS1 + tid]};
    if (LAST && A.W_keep)  // the last stage's input state, kept for the caller
      reinterpret_cast<double4*>(A.W_keep)[off + v] = make_double4(w0.T, w0.u, w0.v, w0.p);
    gk0 = TG.gptr[v];
    gne = TG.gptr[v + 1] - gk0;
#pragma unroll
    for (int u = 0; u < GE; ++u) ge[u] = (u < gne) ? (unsigned)TG.gent[gk0 + u] : 0x4000u;
#pragma unroll
    for (int u = 0; u < GE; ++u)
      if (u < gne && !(ge[u] & 0x4000u)) { gather_face(ge[u], SF, MF, acc); gnf = u + 1; }
    for (int k = gnf; k < gne; ++k) {
      const unsigned e = TG.gent[gk0 + k];
      if (e & 0x4000u) {
        const int b = TG.tbf[bb0 + (e & 0x3fffu)];
        const Cons Fb = bflux<FLUX, REC>(G, A.fo, gas, b, s, v, w0);
        acc[0] += Fb.m * 0.5; acc[1] += Fb.mx * 0.5; acc[2] += Fb.my * 0.5; acc[3] += Fb.e * 0.5;
      } else {
        gather_face(e, SF, MF, acc);
      }
    }
    const double area = G.dcarea[v];
    if (LAST) {
      const double wt = G.weight[v];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double q = fdiv(acc[k], area);
        part[k] = (q * q) * wt;
      }
    }
    const double a_rk[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
    const double b_rk[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
    const double factor = fdiv(A.dtl ? A.dtl[off + v] : A.ctl[s].dt, 1.0 * area);
    const double b = b_rk[r], a = a_rk[r], c = -factor * b_rk[r];
    Cons U = prim2con(gas, w0);
    U.m *= b; U.mx *= b; U.my *= b; U.e *= b;
    if (r == 0) {
      U.m += 0.0 + acc[0] * c; U.mx += 0.0 + acc[1] * c;
      U.my += 0.0 + acc[2] * c; U.e += 0.0 + acc[3] * c;
    } else {
      const Cons Uo = prim2con(gas, ld_prim(Wn, off + v));
      U.m += Uo.m * a + acc[0] * c; U.mx += Uo.mx * a + acc[1] * c;
      U.my += Uo.my * a + acc[2] * c; U.e += Uo.e * a + acc[3] * c;
    }
    const Prim w1 = con2prim(gas, U);
    reinterpret_cast<double4*>(A.W_out)[off + v] = make_double4(w1.T, w1.u, w1.v, w1.p);
    if (LAST) {  // spectral radius for the next step's dt, src/fv.cc:1511-1553
      const double cm = sound_speed(gas, w1);
#pragma unroll
      for (int u = 0; u < GE; ++u)
        if (u < gnf) lam += face_lambda(ge[u], SN, MF, w1, cm);
      for (int k = gnf; k < gne; ++k) {
        const unsigned e = TG.gent[gk0 + k];
        if (e & 0x4000u) {
          const int bf = TG.tbf[bb0 + (e & 0x3fffu)];
          const double nx = G.bn[2 * bf], ny = G.bn[2 * bf + 1];
          const double ar = fsqrt(nx * nx + ny * ny);
          const double vn = w1.u * nx + w1.v * ny;
          lam += 0.5 * (fabs(vn) + cm * ar);
        } else {
          lam += face_lambda(e, SN, MF, w1, cm);
        }
      }
    }
  }
  if (LAST) {
    double dtv = 1.0e20;
    if (own) {
      dtv = (P.time_step > 0.0) ? P.time_step : fdiv(P.cfl * G.dcarea[v], lam);
      if (A.dtl) A.dtl[off + v] = dtv;  // steady mode: read by this vertex only, next step
    }
    block_partials(part, dtv, wred, &arrive, A.partial, dtpart, (size_t)s * gridDim.x + t);
  }
}

// Residual norm, time advance, loop test and the next step's global dt
// (src/fv.cc:1591-1618, :1801-1879, :2273-2316).
__device__ __forceinline__ void step_final_body(const StageArgs& A, const StepParams& P, int nblk,
                                                int nv, const double* dtpart, int log_slot,
                                                int s) {
  __shared__ double red[64];
  Ctl& c = A.ctl[s];
  if (c.done) return;
  double part[4] = {0, 0, 0, 0};
  double x = 1.0e20;
  const double4* q4 = reinterpret_cast<const double4*>(A.partial) + (size_t)s * nblk;
  const double* dq = dtpart + (size_t)s * nblk;
  int i = threadIdx.x;
  for (; i + 3 * (int)blockDim.x < nblk; i += 4 * blockDim.x) {  // four independent loads in flight
    double4 a[4];
    double d[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { a[u] = q4[i + u * blockDim.x]; d[u] = dq[i + u * blockDim.x]; }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      part[0] += a[u].x; part[1] += a[u].y; part[2] += a[u].z; part[3] += a[u].w;
      x = fmin(x, d[u]);
    }
  }
  for (; i < nblk; i += blockDim.x) {
    const double4 a = q4[i];
    part[0] += a.x; part[1] += a.y; part[2] += a.z; part[3] += a.w;
    x = fmin(x, dq[i]);
  }
  bsum<4>(part, red);
  __syncthreads();
  x = bmin(x, red);
  if (threadIdx.x == 0) {
    double n[4];
    for (int k = 0; k < 4; ++k) {
      n[k] = sqrt(part[k]) / nv;
      c.rn[k] = n[k];
    }
    const double tot = sqrt((n[0] * n[0]) + (n[1] * n[1] + n[2] * n[2]) + (n[3] * n[3]));
    if (c.iter == 0) {
      c.res0 = tot;
      if (c.res0 == 0.0) c.res0 = 1.0;
    }
    c.res_total = tot / c.res0;
    c.iter += 1;
    c.t += c.dt;
    if (c.reached && P.unsteady) { c.counter += 1; c.reached = 0; }
    if (A.log && log_slot >= 0) {
      double* Lg = A.log + ((size_t)s * A.log_cap + log_slot) * 4;
      Lg[0] = (double)c.iter; Lg[1] = c.dt; Lg[2] = c.t; Lg[3] = c.res_total;
    }
    const bool go = c.res_total > P.min_residue && (P.unsteady ? true : c.iter < P.max_iter) &&
                    c.t < P.final_time;
    if (!go) {
      c.done = 1;
    } else {
      double dt = x;
      if (P.unsteady && !(P.time_step > 0.0)) {
        const double target = A.tstamps[min(c.counter, P.n_stamps)];
        if (c.t + dt > target) { dt = target - c.t; c.reached = 1; }
      }
      c.dt = dt;
    }
  }
}
__global__ __launch_bounds__(1024) void k_step_final_tiled(StageArgs A, StepParams P, int nblk,
                                                          int nv, const double* dtpart,
                                                          int log_slot) {
  step_final_body(A, P, nblk, nv, dtpart, log_slot, blockIdx.x);
}
__global__ __launch_bounds__(1024) void k_step_final_packed(const PackLevel* __restrict__ pk, int nl) {
  const int b = blockIdx.x;
  int l = 0;
  while (l + 1 < nl && b >= pk[l + 1].base_final) ++l;
  const PackLevel& L = pk[l];
  step_final_body(L.A, L.P, L.TG.ntiles, L.G.nv, L.dtpart, -1, b - L.base_final);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int tile_grad_block(const TileGeo& T) {
  const int b = ((T.tb + 63) / 64) * 64;
  return std::max(b, T.grad_block);
}

size_t tile_grad_lds(const TileGeo& T) {
  return sizeof(double) * std::max(6 * (size_t)T.max_slots, 8 * (size_t)T.tb);
}

namespace {
template <int FLUX, int REC, bool NS = false>
size_t fu_lds(const TileGeo& T, bool) {
  using L = SlotLayout<FLUX, REC, NS>;
  const size_t MS = T.max_slots, MF = T.max_faces;
  return sizeof(double) * (L::KEEP * MS + std::max((size_t)(L::NQ - L::KEEP) * MS, 7 * MF) + 16);
}
template <int FLUX, int REC>
size_t fused_lds(const TileGeo& T) {
  using FL = FusedLayout<FLUX, REC>;
  const size_t MS1 = T.max_slots1, MS2 = T.max_slots, MF = T.max_faces;
  return sizeof(double) *
         (4 * MS1 + std::max((size_t)(FL::R1 - 4) * MS1 + (size_t)FL::R2 * MS2, 7 * MF) + 16);
}
template <int FLUX, int REC>
void fu_launch(const DevGeo& G, const TileGeo& T, const StageArgs& A, const StepParams& P,
               const double* Wn, double* dtpart, int S, int r, bool last, hipStream_t st) {
  constexpr bool GRAD = SlotLayout<FLUX, REC>::GRAD;
  if (GRAD && T.fused) {  // SSP-RK3 only (the host builds two-ring tiles for it alone)
    const size_t lds = fused_lds<FLUX, REC>(T);
    dim3 grid(T.ntiles, S), blk(tile_fu_block(T));
    if (last) {
      auto k = k_stage_fused<FLUX, REC, 1>;
      (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      hipLaunchKernelGGL(k, grid, blk, lds, st, G, T, A, P, Wn, dtpart, r);
    } else {
      auto k = k_stage_fused<FLUX, REC, 0>;
      (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      hipLaunchKernelGGL(k, grid, blk, lds, st, G, T, A, P, Wn, dtpart, r);
    }
    return;
  }
  dim3 grid(T.ntiles, S), blk(tile_fu_block(T));
  auto go = [&](auto kern, size_t lds) {
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, grid, blk, lds, st, G, T, A, P, Wn, dtpart, r);
  };
  if (A.ns && A.ns_walls) {
    const size_t lds = fu_lds<FLUX, REC, true>(T, last);
    if (last) go(k_fu_tile<FLUX, REC, 1, 2>, lds);
    else go(k_fu_tile<FLUX, REC, 0, 2>, lds);
  } else if (A.ns) {
    const size_t lds = fu_lds<FLUX, REC, true>(T, last);
    if (last) go(k_fu_tile<FLUX, REC, 1, 1>, lds);
    else go(k_fu_tile<FLUX, REC, 0, 1>, lds);
  } else {
    const size_t lds = fu_lds<FLUX, REC, false>(T, last);
    if (last) go(k_fu_tile<FLUX, REC, 1, 0>, lds);
    else go(k_fu_tile<FLUX, REC, 0, 0>, lds);
  }
}
template <int FLUX>
void fu_rec(int rec, const DevGeo& G, const TileGeo& T, const StageArgs& A, const StepParams& P,
            const double* Wn, double* dtpart, int S, int r, bool last, hipStream_t st) {
  switch (rec) {
    case R_FIRST: fu_launch<FLUX, R_FIRST>(G, T, A, P, Wn, dtpart, S, r, last, st); break;
    case R_SECOND: fu_launch<FLUX, R_SECOND>(G, T, A, P, Wn, dtpart, S, r, last, st); break;
    case R_MUSCL: fu_launch<FLUX, R_MUSCL>(G, T, A, P, Wn, dtpart, S, r, last, st); break;
    case R_ALBADA: fu_launch<FLUX, R_ALBADA>(G, T, A, P, Wn, dtpart, S, r, last, st); break;
    case R_MINMOD: fu_launch<FLUX, R_MINMOD>(G, T, A, P, Wn, dtpart, S, r, last, st); break;
    case R_TVB: fu_launch<FLUX, R_TVB>(G, T, A, P, Wn, dtpart, S, r, last, st); break;
    default: fu_launch<FLUX, R_ENO2>(G, T, A, P, Wn, dtpart, S, r, last, st); break;
  }
}
}  // namespace

int tile_fu_block(const TileGeo& T) {
  int b = std::max(T.tb, (T.max_faces + FPT - 1) / FPT);
  b = ((b + 63) / 64) * 64;
  return std::max(b, T.fu_block);
}
int tile_fu_block_max() { return 512; }

size_t tile_fu_lds_ns(const TileGeo& T, int flux, int rec) {
  if (flux == F_ROE) return rec == R_FIRST ? fu_lds<F_ROE, R_FIRST, true>(T, true) : fu_lds<F_ROE, R_MINMOD, true>(T, true);
  if (flux == F_RUS) return rec == R_FIRST ? fu_lds<F_RUS, R_FIRST, true>(T, true) : fu_lds<F_RUS, R_MINMOD, true>(T, true);
  return fu_lds<F_KEPEC, R_FIRST, true>(T, true);
}

size_t tile_fu_lds(const TileGeo& T, int flux, int rec) {
  if (flux == F_ROE) {
    if (rec == R_FIRST) return fu_lds<F_ROE, R_FIRST>(T, true);
    return T.fused ? fused_lds<F_ROE, R_MINMOD>(T) : fu_lds<F_ROE, R_MINMOD>(T, true);
  }
  if (flux == F_RUS) {
    if (rec == R_FIRST) return fu_lds<F_RUS, R_FIRST>(T, true);
    return T.fused ? fused_lds<F_RUS, R_MINMOD>(T) : fu_lds<F_RUS, R_MINMOD>(T, true);
  }
  return fu_lds<F_KEPEC, R_FIRST>(T, true);
}

bool tile_needs_gradient_pass(const TileGeo& T, int flux, int rec) {
  const bool grad = (flux == F_ROE || flux == F_RUS) && rec != R_FIRST;
  return grad && !T.fused;
}

// A gradient pass of at most two tiles per CU cannot fill the GPU, so the
// seven-wave occupancy of k_grad_tile<false> buys nothing there, while each
// launch costs several microseconds: on meshes whose periodic sets would split
// the pass into two launches (periodic tiles, then the rest) and that leave no
// vertex to k_grad_bnd, such grids take one launch of the periodic variant
// (its set exchange is skipped for tiles without sets; the same sums in the
// same order).  Measured on the 129^2 periodic vortex: 63 us per step against
// 82.  With boundary vertices the split stays: forming their corrections
// inside the boundary tiles makes those tiles the pass's critical path (93 us
// per step on the 129^2 Riemann mesh against 112).
static bool gradient_in_one_launch(const TileGeo& T, int S) {
  return T.n_gbv == 0 && T.n_periodic_tiles > 0 && (long long)T.ntiles * S <= 512;
}

void launch_tile_gradient(const DevGeo& G, const TileGeo& T, const StageArgs& A, int S,
                          hipStream_t st) {
  const size_t lds = tile_grad_lds(T);
  if (gradient_in_one_launch(T, S)) {
    (void)hipFuncSetAttribute((const void*)k_grad_tile<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_grad_tile<true>, dim3(T.ntiles, S), dim3(tile_grad_block(T)), lds, st, G, T,
                       A, 0);
    return;
  }
  const int np = T.n_periodic_tiles, nrest = T.ntiles - np;
  if (np > 0) {
    (void)hipFuncSetAttribute((const void*)k_grad_tile<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_grad_tile<true>, dim3(np, S), dim3(tile_grad_block(T)), lds, st, G, T, A, 0);
  }
  if (nrest > 0) {
    (void)hipFuncSetAttribute((const void*)k_grad_tile<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_grad_tile<false>, dim3(nrest, S), dim3(tile_grad_block(T)), lds, st, G, T,
                       A, np);
  }

}

namespace {
__global__ void k_gb_count(int n, const int* gbv, const int* cptr, int* cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cnt[i] = cptr[gbv[i] + 1] - cptr[gbv[i]];
}
__global__ void k_gb_fill(int n, const int* gbv, const int* cptr, const int* cgid, const int* cv,
                          const int* cbf, const int* bv, const int* ptr, GbEntry* out, int* bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = gbv[i];
  int o = ptr[i];
  for (int k = cptr[v]; k < cptr[v + 1]; ++k) {
    const int c = cgid[k];
    GbEntry E;
    E.a = cv[3 * c]; E.b = cv[3 * c + 1]; E.d = cv[3 * c + 2];
    E.bf0 = cbf[2 * c];
    E.bf1 = E.bf0 >= 0 ? cbf[2 * c + 1] : -1;
    E.ends = 0; E.pad0 = E.pad1 = 0;
    for (int z = 0; z < 2; ++z) {
      const int bf = z ? E.bf1 : E.bf0;
      if (bf < 0) break;
      for (int e = 0; e < 2; ++e) {
        const int u = bv[2 * bf + e];
        const int q = u == E.a ? 0 : (u == E.b ? 1 : (u == E.d ? 2 : 3));
        if (q == 3) atomicOr(bad, 1);
        E.ends |= (q & 3) << (2 * (2 * z + e));
      }
    }
    out[o++] = E;
  }
}
}  // namespace

int build_grad_bnd_table(const DevGeo& G, TileGeo& T, const std::function<int(void**, size_t)>& dalloc,
                         hipStream_t st, std::string& msg) {
  const int n = T.n_gbv;
  if (n <= 0) return TSM_OK;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess) msg = std::string("boundary gradient table: ") + hipGetErrorString(e);
    return e == hipSuccess;
  };
  int* cnt = nullptr;
  int* ptr = nullptr;
  int* bad = nullptr;
  GbEntry* ent = nullptr;
  if (dalloc((void**)&cnt, (size_t)n * sizeof(int)) || dalloc((void**)&ptr, (size_t)(n + 1) * sizeof(int)) ||
      dalloc((void**)&bad, sizeof(int))) {
    msg = "boundary gradient table: allocation";
    return TSM_E_HIP;
  }
  hipLaunchKernelGGL(k_gb_count, dim3((n + 255) / 256), dim3(256), 0, st, n, T.gbv, T.cptr, cnt);
  std::vector<int> h(n), hp(n + 1, 0);
  if (!chk(hipMemcpyAsync(h.data(), cnt, n * sizeof(int), hipMemcpyDeviceToHost, st)) ||
      !chk(hipStreamSynchronize(st)))
    return TSM_E_HIP;
  for (int i = 0; i < n; ++i) hp[i + 1] = hp[i] + h[i];
  if (dalloc((void**)&ent, std::max<size_t>(hp[n], 1) * sizeof(GbEntry))) {
    msg = "boundary gradient table: allocation";
    return TSM_E_HIP;
  }
  int hbad = 0;
  if (!chk(hipMemcpyAsync(ptr, hp.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice, st)) ||
      !chk(hipMemsetAsync(bad, 0, sizeof(int), st)))
    return TSM_E_HIP;
  hipLaunchKernelGGL(k_gb_fill, dim3((n + 255) / 256), dim3(256), 0, st, n, T.gbv, T.cptr, T.cgid,
                     G.cv, G.cbf, G.bv, ptr, ent, bad);
  if (!chk(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st)) ||
      !chk(hipStreamSynchronize(st)))
    return TSM_E_HIP;
  if (hbad) {
    msg = "boundary gradient table: a boundary face end is not a corner of its cell";
    return TSM_E_MESH;
  }
  T.gb_ptr = ptr;
  T.gb_ent = ent;
  return TSM_OK;
}

void launch_tile_gradient_bnd(const DevGeo& G, const TileGeo& T, const StageArgs& A, int S,
                              hipStream_t st) {
  if (T.n_gbv > 0)
    hipLaunchKernelGGL(k_grad_bnd, dim3((T.n_gbv * GB_LANES + 255) / 256, S), dim3(256), 0, st, G,
                       T, A, T.n_gbv);
}

void launch_tile_stage(int flux, int rec, const DevGeo& G, const TileGeo& T, const StageArgs& A,
                       const StepParams& P, const double* Wn, double* dtpart, int S, int r,
                       bool last, hipStream_t st) {
  switch (flux) {
    case F_ROE: fu_rec<F_ROE>(rec, G, T, A, P, Wn, dtpart, S, r, last, st); break;
    case F_RUS: fu_rec<F_RUS>(rec, G, T, A, P, Wn, dtpart, S, r, last, st); break;
    case F_KEPEC: fu_launch<F_KEPEC, R_FIRST>(G, T, A, P, Wn, dtpart, S, r, last, st); break;
    case F_ROEEC: fu_launch<F_ROEEC, R_FIRST>(G, T, A, P, Wn, dtpart, S, r, last, st); break;
    case F_KEP: fu_launch<F_KEP, R_FIRST>(G, T, A, P, Wn, dtpart, S, r, last, st); break;
    default: fu_launch<F_AVG, R_FIRST>(G, T, A, P, Wn, dtpart, S, r, last, st); break;
  }
}

void launch_step_final_tiled(const DevGeo& G, const StageArgs& A, const StepParams& P, int S,
                             int nblk, const double* dtpart, int log_slot, hipStream_t st) {
  hipLaunchKernelGGL(k_step_final_tiled, dim3(S), dim3(1024), 0, st, A, P, nblk, G.nv, dtpart,
                     log_slot);
}

// ---------------------------------------------------------------------------
// packed launches: every kernel of a step runs once over the tiles and
// samples of all levels (contexts with the same scheme), each block doing
// exactly what it does in its own context's launch
// ---------------------------------------------------------------------------
void pack_plan(PackLevel* lv, int nl, int flux, int rec, PackPlan& p) {
  p = PackPlan{};
  p.nl = nl;
  for (int l = 0; l < nl; ++l) {
    PackLevel& L = lv[l];
    const TileGeo& T = L.TG;
    const int np = T.n_periodic_tiles;
    L.base_fu = p.fu_blocks;         p.fu_blocks += T.ntiles * L.S;
    L.base_gper = p.gper_blocks;     p.gper_blocks += np * L.S;
    L.base_grest = p.grest_blocks;   p.grest_blocks += (T.ntiles - np) * L.S;
    L.base_gbnd = p.gbnd_blocks;     p.gbnd_blocks += (T.n_gbv > 0 ? grad_bnd_blocks(T.n_gbv) : 0) * L.S;
    L.base_final = p.final_blocks;   p.final_blocks += L.S;
    p.fu_block = std::max(p.fu_block, tile_fu_block(T));
    p.grad_block = std::max(p.grad_block, tile_grad_block(T));
    p.fu_lds = std::max(p.fu_lds, tile_fu_lds(T, flux, rec));
    p.grad_lds = std::max(p.grad_lds, tile_grad_lds(T));
    p.grad_pass = tile_needs_gradient_pass(T, flux, rec);
  }
}

void launch_packed_gradient(const PackLevel* d, const PackPlan& p, hipStream_t st) {
  if (!p.grad_pass) return;
  if (p.gbnd_blocks > 0)
    hipLaunchKernelGGL(k_grad_bnd_packed, dim3(p.gbnd_blocks), dim3(GB_BLOCK), 0, st, d, p.nl);
  if (p.gper_blocks > 0) {
    (void)hipFuncSetAttribute((const void*)k_grad_tile_packed<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.grad_lds);
    hipLaunchKernelGGL(k_grad_tile_packed<true>, dim3(p.gper_blocks), dim3(p.grad_block), p.grad_lds,
                       st, d, p.nl);
  }
  if (p.grest_blocks > 0) {
    (void)hipFuncSetAttribute((const void*)k_grad_tile_packed<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.grad_lds);
    hipLaunchKernelGGL(k_grad_tile_packed<false>, dim3(p.grest_blocks), dim3(p.grad_block),
                       p.grad_lds, st, d, p.nl);
  }
}

namespace {
template <int FLUX, int REC>
void packed_fu(const PackLevel* d, const PackPlan& p, int r, bool last, hipStream_t st) {
  auto go = [&](auto kern) {
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)p.fu_lds);
    hipLaunchKernelGGL(kern, dim3(p.fu_blocks), dim3(p.fu_block), p.fu_lds, st, d, p.nl, r);
  };
  if (last) go(k_fu_tile_packed<FLUX, REC, 1, 0>);
  else go(k_fu_tile_packed<FLUX, REC, 0, 0>);
}
template <int FLUX>
void packed_fu_rec(int rec, const PackLevel* d, const PackPlan& p, int r, bool last,
                   hipStream_t st) {
  switch (rec) {
    case R_FIRST: packed_fu<FLUX, R_FIRST>(d, p, r, last, st); break;
    case R_SECOND: packed_fu<FLUX, R_SECOND>(d, p, r, last, st); break;
    case R_MUSCL: packed_fu<FLUX, R_MUSCL>(d, p, r, last, st); break;
    case R_ALBADA: packed_fu<FLUX, R_ALBADA>(d, p, r, last, st); break;
    case R_MINMOD: packed_fu<FLUX, R_MINMOD>(d, p, r, last, st); break;
    case R_TVB: packed_fu<FLUX, R_TVB>(d, p, r, last, st); break;
    default: packed_fu<FLUX, R_ENO2>(d, p, r, last, st); break;
  }
}
}  // namespace

void launch_packed_stage(int flux, int rec, const PackLevel* d, const PackPlan& p, int r,
                         bool last, hipStream_t st) {
  switch (flux) {
    case F_ROE: packed_fu_rec<F_ROE>(rec, d, p, r, last, st); break;
    case F_RUS: packed_fu_rec<F_RUS>(rec, d, p, r, last, st); break;
    case F_KEPEC: packed_fu<F_KEPEC, R_FIRST>(d, p, r, last, st); break;
    case F_ROEEC: packed_fu<F_ROEEC, R_FIRST>(d, p, r, last, st); break;
    case F_KEP: packed_fu<F_KEP, R_FIRST>(d, p, r, last, st); break;
    default: packed_fu<F_AVG, R_FIRST>(d, p, r, last, st); break;
  }
}

void launch_packed_final(const PackLevel* d, const PackPlan& p, hipStream_t st) {
  hipLaunchKernelGGL(k_step_final_packed, dim3(p.final_blocks), dim3(1024), 0, st, d, p.nl);
}

}  // namespace tsm
