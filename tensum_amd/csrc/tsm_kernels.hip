// tsm_kernels.hip -- gfx950 kernels of the FV residual + RK stage loop.
//
// Layout: every per-vertex quantity is structure-of-arrays FP64, one array of
// n_samples*n_vertex values per component (sample-major), so a wavefront of
// 64 consecutive vertices reads 512 contiguous bytes per component.  The
// sample index is the grid's y dimension.  Residuals are summed by gathering
// along a vertex CSR whose entries are in the reference's face order, so each
// vertex adds its face fluxes in exactly the order the reference's face-loop
// scatter does (src/fv.cc:1071-1097); no float atomics are used anywhere and
// two runs give bit-identical results.
#include <hip/hip_runtime.h>

#include "tsm_device.h"
#include "tsm_physics.h"
#include "tprog.h"

namespace tsm {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
// Vertex arrays are array-of-structs: 4 doubles (W, E, U_old, R) or 8 doubles
// (dE) per vertex, so one halo gather touches one or two cache lines.
__device__ __forceinline__ Prim load_prim(const double* W, size_t i) {
  const double4 q = reinterpret_cast<const double4*>(W)[i];
  return Prim{q.x, q.y, q.z, q.w};
}
__device__ __forceinline__ void store_prim(double* W, size_t i, const Prim& w) {
  reinterpret_cast<double4*>(W)[i] = make_double4(w.T, w.u, w.v, w.p);
}
__device__ __forceinline__ Ent load_ent(const double* E, size_t i) {
  const double4 q = reinterpret_cast<const double4*>(E)[i];
  return Ent{q.x, q.y, q.z, q.w};
}
__device__ __forceinline__ void store_ent(double* E, size_t i, const Ent& e) {
  reinterpret_cast<double4*>(E)[i] = make_double4(e.f, e.vx, e.vy, e.l);
}

// Block-wide deterministic sum / min of NV values per thread (blockDim 256).
template <int NV>
__device__ __forceinline__ void block_sum(double* v, double* lds) {
  // wave-level tree (64 lanes), fixed pattern
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    v[k] = x;
  }
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[wid * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    int nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = lds[k];
      for (int w = 1; w < nw; ++w) s += lds[w * NV + k];
      v[k] = s;
    }
  }
}
__device__ __forceinline__ double block_min(double x, double* lds) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmin(x, __shfl_down(x, off, 64));
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    int nw = blockDim.x >> 6;
    for (int w = 1; w < nw; ++w) x = fmin(x, lds[w]);
  }
  return x;
}

// ---------------------------------------------------------------------------
// Cell gradient G_c = sum_j N_cj V(v_j) with the boundary correction of
// src/fv.cc:887-966 (unscaled: G_c = 2 A_c grad V).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cell_gradient(const DevGeo& G, const StageArgs& A, const Gas& gas,
                                              int c, bool corr, double* g) {
  const size_t off = A.off;
  int a = G.cv[3 * c], b = G.cv[3 * c + 1], d = G.cv[3 * c + 2];
  double xa = G.x[a], ya = G.y[a], xb = G.x[b], yb = G.y[b], xd = G.x[d], yd = G.y[d];
  // cell normals, src/grid_preproc.cc:336-362: N0 = dr12, N1 = dr20, N2 = dr01
  double n0x = (yb - yd), n0y = -(xb - xd);
  double n1x = (yd - ya), n1y = -(xd - xa);
  double n2x = (ya - yb), n2y = -(xa - xb);
  Ent ea = load_ent(A.E_in, off + a), eb = load_ent(A.E_in, off + b),
      ed = load_ent(A.E_in, off + d);
  const double qa[4] = {ea.f, ea.vx, ea.vy, ea.l};
  const double qb[4] = {eb.f, eb.vx, eb.vy, eb.l};
  const double qd[4] = {ed.f, ed.vx, ed.vy, ed.l};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    g[2 * k] = n0x * qa[k] + n1x * qb[k] + n2x * qd[k];
    g[2 * k + 1] = n0y * qa[k] + n1y * qb[k] + n2y * qd[k];
  }
  if (corr) {
    for (int t = 0; t < 2; ++t) {
      int bf = G.cbf[2 * c + t];
      if (bf < 0) break;
      int v0 = G.bv[2 * bf], v1 = G.bv[2 * bf + 1];
      double nx = G.bn[2 * bf], ny = G.bn[2 * bf + 1];
      int type = G.btype[bf];
      Ent E0 = load_ent(A.E_in, off + v0), E1 = load_ent(A.E_in, off + v1);
      Ent q0 = E0, q1 = E1;
      // a boundary that leaves the state as it is contributes nothing, unless
      // strong boundary values made the stored entropy variables stale
      if (bc_sets_state(type) || G.strong) {
        Prim s0 = load_prim(A.W_in, off + v0), s1 = load_prim(A.W_in, off + v1);
        bc_apply1(type, bc_values(G, A.sample, bf, 0), nx, ny, s0);
        bc_apply1(type, bc_values(G, A.sample, bf, 1), nx, ny, s1);
        q0 = prim2ent(gas, s0);
        q1 = prim2ent(gas, s1);
      }
      const double dq[4] = {(q0.f + q1.f - E0.f - E1.f) / 2.0, (q0.vx + q1.vx - E0.vx - E1.vx) / 2.0,
                            (q0.vy + q1.vy - E0.vy - E1.vy) / 2.0, (q0.l + q1.l - E0.l - E1.l) / 2.0};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        g[2 * k] += nx * dq[k];
        g[2 * k + 1] += ny * dq[k];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// K1: vertex gradients (src/fv.cc:699-1059, src/reduce_gradient_periodic.cc)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gradient(DevGeo G, StageArgs A, int s_only) {
  const int s = s_only >= 0 ? s_only : (int)blockIdx.y;
  if (!A.ignore_done && A.ctl[s].done) return;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= G.nv) return;
  StageArgs As = A;
  As.off = (size_t)s * G.nv;
  As.sample = s;
  Gas gas{A.gamma[s], A.gasR[s]};
  int p = G.pset[v];
  int m0, m1;
  if (p < 0) { m0 = -1; m1 = 0; } else { m0 = G.pptr[p]; m1 = G.pptr[p + 1]; }
  double tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int mi = m0; mi < m1; ++mi) {
    int m = (mi < 0) ? v : G.pv[mi];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int e = G.vc_ptr[m]; e < G.vc_ptr[m + 1]; ++e) {
      int ce = G.vc[e];
      int c = ce & 0x7fffffff;
      double g[8];
      cell_gradient(G, As, gas, c, ce < 0, g);
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] += g[k];
    }
    if (p < 0) {
#pragma unroll
      for (int k = 0; k < 8; ++k) tot[k] = acc[k];
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k) tot[k] += acc[k] * 1.0;
    }
  }
  double fact = G.gfact[v];
  double* dE = A.dE + 8 * (As.off + v);
#pragma unroll
  for (int k = 0; k < 8; ++k) dE[k] = tot[k] * fact;
}

// ---------------------------------------------------------------------------
// K2: edge fluxes (src/fv.cc:1071-1097, src/material.cc:27-74)
// ---------------------------------------------------------------------------
template <int FLUX, int REC>
__global__ __launch_bounds__(256) void k_flux(DevGeo G, StageArgs A) {
  const int s = blockIdx.y;
  if (!A.ignore_done && A.ctl[s].done) return;
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= G.nf) return;
  const size_t off = (size_t)s * G.nv;
  Gas gas{A.gamma[s], A.gasR[s]};
  int l = G.fv[2 * f], r = G.fv[2 * f + 1];
  double nx = G.fn[2 * f], ny = G.fn[2 * f + 1];
  double area = sqrt(nx * nx + ny * ny);
  double ia = 1.0 / area;
  double unx = nx * ia, uny = ny * ia;
  Prim Wl = load_prim(A.W_in, off + l), Wr = load_prim(A.W_in, off + r);
  Cons F;
  if (FLUX == F_ROE || FLUX == F_RUS) {
    double gl[8], gr[8];
    if (REC != R_FIRST) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        gl[k] = A.dE[8 * (off + l) + k];
        gr[k] = A.dE[8 * (off + r) + k];
      }
    }
    double rx = G.x[r] - G.x[l], ry = G.y[r] - G.y[l];
    double ssw = 0.0;
    if (FLUX == F_ROE && A.ssw) ssw = A.ssw[off + l] + A.ssw[off + r];
    Side Ls = make_side<true, FLUX == F_ROE>(gas, Wl), Rs = make_side<true, FLUX == F_ROE>(gas, Wr);
    double gdl[4] = {0, 0, 0, 0}, gdr[4] = {0, 0, 0, 0};
    if (REC != R_FIRST) {
      project_grad(gl, rx, ry, gdl);
      project_grad(gr, rx, ry, gdr);
    }
    F = tecno_flux<FLUX == F_ROE, REC>(gas, A.fo, rx, ry, Ls, Rs, gdl, gdr, unx, uny, ssw);
  } else if (FLUX == F_KEPEC) {
    F = kepec_flux(gas, Wl, Wr, unx, uny);
  } else if (FLUX == F_KEP) {
    F = kep_flux(gas, Wl, Wr, unx, uny);
  } else if (FLUX == F_ROEEC) {
    F = roe_ec_flux(gas, Wl, Wr, unx, uny);
  } else {
    F = avg_flux(gas, Wl, Wr, unx, uny);
  }
  const size_t fs = A.fs, fo = (size_t)s * G.nf + f;
  A.F[fo] = F.m * area;
  A.F[fs + fo] = F.mx * area;
  A.F[2 * fs + fo] = F.my * area;
  A.F[3 * fs + fo] = F.e * area;
}

// Boundary flux contribution of bface b at its vertex m (src/fv.cc:1100-1189)
template <int FLUX, int REC>
__device__ __forceinline__ Cons bface_flux(const DevGeo& G, const StageArgs& A, const Gas& gas,
                                           int b, int m, const Prim& wm, double xm, double ym) {
  int type = G.btype[b];
  double nx = G.bn[2 * b], ny = G.bn[2 * b + 1];
  Prim s0 = wm, s1 = wm;
  bc_apply2(type, bc_values(G, A.sample, b, bc_end(G, b, m)), nx, ny, s0, s1);
  if (type == TSM_BC_FARFIELD) {
    double area = sqrt(nx * nx + ny * ny);
    double ia = 1.0 / area;
    double unx = nx * ia, uny = ny * ia;
    Cons F;
    if (FLUX == F_ROE || FLUX == F_RUS) {
      const double z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      Side L0 = make_side<true, FLUX == F_ROE>(gas, s0), L1 = make_side<true, FLUX == F_ROE>(gas, s1);
      F = tecno_flux<FLUX == F_ROE, REC>(gas, A.fo, 0.0 * xm, 0.0 * ym, L0, L1, z, z, unx, uny, 0.0);
    } else if (FLUX == F_KEPEC) {
      F = kepec_flux(gas, s0, s1, unx, uny);
    } else if (FLUX == F_KEP) {
      F = kep_flux(gas, s0, s1, unx, uny);
    } else if (FLUX == F_ROEEC) {
      F = roe_ec_flux(gas, s0, s1, unx, uny);
    } else {
      F = avg_flux(gas, s0, s1, unx, uny);
    }
    F.m *= area; F.mx *= area; F.my *= area; F.e *= area;
    return F;
  }
  return inviscid_bflux(gas, s0, s1, nx, ny);
}

// Viscous flux of cell c delivered to its local vertex slot j
// (src/fv.cc:1197-1232, src/material.cc:158-230); a wall cell uses its wall
// state and, for a vertex on its bface b, also returns the wall flux through b
// (src/fv.cc:1268-1308, src/material.cc:236-302).
__device__ __forceinline__ Cons viscous_part(const DevGeo& G, const StageArgs& A, const Gas& gas,
                                             int c, int j, int wall_b = -1, Cons* Fw = nullptr) {
  const size_t off = A.off;
  int a = G.cv[3 * c], b = G.cv[3 * c + 1], d = G.cv[3 * c + 2];
  Ent s;
  if (G.cnb[c]) {
    s = wall_cell_state(G, A.W_in, A.off, A.sample, gas, c);
  } else {
    Ent ea = load_ent(A.E_in, off + a), eb = load_ent(A.E_in, off + b),
        ed = load_ent(A.E_in, off + d);
    const double fact = 1.0 / 3.0;
    s.f = ea.f * fact + eb.f * fact + ed.f * fact;
    s.vx = ea.vx * fact + eb.vx * fact + ed.vx * fact;
    s.vy = ea.vy * fact + eb.vy * fact + ed.vy * fact;
    s.l = ea.l * fact + eb.l * fact + ed.l * fact;
  }
  double dE[8];
  cell_gradient(G, A, gas, c, G.cnb[c] != 0, dE);
  double cf = 1.0 / (2.0 * G.carea[c]);
#pragma unroll
  for (int k = 0; k < 8; ++k) dE[k] *= cf;
  double xa = G.x[a], ya = G.y[a], xb = G.x[b], yb = G.y[b], xd = G.x[d], yd = G.y[d];
  double nx, ny;
  if (j == 0) { nx = (yb - yd); ny = -(xb - xd); }
  else if (j == 1) { nx = (yd - ya); ny = -(xd - xa); }
  else { nx = (ya - yb); ny = -(xa - xb); }
  if (wall_b >= 0)
    *Fw = viscous_cell_flux(gas, A.vp, s, dE, G.bn[2 * wall_b], G.bn[2 * wall_b + 1],
                            bc_adiabatic(G.btype[wall_b]));
  return viscous_cell_flux(gas, A.vp, s, dE, nx, ny);
}

// ---------------------------------------------------------------------------
// K3: residual gather + (optionally) RK update
//   MODE 0: write R only; MODE 1: ssprk3 stage; MODE 2: rk4 stage
// (src/fv.cc:1066-1192 gather side, src/reduce_residual_periodic.cc,
//  src/fv.cc:1635-1707 update, src/fv.cc:1801-1879 norm partials)
// ---------------------------------------------------------------------------
template <int FLUX, int REC, int MODE>
__global__ __launch_bounds__(256) void k_update(DevGeo G, StageArgs A, int r, int last) {
  __shared__ double lds[4 * 4];
  const int s = blockIdx.y;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t off = (size_t)s * G.nv;
  if (!A.ignore_done && A.ctl[s].done) {
    // a finished sample keeps its state across the buffer swap
    if (MODE != 0 && v < G.nv) {
      store_prim(A.W_out, off + v, load_prim(A.W_in, off + v));
      store_ent(A.E_out, off + v, load_ent(A.E_in, off + v));
    }
    return;
  }
  StageArgs As = A;
  As.off = off;
  As.sample = s;
  Gas gas{A.gamma[s], A.gasR[s]};
  double part[4] = {0, 0, 0, 0};
  if (v < G.nv) {
    int p = G.pset[v];
    int m0, m1;
    if (p < 0) { m0 = -1; m1 = 0; } else { m0 = G.pptr[p]; m1 = G.pptr[p + 1]; }
    double tot[4] = {0, 0, 0, 0};
    const size_t fs = A.fs, fo = (size_t)s * G.nf;
    for (int mi = m0; mi < m1; ++mi) {
      int m = (mi < 0) ? v : G.pv[mi];
      double acc[4] = {0, 0, 0, 0};
      Prim wm;
      bool have_wm = false;
      for (int e = G.vfe_ptr[m]; e < G.vfe_ptr[m + 1]; ++e) {
        unsigned ent = (unsigned)G.vfe[e];
        unsigned kind = ent >> 30, idx = ent & 0x3fffffffu;
        if (kind < 2) {
          const double* F = A.F + fo + idx;
          double f0 = F[0], f1 = F[fs], f2 = F[2 * fs], f3 = F[3 * fs];
          if (kind == 0) { acc[0] += f0; acc[1] += f1; acc[2] += f2; acc[3] += f3; }
          else { acc[0] -= f0; acc[1] -= f1; acc[2] -= f2; acc[3] -= f3; }
        } else {
          if (!have_wm) { wm = load_prim(A.W_in, off + m); have_wm = true; }
          Cons Fb = bface_flux<FLUX, REC>(G, As, gas, (int)idx, m, wm, G.x[m], G.y[m]);
          acc[0] += Fb.m * 0.5; acc[1] += Fb.mx * 0.5; acc[2] += Fb.my * 0.5; acc[3] += Fb.e * 0.5;
        }
      }
      if (A.ns) {
        // interior cells in cell order, then wall cells once per bface in
        // bface order (src/fv.cc:1199-1232, :1236-1312)
        for (int e = G.vc_ptr[m]; e < G.vc_ptr[m + 1]; ++e) {
          if (G.vc[e] < 0) continue;  // wall cell
          int c = G.vc[e];
          int j = (G.cv[3 * c] == m) ? 0 : ((G.cv[3 * c + 1] == m) ? 1 : 2);
          Cons Fv = viscous_part(G, As, gas, c, j);
          const double cm = -(0.5 * 1.0);
          acc[0] += Fv.m * cm; acc[1] += Fv.mx * cm; acc[2] += Fv.my * cm; acc[3] += Fv.e * cm;
        }
        for (int e = G.vbe_ptr[m]; e < G.vbe_ptr[m + 1]; ++e) {
          const unsigned ent = (unsigned)G.vbe[e];
          const int b = (int)(ent & 0x7fffffffu);
          const bool on = (ent >> 31) != 0;
          const int c = G.blc[b];
          const int j = (G.cv[3 * c] == m) ? 0 : ((G.cv[3 * c + 1] == m) ? 1 : 2);
          Cons Fw;
          Cons Fv = viscous_part(G, As, gas, c, j, on ? b : -1, &Fw);
          const double cm = -(0.5 * 1.0);
          acc[0] += Fv.m * cm; acc[1] += Fv.mx * cm; acc[2] += Fv.my * cm; acc[3] += Fv.e * cm;
          if (on) {
            const double cw = 0.5 * 1.0;
            acc[0] += Fw.m * cw; acc[1] += Fw.mx * cw; acc[2] += Fw.my * cw; acc[3] += Fw.e * cw;
          }
        }
      }
      if (p < 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) tot[k] = acc[k];
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) tot[k] += acc[k] * 1.0;
      }
    }
    // bc_scheme strong: the last stage zeroes the momentum residual of noslip
    // end points before its norm and update (src/fv.cc:1805-1819)
    if (MODE != 0 && last && G.snoslip && G.snoslip[v]) { tot[1] = 0.0; tot[2] = 0.0; }
    if (MODE == 0) {
      double* R = A.R + 4 * (off + v);
#pragma unroll
      for (int k = 0; k < 4; ++k) R[k] = tot[k];
    } else {
      double area = G.dcarea[v];
      if (last) {
        double w = G.weight[v];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          double q = tot[k] / area;
          part[k] = (q * q) * w;
        }
      }
      // steady mode advances each vertex with its own dt (src/fv.cc:1591-1618)
      double dt = A.dtl ? A.dtl[off + v] : A.ctl[s].dt;
      double factor = dt / (1.0 * area);
      Prim w0 = load_prim(A.W_in, off + v);
      Cons U;
      const double* Uo = A.Uold + 4 * (off + v);
      if (MODE == 1) {
        const double a_rk[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
        const double b_rk[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
        U = prim2con(gas, w0);
        // SSP-RK3, or Heun's method in the two-stage SSP form its coefficients
        // describe (src/fv.cc:1675-1686 with sadd for the overwriting equ)
        double b = A.heun ? (1.0 - r * 0.5) : b_rk[r], a = A.heun ? r * 0.5 : a_rk[r];
        double c = -factor * b;
        U.m *= b; U.mx *= b; U.my *= b; U.e *= b;
        U.m += Uo[0] * a + tot[0] * c;
        U.mx += Uo[1] * a + tot[1] * c;
        U.my += Uo[2] * a + tot[2] * c;
        U.e += Uo[3] * a + tot[3] * c;
      } else {
        const double rks[4] = {1.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};
        double* R2 = A.R2 + 4 * (off + v);
        double ir = 1.0 / rks[r];
        double q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          double r2 = (r == 0 ? 0.0 : R2[k]) + tot[k] * ir;
          R2[k] = r2;
          q[k] = (r < 3) ? tot[k] : r2;
        }
        double c = (r < 3) ? -factor * rks[r < 3 ? r + 1 : 3] : -factor / 6.0;
        U.m = Uo[0] * 1.0 + q[0] * c;
        U.mx = Uo[1] * 1.0 + q[1] * c;
        U.my = Uo[2] * 1.0 + q[2] * c;
        U.e = Uo[3] * 1.0 + q[3] * c;
      }
      Prim w1 = con2prim(gas, U);
      store_prim(A.W_out, off + v, w1);
      store_ent(A.E_out, off + v, prim2ent(gas, w1));
      if (last && A.W_keep) store_prim(A.W_keep, off + v, w0);
    }
  }
  if (MODE != 0 && last) {
    block_sum<4>(part, lds);
    if (threadIdx.x == 0) {
      double* P = A.partial + ((size_t)s * gridDim.x + blockIdx.x) * 4;
#pragma unroll
      for (int k = 0; k < 4; ++k) P[k] = part[k];
    }
  }
}

// ---------------------------------------------------------------------------
// K0: store U_old + local time step + block min (src/fv.cc:1499-1630)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dt(DevGeo G, StageArgs A, StepParams P) {
  __shared__ double lds[4];
  const int s = blockIdx.y;
  Ctl& c = A.ctl[s];
  // loop condition of src/fv.cc:2273-2275, evaluated at the start of a step
  bool done = !(c.res_total > P.min_residue && (P.unsteady ? true : c.iter < P.max_iter) &&
                c.t < P.final_time);
  if (A.ignore_done) done = false;
  if (done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) c.done = 1;
    return;
  }
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t off = (size_t)s * G.nv;
  Gas gas{A.gamma[s], A.gasR[s]};
  double dtv = 1.0e20;
  if (v < G.nv) {
    Prim w = load_prim(A.W_in, off + v);
    Cons U = prim2con(gas, w);
    reinterpret_cast<double4*>(A.Uold)[off + v] = make_double4(U.m, U.mx, U.my, U.e);
    if (P.time_step > 0.0) {
      dtv = P.time_step;
    } else {
      int p = G.pset[v];
      int m0, m1;
      if (p < 0) { m0 = -1; m1 = 0; } else { m0 = G.pptr[p]; m1 = G.pptr[p + 1]; }
      double tot = 0.0;
      for (int mi = m0; mi < m1; ++mi) {
        int m = (mi < 0) ? v : G.pv[mi];
        Prim wm = (m == v) ? w : load_prim(A.W_in, off + m);
        double cm = sound_speed(gas, wm);
        double lam = 0.0;
        for (int e = G.vfe_ptr[m]; e < G.vfe_ptr[m + 1]; ++e) {
          unsigned ent = (unsigned)G.vfe[e];
          unsigned kind = ent >> 30, idx = ent & 0x3fffffffu;
          if (kind < 2) {
            double nx = G.fn[2 * idx], ny = G.fn[2 * idx + 1];
            double area = sqrt(nx * nx + ny * ny);
            double vn = wm.u * nx + wm.v * ny;
            lam += (fabs(vn) + cm * area) * 1.0;
          } else {
            double nx = G.bn[2 * idx], ny = G.bn[2 * idx + 1];
            double area = sqrt(nx * nx + ny * ny);
            double vn = wm.u * nx + wm.v * ny;
            lam += 0.5 * (fabs(vn) + cm * area);
          }
        }
        if (A.ns) {
          double mu;
          double T = wm.T;
          if (A.vp.mu_model == 0) mu = A.vp.mu_ref;
          else if (A.vp.mu_model == 1) mu = A.vp.mu_ref * pow(T, 1.5) / (T + A.vp.T_0);
          else mu = A.vp.mu_ref * pow(T / A.vp.T_ref, A.vp.omega);
          double rho = density(gas, wm);
          for (int e = G.vfe_ptr[m]; e < G.vfe_ptr[m + 1]; ++e) {
            unsigned ent = (unsigned)G.vfe[e];
            unsigned kind = ent >> 30, idx = ent & 0x3fffffffu;
            if (kind >= 2) continue;
            double nx = G.fn[2 * idx], ny = G.fn[2 * idx + 1];
            double area = sqrt(nx * nx + ny * ny);
            int l = G.fv[2 * idx], rr = G.fv[2 * idx + 1];
            double dx = G.x[rr] * 1.0 + G.x[l] * -1.0, dy = G.y[rr] * 1.0 + G.y[l] * -1.0;
            double dl = sqrt(dx * dx + dy * dy);
            lam += (mu * area / dl / rho) * 1.0;
          }
        }
        tot = (p < 0) ? lam : tot + lam * 1.0;
      }
      dtv = P.cfl * G.dcarea[v] / tot;
    }
    if (A.dtl) A.dtl[off + v] = dtv;
  }
  double bm = block_min(dtv, lds);
  if (threadIdx.x == 0) A.partial[(size_t)s * gridDim.x + blockIdx.x] = bm;
}

// Global dt + time-stamp clamp (src/fv.cc:1591-1618), one block per sample.
__global__ __launch_bounds__(256) void k_dt_final(StageArgs A, StepParams P, int nblk) {
  __shared__ double lds[4];
  const int s = blockIdx.x;
  Ctl& c = A.ctl[s];
  if (!A.ignore_done && c.done) return;
  double x = 1.0e20;
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) x = fmin(x, A.partial[(size_t)s * nblk + i]);
  x = block_min(x, lds);
  if (threadIdx.x == 0) {
    double dt = x;
    // a fixed time_step returns before the time-stamp clamp (src/fv.cc:1503-1509)
    if (P.unsteady && !(P.time_step > 0.0)) {
      // the reference indexes past its stamp table if t falls an ulp short of
      // final_time (src/fv.cc:1610); the last stamp is used instead
      double target = A.tstamps[min(c.counter, P.n_stamps)];
      if (c.t + dt > target) {
        dt = target - c.t;
        c.reached = 1;
      }
    }
    c.dt = dt;
  }
}

// Residual norm + time advance (src/fv.cc:1801-1879, :2307-2364)
__global__ __launch_bounds__(256) void k_step_final(StageArgs A, StepParams P, int nblk, int nv,
                                                    int log_slot) {
  __shared__ double lds[4 * 4];
  const int s = blockIdx.x;
  Ctl& c = A.ctl[s];
  if (c.done) return;
  double part[4] = {0, 0, 0, 0};
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
    const double* q = A.partial + ((size_t)s * nblk + i) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) part[k] += q[k];
  }
  block_sum<4>(part, lds);
  if (threadIdx.x == 0) {
    double n[4];
    for (int k = 0; k < 4; ++k) {
      n[k] = sqrt(part[k]) / nv;
      c.rn[k] = n[k];
    }
    double tot = sqrt((n[0] * n[0]) + (n[1] * n[1] + n[2] * n[2]) + (n[3] * n[3]));
    if (c.iter == 0) {
      c.res0 = tot;
      if (c.res0 == 0.0) c.res0 = 1.0;
    }
    c.res_total = tot / c.res0;
    c.iter += 1;
    c.t += c.dt;
    if (c.reached && P.unsteady) { c.counter += 1; c.reached = 0; }
    if (A.log && log_slot >= 0) {
      double* L = A.log + ((size_t)s * A.log_cap + log_slot) * 4;
      L[0] = (double)c.iter; L[1] = c.dt; L[2] = c.t; L[3] = c.res_total;
    }
  }
}

// Forces on a set of boundary faces (src/force.cc:118-217): per selected bface
// (ascending), the pressure force 0.5 (p0 + p1) n_b and, for Navier-Stokes, the
// viscous stress of the mean of the end points' velocity gradients (from the
// vertex gradients of the entropy variables) against n_b.  The boundary
// condition applied to the end-point states is that of bface bq (the
// reference's loop index).  One block: the lanes form the per-face terms, lane 0
// adds them in bface order.
// Mean over a boundary face's two end points of the velocity gradients, and
// with HEAT of the temperature gradient, formed from the entropy-variable
// gradients g (f, vx, vy, l) x (x, y) and the end points' entropy variables:
// gradU = -g[vx]/l + (vx/l^2) g[l] (likewise v), gradT = g[l] / (R l^2)
// (src/force.cc:160-175, src/output_skin_friction.cc:64-74,
// src/output_heatflux.cc:69-76).
struct WallGrad { double gux, guy, gvx, gvy, gTx, gTy; };
template <bool HEAT>
__device__ WallGrad wall_gradients(const Gas& gas, const Ent& e0, const Ent& e1, const double* g0,
                                   const double* g1) {
  const double f0 = fdiv(1.0, e0.l), f1 = fdiv(1.0, e1.l);
  const double a0 = fdiv(f0 * e0.vx, e0.l), a1 = fdiv(f1 * e1.vx, e1.l);
  const double b0 = fdiv(f0 * e0.vy, e0.l), b1 = fdiv(f1 * e1.vy, e1.l);
  const double gu0x = g0[2] * -f0 + g0[6] * a0, gu0y = g0[3] * -f0 + g0[7] * a0;
  const double gu1x = g1[2] * -f1 + g1[6] * a1, gu1y = g1[3] * -f1 + g1[7] * a1;
  const double gv0x = g0[4] * -f0 + g0[6] * b0, gv0y = g0[5] * -f0 + g0[7] * b0;
  const double gv1x = g1[4] * -f1 + g1[6] * b1, gv1y = g1[5] * -f1 + g1[7] * b1;
  WallGrad w;
  w.gux = gu0x * 0.5 + gu1x * 0.5; w.guy = gu0y * 0.5 + gu1y * 0.5;
  w.gvx = gv0x * 0.5 + gv1x * 0.5; w.gvy = gv0y * 0.5 + gv1y * 0.5;
  w.gTx = w.gTy = 0.0;
  if (HEAT) {
    const double t0 = fdiv(1.0, gas.R * (e0.l * e0.l)), t1 = fdiv(1.0, gas.R * (e1.l * e1.l));
    w.gTx = (g0[6] * t0) * 0.5 + (g1[6] * t1) * 0.5;
    w.gTy = (g0[7] * t0) * 0.5 + (g1[7] * t1) * 0.5;
  }
  return w;
}

__device__ double wall_viscosity(const ViscParams& vp, double T) {  // src/material.h:551-568
  if (vp.mu_model == 0) return vp.mu_ref;
  if (vp.mu_model == 1) return fdiv(vp.mu_ref * pow(T, 1.5), T + vp.T_0);
  return vp.mu_ref * pow(fdiv(T, vp.T_ref), vp.omega);
}

__global__ __launch_bounds__(256) void k_forces(DevGeo G, StageArgs A, int s, int bq,
                                                const int* sel, int nsel, int ns_model,
                                                double* terms, double* out) {
  const size_t off = (size_t)s * G.nv;
  const Gas gas{A.gamma[s], A.gasR[s]};
  const int type = G.btype[bq];
  const double qx = G.bn[2 * bq], qy = G.bn[2 * bq + 1];
  for (int j = threadIdx.x; j < nsel; j += blockDim.x) {
    const int b = sel[j];
    const int v0 = G.bv[2 * b], v1 = G.bv[2 * b + 1];
    const double nx = G.bn[2 * b], ny = G.bn[2 * b + 1];
    Prim s0 = load_prim(A.W_in, off + v0), s1 = load_prim(A.W_in, off + v1);
    bc_apply1(type, bc_values(G, s, bq, bc_end(G, bq, v0)), qx, qy, s0);
    bc_apply1(type, bc_values(G, s, bq, bc_end(G, bq, v1)), qx, qy, s1);
    const double p = 0.5 * (s0.p + s1.p);
    double* t = terms + 4 * (size_t)j;
    t[0] = nx * p;
    t[1] = ny * p;
    t[2] = t[3] = 0.0;
    if (ns_model) {
      const Ent e0 = prim2ent(gas, s0), e1 = prim2ent(gas, s1);
      const double T = (s0.T + s1.T) / 2.0;
      const double mu = wall_viscosity(A.vp, T);
      const WallGrad w = wall_gradients<false>(gas, e0, e1, A.dE + 8 * (off + v0), A.dE + 8 * (off + v1));
      const double gux = w.gux, guy = w.guy, gvx = w.gvx, gvy = w.gvy;
      const double div = gux + gvy;
      const double sxx = 2.0 * mu * (gux - (1.0 / 3.0) * div);
      const double syy = 2.0 * mu * (gvy - (1.0 / 3.0) * div);
      const double sxy = mu * (guy + gvx);
      t[2] = -(sxx * nx + sxy * ny);
      t[3] = -(sxy * nx + syy * ny);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double f[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = 0; j < nsel; ++j)
      for (int k = 0; k < 4; ++k) f[k] += terms[4 * (size_t)j + k];
    for (int k = 0; k < 4; ++k) out[k] = f[k];
  }
}

void launch_forces(const DevGeo& G, const StageArgs& A, int s, int bq, const int* sel, int nsel,
                   int ns_model, double* terms, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_forces, dim3(1), dim3(256), 0, st, G, A, s, bq, sel, nsel, ns_model, terms,
                     out);
}

// Skin friction and heat flux per selected boundary face (src/output_skin_friction.cc:
// 43-95, src/output_heatflux.cc:43-90): the end points' states after the face's
// own boundary condition, the gradients in A.dE, the unit normal n; out[2j] =
// f x n with f the viscous stress on n, out[2j+1] = -k gradT.n, k = mu cp / Pr.
__global__ __launch_bounds__(256) void k_surface(DevGeo G, StageArgs A, int s, const int* sel,
                                                 int nsel, double* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsel) return;
  const size_t off = (size_t)s * G.nv;
  const Gas gas{A.gamma[s], A.gasR[s]};
  const int b = sel[j];
  const int type = G.btype[b];
  const int v0 = G.bv[2 * b], v1 = G.bv[2 * b + 1];
  const double bx = G.bn[2 * b], by = G.bn[2 * b + 1];
  Prim s0 = load_prim(A.W_in, off + v0), s1 = load_prim(A.W_in, off + v1);
  bc_apply1(type, bc_values(G, s, b, bc_end(G, b, v0)), bx, by, s0);
  bc_apply1(type, bc_values(G, s, b, bc_end(G, b, v1)), bx, by, s1);
  const Ent e0 = prim2ent(gas, s0), e1 = prim2ent(gas, s1);
  const double T = (s0.T + s1.T) / 2.0;
  const double mu = wall_viscosity(A.vp, T);
  const double cp = fdiv(gas.gamma * gas.R, gas.gamma - 1.0);
  const double kc = fdiv(mu * cp, A.vp.prandtl);
  const WallGrad w = wall_gradients<true>(gas, e0, e1, A.dE + 8 * (off + v0), A.dE + 8 * (off + v1));
  const double div = w.gux + w.gvy;
  const double sxx = 2.0 * mu * (w.gux - (1.0 / 3.0) * div);
  const double syy = 2.0 * mu * (w.gvy - (1.0 / 3.0) * div);
  const double sxy = mu * (w.guy + w.gvx);
  const double im = fdiv(1.0, fsqrt(bx * bx + by * by));
  const double nx = bx * im, ny = by * im;
  const double fx = sxx * nx + sxy * ny;
  const double fy = sxy * nx + syy * ny;
  out[2 * j] = fx * ny - fy * nx;
  out[2 * j + 1] = (w.gTx * nx + w.gTy * ny) * -kc;
}

void launch_surface(const DevGeo& G, const StageArgs& A, int s, const int* sel, int nsel,
                    double* out, hipStream_t st) {
  if (nsel > 0)
    hipLaunchKernelGGL(k_surface, dim3((nsel + 255) / 256), dim3(256), 0, st, G, A, s, sel, nsel,
                       out);
}

// bc_scheme strong (src/fv.cc:1694-1704): after a stage's update each bface, in
// bface order, imposes its boundary condition on its end points' state; one
// thread per boundary vertex applies that vertex's bfaces in that order.  The
// entropy variables keep the update's values, as the reference's do.
__global__ __launch_bounds__(256) void k_strong_bc(DevGeo G, StageArgs A) {
  const int s = blockIdx.y;
  if (A.ctl[s].done) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.nsbv) return;
  const int v = G.sbv[i];
  const size_t off = (size_t)s * G.nv;
  Prim w = load_prim(A.W_out, off + v);
  for (int e = G.sb_ptr[i]; e < G.sb_ptr[i + 1]; ++e) {
    const int b = G.sb[e];
    bc_apply1(G.btype[b], bc_values(G, s, b, bc_end(G, b, v)), G.bn[2 * b], G.bn[2 * b + 1], w);
  }
  store_prim(A.W_out, off + v, w);
}

void launch_strong_bc(const DevGeo& G, const StageArgs& A, int S, hipStream_t st) {
  if (G.nsbv > 0)
    hipLaunchKernelGGL(k_strong_bc, dim3((G.nsbv + 255) / 256, S), dim3(256), 0, st, G, A);
}

// Time-dependent boundary values of Heun's stage r, evaluated at the stage
// time t^n + r dt (src/fv.cc:2296-2298) before the stage's first kernel: the
// reference's boundary conditions evaluate their expressions at (x, y,
// int_step_time) each time they are applied during the stage
// (src/bc.h:411-562).  One lane per (entry, sample) interprets the entry's
// postfix program; the comparisons keep the parser's 1e-12 tolerance and
// & | ! take |v| >= 0.5 as true (expr.cpp).
__device__ __forceinline__ double tp_f1(int f, double a) {
  switch (f) {
    case TF_EXP: return exp(a);
    case TF_EXP2: return exp2(a);
    case TF_LOG: return log(a);
    case TF_LOG10: return log10(a);
    case TF_LOG2: return log2(a);
    case TF_SQRT: return sqrt(a);
    case TF_CBRT: return cbrt(a);
    case TF_SIN: return sin(a);
    case TF_COS: return cos(a);
    case TF_TAN: return tan(a);
    case TF_ASIN: return asin(a);
    case TF_ACOS: return acos(a);
    case TF_ATAN: return atan(a);
    case TF_SINH: return sinh(a);
    case TF_COSH: return cosh(a);
    case TF_TANH: return tanh(a);
    case TF_ABS: return fabs(a);
    case TF_FLOOR: return floor(a);
    case TF_CEIL: return ceil(a);
    case TF_TRUNC: return trunc(a);
    case TF_INT: return floor(a + 0.5);
  }
  return __longlong_as_double(0x7ff8000000000000LL);
}
__device__ __forceinline__ double tp_f2(int f, double a, double b) {
  switch (f) {
    case TF_POW: return pow(a, b);
    case TF_ATAN2: return atan2(a, b);
    case TF_MIN: return a < b ? a : b;
    case TF_MAX: return a > b ? a : b;
    case TF_HYPOT: return hypot(a, b);
  }
  return __longlong_as_double(0x7ff8000000000000LL);
}

__global__ __launch_bounds__(64) void k_bc_time(DevGeo G, BcTime B, StageArgs A, int r) {
  const int s = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= B.n) return;
  const Ctl& c = A.ctl[s];
  if (!A.ignore_done && c.done) return;
  const double t = c.t + (double)r * c.dt;
  const double* sv = B.slots + B.slot[j];
  double st[TP_MAX_DEPTH];
  int sp = 0;
  const auto truth = [](double v) { return fabs(v) >= 0.5; };
  for (int pc = B.prog[j];; ++pc) {
    const int ins = B.code[pc], op = ins & 0xff, arg = ins >> 8;
    if (op == TP_END) break;
    if (op == TP_SLOT) { st[sp++] = sv[arg]; continue; }
    if (op == TP_T) { st[sp++] = t; continue; }
    if (op == TP_NEG) { st[sp - 1] = -st[sp - 1]; continue; }
    if (op == TP_NOT) { st[sp - 1] = truth(st[sp - 1]) ? 0.0 : 1.0; continue; }
    if (op == TP_F1) { st[sp - 1] = tp_f1(arg, st[sp - 1]); continue; }
    if (op == TP_IF) {
      sp -= 2;
      st[sp - 1] = truth(st[sp - 1]) ? st[sp] : st[sp + 1];
      continue;
    }
    const double a = st[sp - 2], b = st[sp - 1];
    double v;
    switch (op) {
      case TP_ADD: v = a + b; break;
      case TP_SUB: v = a - b; break;
      case TP_MUL: v = a * b; break;
      case TP_DIV: v = a / b; break;
      case TP_MOD: v = fmod(a, b); break;
      case TP_POW: v = pow(a, b); break;
      case TP_LT: v = (a < b - 1e-12) ? 1.0 : 0.0; break;
      case TP_GT: v = (b < a - 1e-12) ? 1.0 : 0.0; break;
      case TP_LE: v = (a <= b + 1e-12) ? 1.0 : 0.0; break;
      case TP_GE: v = (b <= a + 1e-12) ? 1.0 : 0.0; break;
      case TP_EQ: v = (fabs(a - b) <= 1e-12) ? 1.0 : 0.0; break;
      case TP_NE: v = (fabs(a - b) > 1e-12) ? 1.0 : 0.0; break;
      case TP_AND: v = (truth(a) && truth(b)) ? 1.0 : 0.0; break;
      case TP_OR: v = (truth(a) || truth(b)) ? 1.0 : 0.0; break;
      default: v = tp_f2(arg, a, b); break;  // TP_F2
    }
    st[--sp - 1] = v;
  }
  double* bs = const_cast<double*>(G.bstate) + (size_t)s * G.bs_sample;
  bs[B.target[j]] = st[0];
}

void launch_bc_time(const DevGeo& G, const BcTime& B, const StageArgs& A, int S, int r,
                    hipStream_t st) {
  if (B.n > 0)
    hipLaunchKernelGGL(k_bc_time, dim3((B.n + 63) / 64, S), dim3(64), 0, st, G, B, A, r);
}

// Solution bounds (src/fv.cc:2099-2171): per block the minimum and maximum of
// T, p, u, v over its vertices of sample s, out[8 * block + 2q] = min,
// out[8 * block + 2q + 1] = max for q = T, p, u, v; the host folds the blocks.
__global__ __launch_bounds__(256) void k_bounds(DevGeo G, StageArgs A, int s, double* out) {
  __shared__ double red[8][256];
  const size_t off = (size_t)s * G.nv;
  double lo[4] = {1.0e20, 1.0e20, 1.0e20, 1.0e20}, hi[4] = {-1.0e20, -1.0e20, -1.0e20, -1.0e20};
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < G.nv; v += gridDim.x * blockDim.x) {
    const Prim w = load_prim(A.W_in, off + v);
    const double q[4] = {w.T, w.p, w.u, w.v};
#pragma unroll
    for (int k = 0; k < 4; ++k) {  // std::min / std::max: a NaN never replaces the running value
      lo[k] = q[k] < lo[k] ? q[k] : lo[k];
      hi[k] = hi[k] < q[k] ? q[k] : hi[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) { red[2 * k][threadIdx.x] = lo[k]; red[2 * k + 1][threadIdx.x] = hi[k]; }
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double a = red[2 * k][threadIdx.x + h], b = red[2 * k + 1][threadIdx.x + h];
        if (a < red[2 * k][threadIdx.x]) red[2 * k][threadIdx.x] = a;
        if (red[2 * k + 1][threadIdx.x] < b) red[2 * k + 1][threadIdx.x] = b;
      }
    __syncthreads();
  }
  if (threadIdx.x < 8) out[8 * (size_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

void launch_bounds(const DevGeo& G, const StageArgs& A, int s, double* out, int nblocks,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_bounds, dim3(nblocks), dim3(256), 0, st, G, A, s, out);
}

// Liou carbuncle sensor (src/compute_ssw.cc:7-39)
__global__ void k_ssw_zero(DevGeo G, StageArgs A) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)G.nv * A.ns_samples) A.ssw[i] = 0.0;
}
__global__ __launch_bounds__(256) void k_ssw(DevGeo G, StageArgs A) {
  const int s = blockIdx.y;
  if (!A.ignore_done && A.ctl[s].done) return;
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= G.nf) return;
  const size_t off = (size_t)s * G.nv;
  Gas gas{A.gamma[s], A.gasR[s]};
  int l = G.fv[2 * f], r = G.fv[2 * f + 1];
  double nx = G.fn[2 * f], ny = G.fn[2 * f + 1];
  Prim wl = load_prim(A.W_in, off + l), wr = load_prim(A.W_in, off + r);
  double ui = wl.u * nx + wl.v * ny, uj = wr.u * nx + wr.v * ny;
  double ai = sound_speed(gas, wl), aj = sound_speed(gas, wr);
  double area = sqrt(nx * nx + ny * ny);
  double l1i = ui - ai * area, l1j = uj - aj * area;
  double l5i = ui + ai * area, l5j = uj + aj * area;
  if ((l1i > 0 && l1j < 0) || (l5i > 0 && l5j < 0)) {
    A.ssw[off + l] = 1.0;
    A.ssw[off + r] = 1.0;
  }
}

// Entropy variables of a freshly set state
__global__ void k_prim2ent(DevGeo G, StageArgs A, int s) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= G.nv) return;
  const size_t off = (size_t)s * G.nv;
  Gas gas{A.gamma[s], A.gasR[s]};
  Prim w = load_prim(A.W_in, off + v);
  store_ent(A.E_in, off + v, prim2ent(gas, w));
}

// Global integrals (src/fv.cc:2175-2225) and error norms (src/fv.cc:1712-1795)
__global__ __launch_bounds__(256) void k_global(DevGeo G, StageArgs A, int s, double* out) {
  __shared__ double lds[4 * 2];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t off = (size_t)s * G.nv;
  Gas gas{A.gamma[s], A.gasR[s]};
  double q[2] = {0, 0};
  if (v < G.nv) {
    Prim w = load_prim(A.W_in, off + v);
    double wt = G.weight[v], da = G.dcarea[v];
    double rho = density(gas, w);
    q[0] = (0.5 * rho * (w.u * w.u + w.v * w.v) * da) * wt;
    q[1] = (-rho * entropy(gas, w) / (gas.gamma - 1.0)) * da * wt;
  }
  block_sum<2>(q, lds);
  if (threadIdx.x == 0) { out[2 * blockIdx.x] = q[0]; out[2 * blockIdx.x + 1] = q[1]; }
}

__global__ __launch_bounds__(256) void k_errnorm(DevGeo G, StageArgs A, int s, const double* ex,
                                                 double* out) {
  __shared__ double lds[4 * 24];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t off = (size_t)s * G.nv;
  Gas gas{A.gamma[s], A.gasR[s]};
  double q[24];
#pragma unroll
  for (int k = 0; k < 24; ++k) q[k] = 0.0;
  double mx[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (v < G.nv) {
    Prim w = load_prim(A.W_in, off + v);
    Prim pe{ex[4 * v], ex[4 * v + 1], ex[4 * v + 2], ex[4 * v + 3]};
    double aw = G.dcarea[v], wt = G.weight[v];
    double e[8];
    e[0] = fabs(pe.T - w.T); e[1] = fabs(pe.p - w.p); e[2] = fabs(pe.u - w.u);
    e[3] = fabs(pe.v - w.v); e[4] = 0.0;
    e[5] = fabs(density(gas, pe) - density(gas, w));
    double me = sqrt(pe.u * pe.u + pe.v * pe.v) / sound_speed(gas, pe);
    double mw = sqrt(w.u * w.u + w.v * w.v) / sound_speed(gas, w);
    e[6] = fabs(me - mw);
    e[7] = fabs(entropy(gas, pe) - entropy(gas, w));
#pragma unroll
    for (int k = 0; k < 5; ++k) { q[k] = e[k] * (aw * wt); q[8 + k] = e[k] * e[k] * (aw * wt); mx[k] = e[k]; }
#pragma unroll
    for (int k = 5; k < 8; ++k) { q[k] = e[k] * aw * wt; q[8 + k] = e[k] * e[k] * aw * wt; mx[k] = e[k]; }
  }
  // sums of the 16 L1/L2 terms; maxima through the same tree with fmax
  block_sum<16>(q, lds);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    double x = mx[k];
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_down(x, o, 64));
    mx[k] = x;
  }
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 8; ++k) lds[wid * 8 + k] = mx[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < 8; ++k) {
      double x = lds[k];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) x = fmax(x, lds[w * 8 + k]);
      mx[k] = x;
    }
    double* o = out + (size_t)blockIdx.x * 24;
    for (int k = 0; k < 16; ++k) o[k] = q[k];
    for (int k = 0; k < 8; ++k) o[16 + k] = mx[k];
  }
}

// ---------------------------------------------------------------------------
// launch dispatch
// ---------------------------------------------------------------------------
template <int FLUX, int REC>
static void launch_stage_t(const DevGeo& G, const StageArgs& A, int S, int r, int last, int mode,
                           hipStream_t st, bool do_flux, bool do_update) {
  dim3 blk(256);
  dim3 gv((G.nv + 255) / 256, S), gf((G.nf + 255) / 256, S);
  if (do_flux) hipLaunchKernelGGL((k_flux<FLUX, REC>), gf, blk, 0, st, G, A);
  if (do_update) {
    if (mode == 0) hipLaunchKernelGGL((k_update<FLUX, REC, 0>), gv, blk, 0, st, G, A, r, last);
    else if (mode == 1) hipLaunchKernelGGL((k_update<FLUX, REC, 1>), gv, blk, 0, st, G, A, r, last);
    else hipLaunchKernelGGL((k_update<FLUX, REC, 2>), gv, blk, 0, st, G, A, r, last);
  }
}

template <int FLUX>
static void launch_rec(int rec, const DevGeo& G, const StageArgs& A, int S, int r, int last,
                       int mode, hipStream_t st, bool df, bool du) {
  switch (rec) {
    case R_FIRST: launch_stage_t<FLUX, R_FIRST>(G, A, S, r, last, mode, st, df, du); break;
    case R_SECOND: launch_stage_t<FLUX, R_SECOND>(G, A, S, r, last, mode, st, df, du); break;
    case R_MUSCL: launch_stage_t<FLUX, R_MUSCL>(G, A, S, r, last, mode, st, df, du); break;
    case R_ALBADA: launch_stage_t<FLUX, R_ALBADA>(G, A, S, r, last, mode, st, df, du); break;
    case R_MINMOD: launch_stage_t<FLUX, R_MINMOD>(G, A, S, r, last, mode, st, df, du); break;
    case R_TVB: launch_stage_t<FLUX, R_TVB>(G, A, S, r, last, mode, st, df, du); break;
    default: launch_stage_t<FLUX, R_ENO2>(G, A, S, r, last, mode, st, df, du); break;
  }
}

void launch_flux_update(int flux, int rec, const DevGeo& G, const StageArgs& A, int S, int r,
                        int last, int mode, hipStream_t st, bool df, bool du) {
  switch (flux) {
    case F_ROE: launch_rec<F_ROE>(rec, G, A, S, r, last, mode, st, df, du); break;
    case F_RUS: launch_rec<F_RUS>(rec, G, A, S, r, last, mode, st, df, du); break;
    // the central fluxes do not reconstruct: one instantiation each
    case F_KEPEC: launch_stage_t<F_KEPEC, R_FIRST>(G, A, S, r, last, mode, st, df, du); break;
    case F_ROEEC: launch_stage_t<F_ROEEC, R_FIRST>(G, A, S, r, last, mode, st, df, du); break;
    case F_KEP: launch_stage_t<F_KEP, R_FIRST>(G, A, S, r, last, mode, st, df, du); break;
    default: launch_stage_t<F_AVG, R_FIRST>(G, A, S, r, last, mode, st, df, du); break;
  }
}

void launch_gradient(const DevGeo& G, const StageArgs& A, int S, hipStream_t st) {
  hipLaunchKernelGGL(k_gradient, dim3((G.nv + 255) / 256, S), dim3(256), 0, st, G, A, -1);
}
void launch_gradient_one(const DevGeo& G, const StageArgs& A, int s, hipStream_t st) {
  hipLaunchKernelGGL(k_gradient, dim3((G.nv + 255) / 256, 1), dim3(256), 0, st, G, A, s);
}
void launch_dt(const DevGeo& G, const StageArgs& A, const StepParams& P, int S, hipStream_t st) {
  int nb = (G.nv + 255) / 256;
  hipLaunchKernelGGL(k_dt, dim3(nb, S), dim3(256), 0, st, G, A, P);
  hipLaunchKernelGGL(k_dt_final, dim3(S), dim3(256), 0, st, A, P, nb);
}
// Caller's stop time for the next step: a running sample whose t + dt would
// pass t_target steps to t_target exactly (the stamp flag stays as it is).
__global__ void k_clamp_dt(StageArgs A, int S, double t_target) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  Ctl& c = A.ctl[s];
  if (c.done || !(c.t < t_target)) return;
  if (c.t + c.dt > t_target) c.dt = t_target - c.t;
}
void launch_clamp_dt(const StageArgs& A, int S, double t_target, hipStream_t st) {
  hipLaunchKernelGGL(k_clamp_dt, dim3((S + 63) / 64), dim3(64), 0, st, A, S, t_target);
}

void launch_step_final(const DevGeo& G, const StageArgs& A, const StepParams& P, int S,
                       int log_slot, hipStream_t st) {
  int nb = (G.nv + 255) / 256;
  hipLaunchKernelGGL(k_step_final, dim3(S), dim3(256), 0, st, A, P, nb, G.nv, log_slot);
}
void launch_ssw(const DevGeo& G, const StageArgs& A, int S, hipStream_t st) {
  size_t n = (size_t)G.nv * S;
  hipLaunchKernelGGL(k_ssw_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, G, A);
  hipLaunchKernelGGL(k_ssw, dim3((G.nf + 255) / 256, S), dim3(256), 0, st, G, A);
}
void launch_prim2ent(const DevGeo& G, const StageArgs& A, int s, hipStream_t st) {
  hipLaunchKernelGGL(k_prim2ent, dim3((G.nv + 255) / 256), dim3(256), 0, st, G, A, s);
}
// Enstrophy partials, 1/2 omega^2 dcarea w per vertex summed per block, with
// omega = R T (dE[vy].x + v dE[l].x - dE[vx].y - u dE[l].y) from the sample's
// entropy-variable gradients (Material::Vorticity, src/material.h:622-627).
__global__ __launch_bounds__(256) void k_enstrophy(DevGeo G, StageArgs A, int s, double* out) {
  __shared__ double lds[4];
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t off = (size_t)s * G.nv;
  double q[1] = {0};
  if (v < G.nv) {
    const Prim w = load_prim(A.W_in, off + v);
    const double* g = A.dE + 8 * (off + v);
    const double om = A.gasR[s] * w.T * (g[4] + w.v * g[6] - g[3] - w.u * g[7]);
    q[0] = (0.5 * om * om * G.dcarea[v]) * G.weight[v];
  }
  block_sum<1>(q, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = q[0];
}
void launch_enstrophy(const DevGeo& G, const StageArgs& A, int s, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_enstrophy, dim3((G.nv + 255) / 256), dim3(256), 0, st, G, A, s, out);
}

void launch_global(const DevGeo& G, const StageArgs& A, int s, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_global, dim3((G.nv + 255) / 256), dim3(256), 0, st, G, A, s, out);
}
void launch_errnorm(const DevGeo& G, const StageArgs& A, int s, const double* ex, double* out,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_errnorm, dim3((G.nv + 255) / 256), dim3(256), 0, st, G, A, s, ex, out);
}

}  // namespace tsm

namespace tsm {

// ---------------------------------------------------------------------------
// Monte Carlo reductions over the sample dimension
// ---------------------------------------------------------------------------
// Per-vertex mean, M2 (sum of squared deviations) and sum of squares of
// (T, u, v, p, rho) over the samples in sample order, with the update of
// FiniteVolume::increment_moments (src/fv.cc:634-690): M_k = x/k + M_{k-1}(k-1)/k,
// M2 += (x - M_{k-1})(x - M_k), S2 += x^2; the density mean uses
// M_k = M_{k-1} + (x - M_{k-1})/k as the reference does for derived fields.
__global__ __launch_bounds__(256) void k_moments(DevGeo G, StageArgs A, int S, double* mean,
                                                 double* m2, double* s2) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= G.nv) return;
  double M[5], Q[5], Z[5];
  for (int s = 0; s < S; ++s) {
    const Gas gas{A.gamma[s], A.gasR[s]};
    const Prim w = load_prim(A.W_in, (size_t)s * G.nv + v);
    const double x[5] = {w.T, w.u, w.v, w.p, density(gas, w)};
    if (s == 0) {
#pragma unroll
      for (int q = 0; q < 5; ++q) { M[q] = x[q]; Z[q] = x[q] * x[q]; Q[q] = 0.0; }
    } else {
      const double k = s + 1.0;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const double mk1 = M[q];
        // primitives: PrimVar::equ (src/primvar.h:213-221); density: scalar form (src/fv.cc:659-661)
        const double mk = q < 4 ? x[q] * (1.0 / k) + mk1 * ((k - 1.0) / k) : mk1 + (x[q] - mk1) / k;
        M[q] = mk;
        Z[q] += x[q] * x[q];
        Q[q] += (x[q] - mk1) * (x[q] - mk);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    mean[5 * (size_t)v + q] = M[q];
    m2[5 * (size_t)v + q] = Q[q];
    s2[5 * (size_t)v + q] = Z[q];
  }
}

// Density at probe points: linear interpolation in the containing triangle
// (corner ids pc, barycentric weights pw, both from the host) for every sample.
__global__ void k_probes(StageArgs A, int nv, int np, const int* pc, const double* pw, int S,
                         double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np * S) return;
  const int s = i / np, p = i % np;
  const Gas gas{A.gamma[s], A.gasR[s]};
  double r = 0.0;
  for (int k = 0; k < 3; ++k) {
    const Prim w = load_prim(A.W_in, (size_t)s * nv + pc[3 * p + k]);
    r += pw[3 * p + k] * density(gas, w);
  }
  out[i] = r;
}

// Kinetic energy and entropy integrals of every sample (block partials).
__global__ __launch_bounds__(256) void k_global_all(DevGeo G, StageArgs A, double* out) {
  __shared__ double lds[4 * 2];
  const int v = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  const Gas gas{A.gamma[s], A.gasR[s]};
  double q[2] = {0, 0};
  if (v < G.nv) {
    const Prim w = load_prim(A.W_in, (size_t)s * G.nv + v);
    const double wt = G.weight[v], da = G.dcarea[v];
    const double rho = density(gas, w);
    q[0] = (0.5 * rho * (w.u * w.u + w.v * w.v) * da) * wt;
    q[1] = (-rho * entropy(gas, w) / (gas.gamma - 1.0)) * da * wt;
  }
  block_sum<2>(q, lds);
  if (threadIdx.x == 0) {
    double* o = out + ((size_t)s * gridDim.x + blockIdx.x) * 2;
    o[0] = q[0];
    o[1] = q[1];
  }
}

void launch_moments(const DevGeo& G, const StageArgs& A, int S, double* mean, double* m2,
                    double* s2, hipStream_t st) {
  hipLaunchKernelGGL(k_moments, dim3((G.nv + 255) / 256), dim3(256), 0, st, G, A, S, mean, m2, s2);
}
// The stage kernels' quotient and square root (fdiv, fsqrt) beside the
// compiler's IEEE `/` and sqrt() on the same operands: out[4i..4i+3] =
// fdiv(a, b), a / b, fsqrt(a), sqrt(a).
__global__ void k_check_arith(long n, const double* a, const double* b, double* out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = a[i], y = b[i];
  out[4 * i] = fdiv(x, y);
  out[4 * i + 1] = x / y;
  out[4 * i + 2] = fsqrt(x);
  out[4 * i + 3] = sqrt(x);
}
void launch_check_arith(long n, const double* a, const double* b, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_check_arith, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, b, out);
}

void launch_probes(const StageArgs& A, int nv, int np, const int* pc, const double* pw, int S,
                   double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_probes, dim3((np * S + 255) / 256), dim3(256), 0, st, A, nv, np, pc, pw, S, out);
}
void launch_global_all(const DevGeo& G, const StageArgs& A, int S, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_global_all, dim3((G.nv + 255) / 256, S), dim3(256), 0, st, G, A, out);
}

}  // namespace tsm
