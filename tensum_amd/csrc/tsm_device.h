// tsm_device.h -- device-side structures shared by the kernels and the host
// context of libtsm.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <functional>
#include <string>

#include "../../include/tsm.h"
#include "tsm_physics.h"

namespace tsm {

// Per-sample loop control, resident in device memory so that whole steps can
// be queued without a host round trip (src/fv.cc:2230-2395 loop variables).
struct Ctl {
  double t, dt, res0, res_total;
  long long iter;
  int counter, reached, done, pad;
  double rn[4];  // last step's residual norms: mass, x/y momentum, energy
};

// Read-only mesh arrays (device pointers).
struct DevGeo {
  int nv, nf, nb, nc;
  const double *x, *y, *dcarea, *gfact, *weight, *carea;
  const int* fv;        // 2*nf: left, right vertex of each face (reference order)
  const double* fn;     // 2*nf: dual-face normal
  const int* vfe_ptr;   // nv+1
  const int* vfe;       // per vertex: faces (kind 0 left / 1 right) then bfaces (kind 2), bits 31..30
  const int* vc_ptr;    // nv+1
  const int* vc;        // per vertex: cells ascending; sign bit = cell has boundary corrections
  const int* cv;        // 3*nc
  const int* cbf;       // 2*nc: non-periodic bfaces of the cell (bface order), -1 padded
  const int* cnb;       // nc: cell has a non-periodic bface
  // Navier-Stokes wall cells (cells with a non-periodic bface)
  const int* blc;       // nb: the bface's cell
  const int* cwf;       // 2*nc: the cell's boundary edges as face ids, cell slot order, -1 padded
  const int* cwb;       // 2*nc: the bface of each of those edges
  const int* vbe_ptr;   // nv+1
  const int* vbe;       // per vertex: bfaces whose cell holds it, ascending; bit 31 = vertex on the bface
  const int* bv;        // 2*nb (oriented)
  const double* bn;     // 2*nb outward edge normal
  const int* btype;     // nb
  const double* bstate; // boundary values (T,u,v,p) per bface end point: 8*nb per sample
  long long bs_sample;  // doubles between samples' boundary values (0: all samples share)
  const int* pset;      // nv: periodic set id or -1
  const int* pptr;      // nsets+1
  const int* pv;        // members
  // strong boundary conditions (null when weak): the vertices on bfaces, each
  // with its bfaces in bface order, and per vertex whether it ends a noslip bface
  int strong;           // bc_scheme strong: stored entropy variables of boundary
                        // vertices may predate the imposed boundary values
  int nsbv;
  const int* sbv;
  const int* sb_ptr;    // nsbv+1
  const int* sb;
  const unsigned char* snoslip;  // nv
};

// One operand record per (vertex of k_grad_bnd, cell of that vertex), in the
// vertex's cell order: the cell's corners (new ids), its non-periodic boundary
// faces in bface order (-1: none) and, 2 bits each, the corner (0..2) at each
// end of each of them (bf0 end 0, bf0 end 1, bf1 end 0, bf1 end 1).
struct alignas(16) GbEntry {
  int a, b, d, ends;
  int bf0, bf1, pad0, pad1;
};

// Tile arrays of the LDS-staged kernels (device pointers, new vertex ids).
struct TileGeo {
  int ntiles, tb, max_slots, max_faces;
  int grad_block, fu_block;  // requested threads per tile (0: smallest that fits)
  int n_periodic_tiles;      // tiles [0, n) hold the periodic sets
  int n_gbv;                 // vertices of k_grad_bnd
  const int* gbv;            // their ids
  const int* gb_ptr;         // n_gbv+1: their records [gb_ptr[i], gb_ptr[i+1])
  const GbEntry* gb_ent;
  int fused, max_slots1;     // two-ring tiles: gradients formed inside the stage kernel
  const int *h2beg, *halo2, *g1beg, *g1ptr, *g1cgid;
  const unsigned* g1ent;
  const int *vbeg, *hbeg, *halo, *fbeg, *bbeg, *tbf, *gptr, *cptr, *cgid, *tper, *pinfo;
  const unsigned *f_lr, *cent;
  const double* f_n;
  const unsigned short* gent;
};

// Boundary values that depend on time: entry j writes the value at offset
// target[j] of each sample's boundary block, evaluating the program that
// starts at code[prog[j]] over the slot values from slots[slot[j]] and the
// stage time (tprog.h).
struct BcTime {
  int n;
  const int* target;
  const int* prog;
  const int* slot;
  const int* code;
  const double* slots;
};

struct ViscParams {
  double mu_ref, T_ref, T_0, omega, prandtl;
  int mu_model, visc_active, heat_active;
};


// Boundary values of sample s at end point e (0: bv[2b], 1: bv[2b+1]) of bface b:
// the reference evaluates a boundary's expressions at each end point's
// coordinates with the sample's random numbers (src/bc.h:411-562).
__device__ __forceinline__ const double* bc_values(const DevGeo& G, int s, int b, int e) {
  return G.bstate + (size_t)s * G.bs_sample + 8 * (size_t)b + 4 * e;
}
__device__ __forceinline__ int bc_end(const DevGeo& G, int b, int v) {
  return G.bv[2 * b + 1] == v ? 1 : 0;
}

// Device-only boundary code: a noslip wall without a temperature of its own
// (BoundaryCondition::adiabatic, src/bc.h:246-276).  Slip walls are adiabatic
// by construction.
constexpr int BC_NOSLIP_ADIABATIC = 8;

__device__ __forceinline__ bool bc_adiabatic(int type) {
  return type == TSM_BC_SLIP || type == BC_NOSLIP_ADIABATIC;
}
// types whose one-state apply changes the state (the gradient's boundary
// correction recomputes the entropy variables only for these)
__device__ __forceinline__ bool bc_sets_state(int type) {
  return type == TSM_BC_SLIP || type == TSM_BC_NOSLIP || type == BC_NOSLIP_ADIABATIC ||
         type == TSM_BC_PRESSURE || type == TSM_BC_INLET;
}

// One-state boundary condition (src/bc.h:615-654, :397-451); (nx, ny) is the
// normal of the face the state belongs to.
__device__ __forceinline__ void bc_apply1(int type, const double* bs, double nx, double ny,
                                          Prim& st) {
  switch (type) {
    case TSM_BC_SLIP: {
      double im = fdiv(1.0, fsqrt(nx * nx + ny * ny));
      double ux = nx * im, uy = ny * im;
      double fact = st.u * ux + st.v * uy;
      st.u += ux * -fact;
      st.v += uy * -fact;
      break;
    }
    case TSM_BC_NOSLIP: st.u = bs[1]; st.v = bs[2]; st.T = bs[0]; break;
    case BC_NOSLIP_ADIABATIC: st.u = bs[1]; st.v = bs[2]; break;
    case TSM_BC_PRESSURE: st.p = bs[3]; break;
    case TSM_BC_INLET: st = Prim{bs[0], bs[1], bs[2], bs[3]}; break;
    default: break;
  }
}
// Two-state boundary condition (src/bc.h:569-610, :385-431)
__device__ __forceinline__ void bc_apply2(int type, const double* bs, double nx, double ny,
                                          Prim& s0, Prim& s1) {
  switch (type) {
    case TSM_BC_SLIP: {
      double im = fdiv(1.0, fsqrt(nx * nx + ny * ny));
      double ux = nx * im, uy = ny * im;
      double fact = s0.u * ux + s0.v * uy;
      s0.u += ux * -fact;
      s0.v += uy * -fact;
      s1 = s0;
      break;
    }
    case TSM_BC_NOSLIP:
      s0.u = bs[1]; s0.v = bs[2];
      s1.u = s0.u; s1.v = s0.v; s1.p = s0.p;
      s0.T = bs[0]; s1.T = bs[0];
      break;
    case BC_NOSLIP_ADIABATIC:
      s0.u = bs[1]; s0.v = bs[2];
      s1.u = s0.u; s1.v = s0.v; s1.p = s0.p;
      s1.T = s0.T;
      break;
    case TSM_BC_PRESSURE: s1.p = bs[3]; break;
    case TSM_BC_INLET: s0 = Prim{bs[0], bs[1], bs[2], bs[3]}; s1 = s0; break;
    case TSM_BC_OUTLET: s1 = s0; break;
    case TSM_BC_FARFIELD: s1 = Prim{bs[0], bs[1], bs[2], bs[3]}; break;
    default: break;
  }
}

// Mean entropy variables of a wall cell c (src/fv.cc:1245-1266): the one-state
// boundary states of the ends of the cell's boundary edges (slot order, each
// edge's ends in face order, a slip edge projected with its half dual face
// normal), summed from zero and scaled by 0.5 / (number of edges).  W: the
// stage's array-of-structs state, off: the sample's first vertex.
__device__ __forceinline__ Ent wall_cell_state(const DevGeo& G, const double* W, size_t off,
                                               int s, const Gas& gas, int c) {
  Ent e{0.0, 0.0, 0.0, 0.0};
  int n = 0;
  for (int t = 0; t < 2; ++t) {
    const int f = G.cwf[2 * c + t];
    if (f < 0) break;
    const int b = G.cwb[2 * c + t];
    const int type = G.btype[b];
    const double nx = G.fn[2 * f], ny = G.fn[2 * f + 1];
    for (int k = 0; k < 2; ++k) {
      const int vb = G.fv[2 * f + k];
      const double4 q = reinterpret_cast<const double4*>(W)[off + vb];
      Prim st{q.x, q.y, q.z, q.w};
      bc_apply1(type, bc_values(G, s, b, bc_end(G, b, vb)), nx, ny, st);
      const Ent x = prim2ent(gas, st);
      e.f += x.f; e.vx += x.vx; e.vy += x.vy; e.l += x.l;
    }
    ++n;
  }
  const double sc = 0.5 * 1.0 / n;
  e.f *= sc; e.vx *= sc; e.vy *= sc; e.l *= sc;
  return e;
}

// Viscous flux of a cell delivered to one of its vertices (src/material.cc:
// 158-230): s = the mean of the cell's three entropy-variable states, dE =
// the cell gradient of the entropy variables (x, y per component), (nx, ny) =
// the vertex's cell normal.  mu(T) with T = -1/(R s_l), kappa = mu Cp / Pr.
// With adiabatic_wall set it is the wall flux of an adiabatic wall
// (viscous_bflux, src/material.cc:236-302): the heat-conduction share of the
// energy flux is taken out again; a non-adiabatic wall flux is the cell flux
// through the wall's normal.
__device__ __forceinline__ Cons viscous_cell_flux(const Gas& gas, const ViscParams& vp,
                                                  const Ent& s, const double* dE, double nx,
                                                  double ny, bool adiabatic_wall = false) {
  const double R = gas.R;
  double T = fdiv(-1.0, R * s.l);
  double mu;
  if (vp.mu_model == 0) mu = vp.mu_ref;
  else if (vp.mu_model == 1) mu = fdiv(vp.mu_ref * pow(T, 1.5), T + vp.T_0);
  else mu = vp.mu_ref * pow(fdiv(T, vp.T_ref), vp.omega);
  double Cp = fdiv(gas.gamma * gas.R, gas.gamma - 1.0);
  double k = fdiv(mu * Cp, vp.prandtl);
  if (!vp.visc_active) mu = 0.0;
  if (!vp.heat_active) k = 0.0;
  double e1 = fdiv(s.vx, s.l), e2 = fdiv(s.vy, s.l), e4 = fdiv(1.0, s.l);
  double c1 = fdiv(4.0 * mu, 3.0), c2 = fdiv(k, R), c3 = fdiv(2 * mu, 3.0);
  // dE components: [0,1] f, [2,3] vx, [4,5] vy, [6,7] l  (x, y)
  double d1x = dE[2], d1y = dE[3], d2x = dE[4], d2y = dE[5], d4x = dE[6], d4y = dE[7];
  double K11_1 = c1 * (-d1x + e1 * d4x);
  double K11_2 = mu * (-d2x + e2 * d4x);
  double K11_4 = c1 * e1 * d1x + mu * (e2 * d2x) + (-c1 * e1 * e1 - mu * (e2 * e2) + c2 * e4) * d4x;
  double K12_1 = c3 * (d2y - e2 * d4y);
  double K12_2 = mu * (-d1y + e1 * d4y);
  double K12_4 = mu * e2 * d1y - c3 * e1 * d2y - (mu - c3) * e1 * e2 * d4y;
  double K21_1 = mu * (-d2x + e2 * d4x);
  double K21_2 = c3 * (d1x - e1 * d4x);
  double K21_4 = -c3 * e2 * d1x + mu * e1 * d2x - (mu - c3) * e1 * e2 * d4x;
  double K22_1 = mu * (-d1y + e1 * d4y);
  double K22_2 = c1 * (-d2y + e2 * d4y);
  double K22_4 = mu * (e1 * d1y) + c1 * e2 * d2y + (-mu * (e1 * e1) - c1 * e2 * e2 + c2 * e4) * d4y;
  Cons F;
  F.m = 0.0;
  F.mx = -((K11_1 + K12_1) * nx + (K21_1 + K22_1) * ny) * e4;
  F.my = -((K11_2 + K12_2) * nx + (K21_2 + K22_2) * ny) * e4;
  F.e = -((K11_4 + K12_4) * nx + (K21_4 + K22_4) * ny) * e4;
  if (adiabatic_wall) {  // src/material.cc:260-264, :300
    const double Cor = -1.0 * (0.0 + c2 * e4 * e4 * (d4x * nx + d4y * ny));
    F.e = F.e - Cor;
  }
  return F;
}

struct StepParams {
  double min_residue, final_time, time_step, cfl;
  long long max_iter;
  int unsteady;
  int n_stamps;
};

// Mutable state and scratch (device pointers).  Vertex arrays are array-of-
// structs (4 or 8 doubles per vertex, sample-major, cs = n_samples*nv
// vertices); face arrays are structure-of-arrays strided by fs = n_samples*nf.
struct StageArgs {
  size_t cs, fs, off;
  const double* W_in; double* W_out;
  double* E_in; double* E_out;
  double* Uold;
  double* dE;      // 8 components
  double* F;       // 4 components per face
  double* R;       // 4 components (residual export)
  double* R2;      // 4 components (rk4)
  double* ssw;
  double* partial; // per-block reduction scratch
  double* dtl;     // per-vertex local time step (steady mode), null in unsteady mode
  double* W_keep;  // last stage's input state of each sample's latest step, null unless kept
  Ctl* ctl;
  const double* gamma; const double* gasR;
  const double* tstamps;
  double* log; int log_cap;
  FluxOpts fo;
  ViscParams vp;
  int ns;          // Navier-Stokes model active
  int ns_walls;    // ... on a mesh with wall cells (cells with a non-periodic bface)
  int ns_samples;  // number of samples
  int ignore_done; // single-call API entry points compute regardless of the loop state
  int sample;      // the sample the CSR helpers work on (set per thread block)
  int heun;        // two-stage Heun (SSP-RK2) instead of SSP-RK3 in the SSP update
  int scheme;      // 0 SSP-RK3, 1 RK4, 2 Heun (the tiled stage kernel's update)
};

// One context (a level of a multilevel run) in a packed launch: its geometry,
// tiling, stage arguments (the stage's input and output buffers) and step
// parameters, and where its blocks start in each packed grid.
struct PackLevel {
  DevGeo G;
  TileGeo TG;
  StageArgs A;
  StepParams P;
  const double* Wn;   // W^n (read by every stage for U_old)
  double* dtpart;     // per-block dt partials
  int S;              // samples
  int base_fu, base_gper, base_grest, base_gbnd, base_final;
};
// Grid totals of a packed step over nl levels (computed by pack_plan).
struct PackPlan {
  int nl;
  int fu_blocks, gper_blocks, grest_blocks, gbnd_blocks, final_blocks;
  int fu_block, grad_block;   // threads per block: the largest any level needs
  size_t fu_lds, grad_lds;    // dynamic LDS: the largest any level needs
  bool grad_pass;
};
// Fills the block bases of lv[0..nl) and the plan (host).
void pack_plan(PackLevel* lv, int nl, int flux, int rec, PackPlan& plan);
void launch_packed_gradient(const PackLevel* d, const PackPlan& p, hipStream_t st);
void launch_packed_stage(int flux, int rec, const PackLevel* d, const PackPlan& p, int r,
                         bool last, hipStream_t st);
void launch_packed_final(const PackLevel* d, const PackPlan& p, hipStream_t st);

// LDS-staged tile kernels (Euler, SSP-RK3)
size_t tile_grad_lds(const TileGeo& T);
int tile_grad_block(const TileGeo& T);
int tile_fu_block(const TileGeo& T);
int tile_fu_block_max();
bool tile_needs_gradient_pass(const TileGeo& T, int flux, int rec);
// Builds gb_ptr / gb_ent of T (n_gbv > 0) from its cell entries and G on the
// stream; memory from dalloc.  TSM_OK, or TSM_E_* with msg.
int build_grad_bnd_table(const DevGeo& G, TileGeo& T, const std::function<int(void**, size_t)>& dalloc,
                         hipStream_t st, std::string& msg);
void launch_tile_gradient_bnd(const DevGeo& G, const TileGeo& T, const StageArgs& A, int S,
                              hipStream_t st);
size_t tile_fu_lds(const TileGeo& T, int flux, int rec);
size_t tile_fu_lds_ns(const TileGeo& T, int flux, int rec);
void launch_tile_gradient(const DevGeo& G, const TileGeo& T, const StageArgs& A, int S,
                          hipStream_t st);
void launch_tile_stage(int flux, int rec, const DevGeo& G, const TileGeo& T, const StageArgs& A,
                       const StepParams& P, const double* Wn, double* dtpart, int S, int r,
                       bool last, hipStream_t st);
void launch_step_final_tiled(const DevGeo& G, const StageArgs& A, const StepParams& P, int S,
                             int nblk, const double* dtpart, int log_slot, hipStream_t st);

void launch_forces(const DevGeo& G, const StageArgs& A, int s, int bq, const int* sel, int nsel,
                   int ns_model, double* terms, double* out, hipStream_t st);
void launch_flux_update(int flux, int rec, const DevGeo& G, const StageArgs& A, int S, int r,
                        int last, int mode, hipStream_t st, bool do_flux, bool do_update);
void launch_gradient(const DevGeo& G, const StageArgs& A, int S, hipStream_t st);
// vertex gradients of sample s alone (into the sample's slice of A.dE)
void launch_gradient_one(const DevGeo& G, const StageArgs& A, int s, hipStream_t st);
void launch_surface(const DevGeo& G, const StageArgs& A, int s, const int* sel, int nsel,
                    double* out, hipStream_t st);
// per-block minima and maxima of T, p, u, v of sample s (8 doubles per block)
void launch_bounds(const DevGeo& G, const StageArgs& A, int s, double* out, int nblocks,
                   hipStream_t st);
// strong boundary conditions on the stage output A.W_out of every live sample
void launch_strong_bc(const DevGeo& G, const StageArgs& A, int S, hipStream_t st);
// Evaluates the time-dependent boundary values of every sample at its stage-r
// time t + r dt (Heun).
void launch_bc_time(const DevGeo& G, const BcTime& B, const StageArgs& A, int S, int r,
                    hipStream_t st);
void launch_dt(const DevGeo& G, const StageArgs& A, const StepParams& P, int S, hipStream_t st);
void launch_step_final(const DevGeo& G, const StageArgs& A, const StepParams& P, int S,
                       int log_slot, hipStream_t st);
void launch_ssw(const DevGeo& G, const StageArgs& A, int S, hipStream_t st);
void launch_prim2ent(const DevGeo& G, const StageArgs& A, int s, hipStream_t st);
void launch_global(const DevGeo& G, const StageArgs& A, int s, double* out, hipStream_t st);
void launch_enstrophy(const DevGeo& G, const StageArgs& A, int s, double* out, hipStream_t st);
void launch_errnorm(const DevGeo& G, const StageArgs& A, int s, const double* ex, double* out,
                    hipStream_t st);

void launch_clamp_dt(const StageArgs& A, int S, double t_target, hipStream_t st);
void launch_moments(const DevGeo& G, const StageArgs& A, int S, double* mean, double* m2,
                    double* s2, hipStream_t st);
void launch_probes(const StageArgs& A, int nv, int np, const int* pc, const double* pw, int S,
                   double* out, hipStream_t st);
void launch_global_all(const DevGeo& G, const StageArgs& A, int S, double* out, hipStream_t st);
void launch_check_arith(long n, const double* a, const double* b, double* out, hipStream_t st);
// tsm_noalloc.hip: the fixed scratch array (its device address on the current
// device; 0 or a hipError_t) and the device point location of probe points
size_t noalloc_scratch_bytes();
int noalloc_scratch(void** out);
void launch_probe_locate(const DevGeo& G, int np, const double* px, const double* py, int* pc,
                         double* pw, int* cell, hipStream_t st);

}  // namespace tsm
