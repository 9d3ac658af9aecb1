/*
 * tsm_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++ restatement of the
 * reference solver's vertex-centred finite-volume residual and Runge-Kutta
 * loop (deepray/TENSUM, src/fv.cc and the files it calls), written in the
 * reference's own data order (array-of-structs, face-loop scatter).  It is
 * loaded only by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg, always as the checker and never as the thing measured or
 * shipped.  The product library (libtsm) never links or calls it.
 *
 * Every function here is pinned by the checks listed in tests/test_oracle_*.py
 * (closed forms, conservation and symmetry invariants, brute force on tiny
 * inputs); the reference binary itself cannot be built here because every
 * translation unit of the path includes <mpi.h> (src/mpi_utils.h:4).
 */
#ifndef TSM_ORACLE_H
#define TSM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters: same numeric codes as include/tsm.h (tsm_params). */
typedef struct orc_params {
  double gamma, gas_const, prandtl;
  double mu_ref, T_ref, T_0, omega;
  int model;          /* 0 euler, 1 ns */
  int mu_model;       /* 0 constant, 1 sutherland, 2 power */
  int flux;           /* 0 kepes_tecno_roe 1 kepes_tecno_rusanov 2 kepec 3 roe_ec 4 kep 5 avg */
  int reconstruct;    /* 0 first 1 second 2 second_muscl 3 van_albada 4 minmod 5 tvb_minmod 6 eno2 */
  double tvb_M;
  double ev_alpha, ev_beta;
  int nd_active, visc_active, heat_active;
  int time_scheme;    /* 0 ssprk3, 1 rk4, 2 heuns (the two-stage SSP scheme) */
  int unsteady;       /* 1 unsteady, 0 steady */
  double cfl, time_step, final_time, min_residue;
  long max_iter;
  int n_time_stamps;
  int cell_type;      /* 0 median, 1 voronoi */
  int liou_fix;
  int strong_bc;      /* bc_scheme: 0 weak, 1 strong */
} orc_params;

typedef struct orc_mesh {
  int n_vertex;
  const double* x; const double* y;
  int n_bface;
  const int* bface_tag;      /* n_bface */
  const int* bface_v;        /* 2*n_bface (gids) */
  int n_cell;
  const int* cell_v;         /* 3*n_cell, CCW */
  int n_periodic;
  const int* periodic_ptr;   /* n_periodic+1 */
  const int* periodic_v;     /* members */
  /* boundary condition table: tag -> type (codes of tsm.h: 1 slip, 2 noslip,
     3 farfield, 4 inlet, 5 outlet, 6 pressure, 7 periodic) plus a constant
     state (T,u,v,p) used by farfield/inlet/pressure/noslip */
  int n_bc;
  const int* bc_tag; const int* bc_type; const double* bc_state; /* 4*n_bc */
  /* n_bc or NULL: BoundaryCondition::adiabatic (src/bc.h:226-276); slip walls
     are always adiabatic, NULL leaves noslip walls isothermal */
  const int* bc_adiabatic;
} orc_mesh;

typedef struct orc_solver orc_solver;

orc_solver* orc_create(const orc_mesh* m, const orc_params* p);
void orc_destroy(orc_solver* s);
const char* orc_error(void);

/* geometry export */
int orc_n_face(const orc_solver* s);
int orc_n_bface(const orc_solver* s);
/* face_v: 2*nf, face_n: 2*nf (normal x,y), face_type: nf */
void orc_faces(const orc_solver* s, int* face_v, double* face_n, int* face_type);
void orc_bfaces(const orc_solver* s, int* bface_v, double* bface_n, int* lcell);
void orc_areas(const orc_solver* s, double* dcarea, double* mcarea, double* cell_area);

/* state (primitive T,u,v,p per vertex, 4*V) */
void orc_set_state(orc_solver* s, const double* prim);
void orc_get_state(const orc_solver* s, double* prim);

/* per-call pieces at the current state */
void orc_gradients(orc_solver* s, double* dE /* V*8: (f,vx,vy,l) x (x,y) */);
void orc_residual(orc_solver* s, double* R /* V*4 */);
double orc_compute_dt(orc_solver* s);
/* per-vertex time steps of the last compute_dt (local in steady mode, all
   equal to the global one in unsteady mode), V */
void orc_local_dt(const orc_solver* s, double* dt);

/* time loop: run up to nsteps of the reference while-loop body; returns steps
   done.  log (optional) receives 4 doubles per step: iter, dt, t, res_total */
long orc_run(orc_solver* s, long nsteps, double* log);
double orc_time(const orc_solver* s);
/* pressure (out[0..1]) and viscous (out[2..3]) force on the boundary faces with
   one of the tags, current state (src/force.cc:118-217) */
void orc_forces(orc_solver* s, int force_index, const int* tags, int ntags, double* out);
/* The entropy-variable gradients the reference's dE holds: after a step those of
   its last stage's input state; before any step those of the current state. */
void orc_stage_gradients(orc_solver* s, double* out);
/* Loop clock of a restart file: iteration, stamp counter, time, first-step
   residual norm (src/fv.cc:334-338). */
void orc_get_clock(const orc_solver* s, long* iter, int* counter, double* t, double* res0);
void orc_set_clock(orc_solver* s, long iter, int counter, double t, double res0);
/* Skin friction and heat flux (2 per bface whose tag is listed, bface order)
   from the gradients dE holds (src/output_skin_friction.cc, src/output_heatflux.cc). */
void orc_surface(orc_solver* s, const int* tags, int ntags, double* out);
/* boundary states of this run per input bface b and end point k (the vertex
   bface_v[2b+k]): state[8b+4k..] = (T,u,v,p), the boundary expressions
   evaluated there; replaces the tags' constant states.  0 or -1. */
int orc_set_bc_state(orc_solver* s, const int* bface_v, const double* state);
/* Boundary values of time: in Heun runs, before the residual of stage r, the
   callback receives the stage time t = elapsed + r*dt_global
   (src/fv.cc:2296-2298) and the current values in orc_set_bc_state's layout
   (8 per input bface), which it updates in place; they then replace the
   boundary states.  NULL removes it. */
typedef void (*orc_bc_time_fn)(void* user, double t, double* state);
int orc_set_bc_time_fn(orc_solver* s, const int* bface_v, orc_bc_time_fn fn, void* user);
/* error norms vs a given exact primitive field: out[24] =
   (L1,L2,Linf) x (T,p,u,v,w,rho,mach,entropy) */
void orc_error_norms(const orc_solver* s, const double* exact_prim, double* out);
/* global KE and entropy integrals */
void orc_global(const orc_solver* s, double* ke, double* ent);
/* 1/2 sum omega^2 dcarea w over the vertices, omega from the entropy-variable
   gradients of the current state (Material::Vorticity, src/material.h:622-627). */
double orc_enstrophy(orc_solver* s);
/* Liou carbuncle sensor of the current state (src/compute_ssw.cc:7-39; only
   for kepes_tecno_roe): out[nv] = 1 where a face raises it, else 0. */
void orc_ssw(orc_solver* s, int* out);

/* Monte Carlo moments of S sample states (prim: S x nv x 4, sample-major),
   fields (T,u,v,p,rho), in sample order: mean, M2, S2 each nv x 5 */
void orc_moments(int nv, int S, const double* prim, const double* gamma, const double* gas_const,
                 double* mean, double* m2, double* s2);

/* pointwise pieces */
double orc_logavg(double a, double b);
/* viscous flux through normal n of a state with entropy variables ent[4]
   (f, vx, vy, l) and gradient dE[8]; wall = 1: the wall flux
   (src/material.cc:236-302) with the adiabatic correction when adiabatic */
void orc_viscous_flux(const orc_params* p, const double* ent, const double* dE, const double* n,
                      int wall, int adiabatic, double* F);
/* primitive (T,u,v,p) left/right, gradients 8 each, coords, normal, ssw */
void orc_num_flux(const orc_params* p, const double* xl, const double* xr,
                  const double* Wl, const double* Wr, const double* dEl,
                  const double* dEr, const double* n, double ssw, double* F);
double orc_limited_slope(const orc_params* p, double ul, double ur, double h);
void orc_reconstruct(const orc_params* p, double h, const double* zl,
                     const double* zr, const double* gl, const double* gr,
                     double* rl, double* rr); /* 4 components each */

#ifdef __cplusplus
}
#endif
#endif
