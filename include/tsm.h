/*
 * tsm.h -- C ABI of libtsm, the MI355X-native vertex-centred finite-volume
 * residual + Runge-Kutta library (drop-in for the time loop of deepray/TENSUM).
 *
 * The reference exposes no FFI: its hot path is the private member functions
 * of FiniteVolume (src/fv.h:351-403) called in-process on std::vector state.
 * Each entry point below names the reference function it replaces.  All
 * signatures use plain C types and host pointers; device memory, streams and
 * the vertex renumbering stay inside the opaque context.
 *
 * Conventions: every call returns TSM_OK (0) or a negative TSM_E* status and
 * never aborts or exits (the reference calls MPI_Abort / exit(0) instead,
 * src/mpi_utils.h:23-41).  The message of the last failure on the calling
 * thread is returned by tsm_last_error().  One host thread per context.
 * Primitive state is (T, u, v, p) per vertex, vertex order = part-file order.
 */
#ifndef TSM_H
#define TSM_H

#ifdef __cplusplus
extern "C" {
#endif

#define TSM_OK 0
#define TSM_E_ARG -1       /* bad argument / unsupported option          */
#define TSM_E_MESH -2      /* mesh invalid (non-positive area, tags ...) */
#define TSM_E_HIP -3       /* HIP runtime failure or no device           */
#define TSM_E_IO -4        /* file could not be read / written           */
#define TSM_E_STATE -5     /* non-finite or non-positive T/p detected    */

/* Flux codes follow Material::FluxScheme (src/material.h:70-82). */
enum { TSM_FLUX_KEPES_TECNO_ROE = 0, TSM_FLUX_KEPES_TECNO_RUSANOV = 1,
       TSM_FLUX_KEPEC = 2, TSM_FLUX_ROE_EC = 3, TSM_FLUX_KEP = 4,
       TSM_FLUX_AVG = 5 };
/* Reconstruction codes follow Material::ReconstructScheme (src/material.h:89-98). */
enum { TSM_REC_FIRST = 0, TSM_REC_SECOND = 1, TSM_REC_SECOND_MUSCL = 2,
       TSM_REC_VAN_ALBADA = 3, TSM_REC_MINMOD = 4, TSM_REC_TVB_MINMOD = 5,
       TSM_REC_ENO2 = 6 };
/* Boundary types follow BC::BCType (src/bc.h:14-23). */
enum { TSM_BC_NONE = 0, TSM_BC_SLIP = 1, TSM_BC_NOSLIP = 2,
       TSM_BC_FARFIELD = 3, TSM_BC_INLET = 4, TSM_BC_OUTLET = 5,
       TSM_BC_PRESSURE = 6, TSM_BC_PERIODIC = 7 };

/* Solver parameters: the numeric/material/output fields of Parameter
   (src/parameter.h), already parsed. */
typedef struct tsm_params {
  double gamma, gas_const, prandtl;
  double mu_ref, T_ref, T_0, omega;
  int model;          /* 0 euler, 1 ns */
  int mu_model;       /* 0 constant, 1 sutherland, 2 power */
  int flux;           /* TSM_FLUX_* */
  int reconstruct;    /* TSM_REC_* */
  double tvb_M;       /* M_value of tvb_minmod */
  double ev_alpha, ev_beta;  /* flux_par alpha / beta */
  int nd_active, visc_active, heat_active;
  int time_scheme;    /* 0 ssprk3, 1 rk4, 2 heuns: Heun's two-stage SSP scheme as
                         src/fv.cc:1675-1686 means it (the reference's own update
                         overwrites a term and gives NaN after one step) */
  int unsteady;       /* 1 unsteady, 0 steady */
  double cfl, time_step, final_time, min_residue;
  long max_iter;
  int n_time_stamps;
  int cell_type;      /* 0 median, 1 voronoi */
  int liou_fix;
  /* Kernel family (no counterpart in the reference): 0 = LDS-tiled stage
     kernels where the configuration allows (every model, scheme and flux
     option except bc_scheme strong), CSR-gather kernels otherwise; 1 = always
     the CSR-gather kernels.  Both give the reference's summation order. */
  int kernel_path;
  int tile_vertices;  /* owned vertices per tile of the tiled kernels (0: 128) */
  int block_grad;     /* threads per tile of the tiled gradient kernel (0: auto) */
  int block_fu;       /* threads per tile of the tiled flux/update kernel (0: auto) */
  int tile_fused_gradient; /* 1: on meshes without periodic sets, form the
                              gradients inside the stage kernel (two halo rings)
                              instead of in a pass of their own */
  int graph_steps;    /* tiled path: capture this many steps in one HIP graph and
                         replay it (0: launch every kernel from the host) */
  int host_preprocess; /* 1: build the dual-mesh geometry and the tiling on the
                          host; 0: on the device for median duals (the tiling
                          unless the fused kernel's second halo ring is asked
                          for; the same arrays bit for bit) */
  int strong_bc;      /* bc_scheme (src/parameter.h:31-34): 0 weak; 1 strong: after
                         every stage update each bface, in bface order, imposes its
                         boundary condition on its end points' state (src/fv.cc:
                         1694-1704) and the last stage zeroes the momentum residual
                         of noslip end points (src/fv.cc:1805-1819); CSR kernels */
} tsm_params;

/* NPART = 1 part file content (readme/README.md:155-201), 0-based gids. */
typedef struct tsm_mesh {
  int n_vertex;
  const double* x; const double* y;
  int n_bface;
  const int* bface_tag;      /* n_bface */
  const int* bface_v;        /* 2*n_bface */
  int n_cell;
  const int* cell_v;         /* 3*n_cell, counter-clockwise */
  int n_periodic;
  const int* periodic_ptr;   /* n_periodic+1 */
  const int* periodic_v;     /* set members */
  int n_bc;                  /* boundary table: tag -> type + constant state */
  const int* bc_tag; const int* bc_type; const double* bc_state; /* 4*n_bc (T,u,v,p) */
  /* n_bc or NULL: BoundaryCondition::adiabatic (src/bc.h:226-276).  An
     adiabatic noslip wall keeps the fluid's temperature (no temperature in
     its param.in block) and its viscous wall flux carries no heat; slip walls
     are always adiabatic; NULL makes noslip walls isothermal at their T. */
  const int* bc_adiabatic;
} tsm_mesh;

typedef struct tsm_ctx tsm_ctx;

const char* tsm_last_error(void);

/* Builds geometry (replaces Grid::preproc, src/grid_preproc.cc:1147-1194),
   renumbers, uploads, allocates state for n_samples independent samples. */
int tsm_create(const tsm_mesh* mesh, const tsm_params* params, int n_samples,
               tsm_ctx** out);
void tsm_destroy(tsm_ctx* ctx);

/* Per-sample gamma / gas constant (batched Monte Carlo over material constants;
   the reference can vary only IC/BC expressions per sample, src/perturb.cc). */
int tsm_set_sample_material(tsm_ctx* ctx, int sample, double gamma, double gas_const);

/* Boundary values of one sample, replacing the boundary table's constant
   states: state[8b+4k .. 8b+4k+3] = (T,u,v,p) of bface b (mesh order) at its
   end point k (the vertex bface_v[2b+k]).  The reference evaluates each
   boundary's expressions at every end point's coordinates with the sample's
   random numbers (BoundaryCondition::apply, src/bc.h:411-562; pert,
   src/perturb.cc:16-209); values that do not depend on time are passed
   here.  Only the fields a boundary type reads are used; they must be finite,
   with T > 0 and p > 0 where the type reads them. */
int tsm_set_bc_state(tsm_ctx* ctx, int sample, const double* state);

/* Boundary values that depend on time (replaces the time argument of
   BoundaryCondition::apply, src/bc.h:411-562, which the reference sets to
   int_step_time = elapsed_time + r*dt_global before stage r of Heun's scheme,
   src/fv.cc:2296-2298; the other schemes never set it, so these values are
   refused there).  bface_group[b] (mesh order) names bface b's group, -1 for
   none; expr[4g+q] is the expression of field q (T, u, v, p) of group g in the
   parameter file's syntax of x, y and t, NULL or "" to keep the value
   tsm_set_bc_state or the boundary table gave; the constants name the
   parameter file's user constants (pi is built in; pert() is not available).
   Every sample's values are evaluated on the device at every end point before
   each stage at that sample's stage time, and once now at its current time.
   n_groups = 0 removes them (the values they wrote last stay in force). */
int tsm_set_bc_time_expr(tsm_ctx* ctx, const int* bface_group, int n_groups,
                         const char* const* expr, int n_const,
                         const char* const* const_name, const double* const_value);

/* Initial state at t = 0 (replaces the IC / restart branch of
   FiniteVolume::initialize_sample_run, src/fv.cc:320-402). prim: 4*n_vertex. */
int tsm_set_state(tsm_ctx* ctx, int sample, const double* prim);
int tsm_get_state(tsm_ctx* ctx, int sample, double* prim);

/* Force integral on the current state (FiniteVolume::compute_forces,
   src/force.cc:118-217): over the boundary faces whose tag is one of tags[ntags]
   (in bface order), out[0..1] = sum 0.5 (p0 + p1) n_b and, for Navier-Stokes,
   out[2..3] = -sum tau n_b with tau from the mean of the end points' velocity
   gradients (vertex gradients of the current state).  force_index is the
   force's position in the param.in list: as the reference does, the end-point
   states get the boundary condition of bface number force_index. */
int tsm_forces(tsm_ctx* ctx, int sample, int force_index, const int* tags, int ntags,
               double* out);

/* The reference's outputs after a step read the gradient array dE as the loop
   body left it: the gradients of the last Runge-Kutta stage's input state
   (compute_residual recomputes them every stage, src/fv.cc:1476, :2295-2304),
   unless a Navier-Stokes force evaluation recomputed them on the current state
   (src/force.cc:124).  tsm_keep_stage_state(ctx, 1) makes every step keep each
   sample's last-stage input state (4 doubles per vertex and sample, one more
   store per vertex in the last stage); before any step, and after
   tsm_set_state, the kept state is the current one.  Call it before the steps
   whose outputs need it; 0 stops keeping. */
int tsm_keep_stage_state(tsm_ctx* ctx, int on);
/* Gradients of the kept last-stage input state, the reference's dE after a
   step: the vorticity output and moments (src/writer.cc:241-255,
   src/fv.cc:624-679) use them.  dE: 8*n_vertex as in tsm_gradients.
   TSM_E_ARG if the state is not kept. */
int tsm_stage_gradients(tsm_ctx* ctx, int sample, double* dE);
/* Surface quantities on the boundary faces whose tag is one of tags[ntags], in
   bface order (FiniteVolume::output_surface_sf / output_surface_hf,
   src/output_skin_friction.cc:14-96, src/output_heatflux.cc:14-92): the end
   points' states after the face's own boundary condition; out[2j] = skin
   friction f x n (f the viscous stress on the unit normal n), out[2j+1] =
   heat flux -k gradT.n, k = mu cp / Pr, from the end points' mean velocity and
   temperature gradients.  stage = 1 takes the gradients of the kept
   last-stage state (tsm_stage_gradients), 0 those of the current state.
   out: 2 doubles per selected face. */
int tsm_surface(tsm_ctx* ctx, int sample, int stage, const int* tags, int ntags, double* out);

/* Residual of the current state (FiniteVolume::compute_residual, src/fv.cc:1467). R: 4*n_vertex. */
int tsm_residual(tsm_ctx* ctx, int sample, double* R);
/* Vertex entropy-variable gradients (FiniteVolume::compute_gradients, src/fv.cc:699).
   dE: 8*n_vertex ordered (f,vx,vy,l) x (x,y). */
int tsm_gradients(tsm_ctx* ctx, int sample, double* dE);
/* Global time step of the current state (FiniteVolume::compute_dt, src/fv.cc:1499). */
int tsm_compute_dt(tsm_ctx* ctx, int sample, double* dt_global);

/* Advances every sample by up to nsteps time steps of the loop body of
   FiniteVolume::solve (src/fv.cc:2273-2368); a sample stops by itself when its
   normalised residual falls to min_residue, when it reaches final_time
   (unsteady) or after max_iter steps (steady).  In steady mode (unsteady = 0)
   every vertex advances with its own time step cfl*area/lambda (local time
   stepping, src/fv.cc:1591-1618) and t advances by the smallest of them.
   done[s] receives the steps sample s performed; log (optional,
   n_samples*nsteps*4, sample-major) receives per step (iter, dt_global, t,
   res_total), zero rows after a sample stopped. */
int tsm_run(tsm_ctx* ctx, long nsteps, long* done, double* log);
/* One pass of the loop body of FiniteVolume::solve (src/fv.cc:2277-2316) for
   every running sample; t_target > 0 additionally clamps the step so that a
   sample's time lands on t_target exactly instead of passing it (as the
   reference does at an output time stamp).  dt_out[n_samples] and
   res_norm_out[n_samples] (optional) receive each sample's last dt and
   res_total (src/fv.cc:1862-1876).  t_target > 0 is refused in steady mode. */
int tsm_step(tsm_ctx* ctx, double t_target, double* dt_out, double* res_norm_out);
/* tsm_run over several contexts at once (meshes of different resolution, e.g.
   the levels of a multilevel Monte Carlo run): each context's steps go to its
   own stream, interleaved, so the levels share the GPU.  done[i] receives the
   steps summed over the samples of context i (0 once all of them stopped).
   No counterpart in the reference, which runs one sample at a time
   (src/fv.cc:2440-2459). */
int tsm_run_group(tsm_ctx** ctxs, int n, long nsteps, long* done);
/* tsm_run_group with every kernel of a step launched once over the tiles and
   samples of all n contexts (one grid spanning every level's meshes, block
   ranges per level; the levels' uneven sizes packed into the same launches),
   on the first context's stream.  Each block computes what it computes in its
   own context's launch, so the results are those of tsm_run bit for bit.
   The contexts must run the Euler equations with SSP-RK3 on the tiled kernels
   (gradients in a pass of their own) and share the flux and reconstruction;
   TSM_E_ARG otherwise.  No counterpart in the reference, which runs its
   samples one after another (src/fv.cc:2440-2459). */
int tsm_run_packed(tsm_ctx** ctxs, int n, long nsteps, long* done);
int tsm_time(tsm_ctx* ctx, int sample, double* t, long* iter);
/* Loop state of one sample: v[0] t, v[1] dt of the last/next step, v[2]
   res_total; iter; counter = index of the next output time stamp
   (src/fv.cc:2324-2343); done = the sample has left its time loop. */
int tsm_status(tsm_ctx* ctx, int sample, double* v, long* iter, int* counter, int* done);
/* tsm_status and tsm_residual_norms of every sample in one transfer (a driver
   stepping a batch reads them once per step): v[3*s..], iter[s], counter[s],
   done[s], norms[4*s..]; any of them may be NULL. */
int tsm_status_all(tsm_ctx* ctx, double* v, long* iter, int* counter, int* done, double* norms);
/* The loop clock a restart file carries after the state (src/writer.cc:404-407,
   read back by src/fv.cc:334-338): the last iteration, the stamp counter, the
   elapsed time and res0, the first step's residual norm that every later one
   is normalised by.  tsm_set_clock continues a run from a restart: after
   tsm_set_state it sets the clock, and the next step starts with the loop's
   initial res_total = 1e20 (src/fv.cc:2234); iter > 0 keeps res0. */
/* Solution bounds of the current state (FiniteVolume::compute_bounds,
   src/fv.cc:2099-2171, the -b option): out[10] = min and max of T, p, u, v and
   of the z velocity (zero in two dimensions), in the reference's column order
   TEMP_MIN, TEMP_MAX, P_MIN, P_MAX, U_MIN, U_MAX, V_MIN, V_MAX, W_MIN, W_MAX. */
int tsm_bounds(tsm_ctx* ctx, int sample, double* out);
int tsm_get_clock(tsm_ctx* ctx, int sample, long* iter, int* counter, double* t, double* res0);
int tsm_set_clock(tsm_ctx* ctx, int sample, long iter, int counter, double t, double res0);
/* Residual norms of the sample's last step before normalisation, as the
   steady-mode residual log prints them (FiniteVolume::compute_residual_norm,
   src/fv.cc:1821-1851; log_messages, src/fv.cc:1944-1957): norms[4] =
   sqrt(sum_i w_i (R_i,q / A_i)^2) / n_vertex for q = mass, x momentum,
   y momentum, energy. */
int tsm_residual_norms(tsm_ctx* ctx, int sample, double* norms);

/* FiniteVolume::compute_error_norm (src/fv.cc:1712-1795): exact primitive
   field in, out[24] = (L1, L2, Linf) for (T, p, u, v, w, rho, Mach, entropy). */
int tsm_error_norms(tsm_ctx* ctx, int sample, const double* exact_prim, double* out);
/* FiniteVolume::compute_global (src/fv.cc:2175-2225): kinetic energy and entropy. */
int tsm_global(tsm_ctx* ctx, int sample, double* ke, double* entropy);

/* Enstrophy of the current state, ens = 1/2 sum_i omega_i^2 dcarea_i w_i with
   the weights of compute_global and the vorticity of Material::Vorticity
   (src/material.h:622-627) from the entropy-variable gradients of the current
   state (as tsm_gradients).  The reference writes vorticity fields
   (src/writer.cc:241-255) but integrates no enstrophy; it sits beside
   compute_global's integrals (src/fv.cc:2175-2225). */
int tsm_enstrophy(tsm_ctx* ctx, int sample, double* ens);

/* Monte Carlo reductions over all samples of the context (current states).
   tsm_moments: per vertex, in part-file order, 5 fields (T, u, v, p, rho):
   mean, M2 = sum of squared deviations and S2 = sum of squares, accumulated in
   sample order with the update of FiniteVolume::increment_moments
   (src/fv.cc:634-690); variance = M2/(n-1), second moment = S2/n
   (src/reduce_moments.cc:115-141).  Arrays of 5*n_vertex each (may be NULL). */
int tsm_moments(tsm_ctx* ctx, double* mean, double* m2, double* s2);
/* Density of every sample at n_probe points (px, py), linearly interpolated in
   the containing triangle: out[n_samples*n_probe] (sample-major).  Returns
   TSM_E_ARG if a point lies outside the mesh. */
int tsm_probe_density(tsm_ctx* ctx, int n_probe, const double* px, const double* py, double* out);
/* tsm_probe_density without device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops).  Probe points, weights and results
   live in a fixed scratch array of 8 MiB that the library holds on each GPU it
   runs on, until the process ends; the containing triangle is found on the
   device (the same cell and weights as tsm_probe_density).  Limit: with
   a = 256-byte rounding, a(8 np) x 2 + a(24 np) + a(8 np n_samples) + a(12 np)
   + a(4 np) <= 8 MiB (e.g. 64 probes x 16,000 samples), else TSM_E_ARG and
   nothing runs.  Calls on one GPU share the array: calls that overlap on one
   GPU (several host threads) race.  Output as tsm_probe_density. */
int tsm_probe_density_noalloc(tsm_ctx* ctx, int n_probe, const double* px, const double* py,
                              double* out);
/* Kinetic energy and entropy integrals (src/fv.cc:2175-2225) of every sample:
   ke[n_samples], entropy[n_samples] (either may be NULL). */
int tsm_global_all(tsm_ctx* ctx, double* ke, double* entropy);

/* Sample inputs.  No counterpart in the reference, which reads its random
   numbers from the rnd_file_loc file (src/fv.cc:502-551); these points feed that
   role.  Sobol points in Gray-code order with the Joe-Kuo direction numbers,
   optionally nested-uniform (Owen) scrambled with a per-dimension seed derived
   from `seed`: out[n*dim], row i = point skip+i in [0,1)^dim.  dim <= 16. */
int tsm_sobol(long n, int dim, long skip, unsigned long long seed, int scramble, double* out);
int tsm_sobol_max_dim(void);

/* pert(type, x, y, p) of param.in expressions with one sample's random numbers
   rnd[n_rnd] (src/perturb.cc:16-209, replaces PERT): 1 and 3 multi-modal band
   perturbations, 2 piecewise constant, 4 the p-th number (1-based).  Returns
   TSM_E_ARG for an unknown type or a PERT4 index outside [1, n_rnd], where the
   reference aborts. */
int tsm_pert(int type, double x, double y, double p, const double* rnd, int n_rnd, double* out);

/* Self-check of the device arithmetic (no counterpart in the reference): the
   quotient and square root the stage kernels use (the refinement steps of the
   IEEE sequences without their range scaling) beside the compiler's `/` and
   sqrt() on the same operands: out[4i..4i+3] = fdiv(a_i, b_i), a_i / b_i,
   fsqrt(a_i), sqrt(a_i). */
int tsm_check_arith(long n, const double* a, const double* b, double* out);

/* Geometry export in reference order, for parity checks. */
int tsm_geometry_sizes(tsm_ctx* ctx, int* n_face, int* n_bface);
int tsm_geometry(tsm_ctx* ctx, int* face_v, double* face_n, double* dcarea, double* mcarea);

/* Host-only preprocessing (no device needed): runs the same geometry build as
   tsm_create and returns it in reference order.  Arrays may be NULL (sizes
   only): face_v/face_n 2*n_face, bface_v/bface_n 2*n_bface, areas n_vertex. */
int tsm_host_geometry(const tsm_mesh* mesh, int cell_type, int* face_v, double* face_n,
                      int* bface_v, double* bface_n, double* dcarea, double* mcarea,
                      int* n_face, int* n_bface);

/* tsm_host_geometry's outputs built on the device (median duals: cell_type 0;
   TSM_E_ARG for meshes the device builder leaves to the host). */
int tsm_device_geometry(const tsm_mesh* mesh, int cell_type, int* face_v, double* face_n,
                        int* bface_v, double* bface_n, double* dcarea, double* mcarea,
                        int* n_face, int* n_bface);

/* The vertex tiling built on the device (median duals, one halo ring: what
   tsm_create builds there unless host_preprocess) against
   the host's build of the same mesh, array by array: *n_diff = differing
   entries over all arrays, which = name of the first array that differs (""
   if none; a buffer of at least 16 bytes). */
int tsm_device_tiling_check(const tsm_mesh* mesh, int tile_vertices, long long* n_diff,
                            char* which);

/* Host-only: builds and bounds-checks the vertex tiling of the tiled kernels
   (two_rings: with the second halo ring of the fused stage kernel, on meshes
   without periodic sets).  stats[10] = tiles, periodic tiles, max slots, max
   local faces, first-ring halo entries, local face entries, two rings built,
   second-ring halo entries, gradient cell entries, max gradient slots. */
int tsm_host_tiling(const tsm_mesh* mesh, int cell_type, int tile_vertices, int two_rings,
                    int* stats);

/* Times nsteps full steps on the context stream with HIP events and reports
   average durations in microseconds.  CSR kernels: us[0] dt pass, us[1]
   gradient, us[2] edge flux, us[3] gather + update (per stage).  Tiled
   kernels: us[0] 0, us[1] gradient, us[2] fused flux + gather + update (per
   stage), us[3] step finalisation (per step).  us[4]: whole step, timed
   without interleaved events. */
int tsm_bench(tsm_ctx* ctx, int nsteps, double* us);

/* Seeded jittered-triangle lattice on [x0,x1]x[y0,y1] with n x n vertices
   (the build's own mesh generator; the reference ships none). Sizes first: */
int tsm_mesh_jittered_sizes(int n, int periodic, int* n_vertex, int* n_bface,
                            int* n_cell, int* n_periodic, int* n_periodic_v);
int tsm_mesh_jittered(int n, double x0, double x1, double y0, double y1,
                      double jitter, unsigned long long seed, int periodic,
                      double* x, double* y, int* bface_tag, int* bface_v,
                      int* cell_v, int* periodic_ptr, int* periodic_v);


/* ---------------------------------------------------------------------------
   Surrogate network (no counterpart in the reference): an MLP
   widths[0] -> ... -> widths[n_layers] with exact GELU between layers, BF16
   weights and activations on MFMA with FP32 accumulation, FP32 master weights
   and Adam (beta1 0.9, beta2 0.999, eps 1e-8).  Loss: mean over rows and
   outputs of (y - t)^2.  Parameters are flat fp32: W_0 (widths[1] x widths[0],
   row-major), b_0, W_1, b_1, ...  Inputs and targets are row-major fp32 host
   arrays.  One context owns one HIP stream.
   --------------------------------------------------------------------------- */
typedef struct tsm_mlp tsm_mlp;
const char* tsm_mlp_last_error(void);
int tsm_mlp_create(int n_layers, const int* widths, int max_batch, unsigned long long seed,
                   tsm_mlp** out);
void tsm_mlp_destroy(tsm_mlp* m);
long tsm_mlp_n_params(const tsm_mlp* m);
int tsm_mlp_get_params(tsm_mlp* m, float* params);
int tsm_mlp_set_params(tsm_mlp* m, const float* params);
/* y[n x widths[L]] = net(x[n x widths[0]]), in chunks of max_batch rows */
int tsm_mlp_forward(tsm_mlp* m, long n, const float* x, float* y);
/* loss and gradient (flat, like the parameters) of one batch, n <= max_batch */
int tsm_mlp_gradient(tsm_mlp* m, int n, const float* x, const float* t, double* loss, float* grad);
/* epochs of minibatch Adam over a seeded shuffle; epoch_loss[epochs] optional */
int tsm_mlp_train(tsm_mlp* m, long n, const float* x, const float* t, int epochs, int batch,
                  float lr, unsigned long long seed, double* epoch_loss);
/* C[M x N] (fp32) = A B through the MLP's MFMA GEMM, bf16 bit patterns in */
int tsm_mlp_gemm_test(int M, int N, int K, const unsigned short* A, int a_kmajor,
                      const unsigned short* B, int b_kmajor, float* C);

#ifdef __cplusplus
}
#endif
#endif
