// tsm_api.cpp -- host context and C ABI of libtsm (include/tsm.h).
//
// Owns the device copies of the geometry and of the per-sample state, one HIP
// stream, and the stage loop that replaces the body of FiniteVolume::solve
// (src/fv.cc:2273-2368).  No PyTorch, no MPI: the whole mesh is resident on
// one GPU and every reduction of the reference (dt, residual norm, periodic
// sums) is a device kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tsm_device.h"
#include "tsm_host.h"
#include "expr.h"
#include "tprog.h"

namespace {
thread_local std::string g_last_error;
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return fail(TSM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));  \
  } while (0)
}  // namespace

struct tsm_ctx {
  tsm_params P{};
  int S = 1;
  tsm::Geometry geo;
  hipStream_t st = nullptr;
  int dev = 0;              // device current at tsm_create (holds the buffers and the stream)
  tsm::DevGeo G{};
  tsm::StageArgs A{};
  tsm::StepParams SP{};
  std::vector<void*> allocs;
  double* W[3] = {nullptr, nullptr, nullptr};
  double* E[2] = {nullptr, nullptr};
  int cur = 0;              // W[cur] / E[cur] hold the current state
  bool tiled = false;       // LDS-tiled stage kernels in use
  bool e_valid = true;      // E[cur] == prim2ent(W[cur])
  tsm::Tiling tiling;
  tsm::TileGeo TG{};
  double* dtpart = nullptr;
  std::vector<int> perm;    // part-file vertex id -> device vertex id
  // input bface b, end point k (bface_v[2b+k]) -> offset of its boundary
  // values in a sample's block of G.bstate; device boundary codes
  std::vector<long long> bmap;
  double* force_buf = nullptr;  // per-face force terms and the sum (tsm_forces)
  int* force_sel = nullptr;
  size_t force_cap = 0;
  double* W_keep = nullptr;     // last stage's input states (tsm_keep_stage_state)
  double* bounds_buf = nullptr; // per-block minima and maxima (tsm_bounds)
  std::vector<int> h_btype;
  double* red = nullptr;   // reduction partials for global / error norms
  size_t red_cap = 0;
  std::vector<double> h_gamma, h_R;
  int nblk_v = 0;
  // captured steps of the tiled path (graph_steps > 0)
  hipGraphExec_t gexec = nullptr;
  int gcur = -1;
  // device copies of the level descriptors of packed multilevel steps
  // (tsm_run_packed with this context first)
  tsm::PackLevel* pack = nullptr;
  size_t pack_cap = 0;
  // time-dependent boundary values (tsm_set_bc_time_expr), re-evaluated
  // before every stage
  tsm::BcTime bt{};

  ~tsm_ctx() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (void* p : allocs) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }
  template <class T>
  int alloc(T** p, size_t n) {
    void* q = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) return fail(TSM_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(q);
    *p = (T*)q;
    return TSM_OK;
  }
  template <class T>
  int upload(const T** dst, const std::vector<T>& src) {
    T* d = nullptr;
    int rc = alloc(&d, src.size());
    if (rc) return rc;
    if (!src.empty()) HIP_TRY(hipMemcpy(d, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    *dst = d;
    return TSM_OK;
  }
  void bind_state() {
    A.W_in = W[cur]; A.W_out = W[cur ^ 1];
    A.E_in = E[cur]; A.E_out = E[cur ^ 1];
  }
  // Makes W[cur] the state of the CSR kernels' double buffer (W[0]/W[1]) and
  // refreshes the entropy variables they read.
  void csr_ready() {
    state_ready();
    if (!e_valid) {
      for (int s = 0; s < S; ++s) tsm::launch_prim2ent(G, A, s, st);
      e_valid = true;
    }
  }
  // W[cur] in the CSR kernels' double buffer, entropy variables as they are
  void state_ready() {
    if (cur == 2) {
      (void)hipMemcpyAsync(W[0], W[2], 4 * A.cs * sizeof(double), hipMemcpyDeviceToDevice, st);
      cur = 0;
      e_valid = false;
    }
    bind_state();
  }
  // ... with sample s's entropy variables current (single-sample queries: the
  // other samples' stay as they are, so a batch of queries costs one sample each)
  void sample_ready(int s) {
    state_ready();
    if (!e_valid) tsm::launch_prim2ent(G, A, s, st);
  }
};

extern "C" const char* tsm_last_error(void) { return g_last_error.c_str(); }

namespace {

int validate_params(const tsm_params& p) {
  if (!(p.gamma > 1.0)) return fail(TSM_E_ARG, "gamma must be > 1");
  if (!(p.gas_const > 0.0)) return fail(TSM_E_ARG, "gas_const must be > 0");
  if (p.flux < 0 || p.flux > 5) return fail(TSM_E_ARG, "unknown flux scheme");
  if (p.reconstruct < 0 || p.reconstruct > 6) return fail(TSM_E_ARG, "unknown reconstruction");
  if (p.time_scheme < 0 || p.time_scheme > 2)
    return fail(TSM_E_ARG, "time_scheme must be ssprk3 (0), rk4 (1) or heuns (2)");
  if (p.unsteady != 0 && p.unsteady != 1) return fail(TSM_E_ARG, "unsteady must be 0 or 1");
  if (p.strong_bc != 0 && p.strong_bc != 1) return fail(TSM_E_ARG, "strong_bc must be 0 (weak) or 1 (strong)");
  if (p.time_step == 0.0 && p.cfl == 0.0) return fail(TSM_E_ARG, "Either time_step or cfl must be non-zero");
  if (p.time_step > 0.0 && p.cfl > 0.0) return fail(TSM_E_ARG, "both time_step and cfl specified");
  if (p.unsteady && !(p.final_time > 0.0))
    return fail(TSM_E_ARG, "final_time must be positive for unsteady flow");
  if (p.unsteady && p.n_time_stamps <= 0) return fail(TSM_E_ARG, "time_stamps must be > 0");
  if (!p.unsteady && p.max_iter <= 0) return fail(TSM_E_ARG, "max_iter must be > 0");
  if (!(p.min_residue >= 0.0)) return fail(TSM_E_ARG, "min_residue must be >= 0");
  if (p.model == 0 && p.mu_ref > 0.0) return fail(TSM_E_ARG, "mu_ref must be 0 for the model=euler");
  return TSM_OK;
}

}  // namespace

extern "C" int tsm_create(const tsm_mesh* mesh, const tsm_params* params, int n_samples,
                          tsm_ctx** out) {
  if (!mesh || !params || !out || n_samples < 1) return fail(TSM_E_ARG, "null argument");
  *out = nullptr;
  int rc = validate_params(*params);
  if (rc) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(TSM_E_HIP, "no HIP device visible");
  tsm_ctx* c = new tsm_ctx();
  c->P = *params;
  c->S = n_samples;
  std::string msg;
  int grc = 1;  // median duals: geometry on the device unless the host is asked for
  if (params->cell_type == 0 && !params->host_preprocess) {
    grc = tsm::build_geometry_device(*mesh, c->geo, msg);
    if (grc < 0) { delete c; return fail(grc, msg); }
  }
  if (grc == 1 && !tsm::build_geometry(*mesh, params->cell_type, c->geo, msg)) {
    delete c;
    return fail(TSM_E_MESH, msg);
  }
  const tsm::Geometry& g = c->geo;
  const int nv = g.nv, nc = g.nc, nf = (int)g.face.size(), nb = (int)g.bface.size();
  bool wall_cells = false;  // Navier-Stokes cells with a non-periodic bface
  if (params->model == 1)
    for (int k = 0; k < nc; ++k) wall_cells = wall_cells || g.cell_has_nonper_bface[k];
  if ((size_t)nf >= (1u << 30) || (size_t)nb >= (1u << 30)) {
    delete c;
    return fail(TSM_E_MESH, "mesh too large for 30-bit face ids");
  }
  // ---- path and vertex numbering -----------------------------------------
  // tiled stage kernels: every scheme, Euler or Navier-Stokes, with or without
  // the Liou fix; strong boundary conditions keep the CSR kernels
  c->tiled = params->kernel_path == 0 && !params->strong_bc;
  std::vector<int> perm(nv), inv(nv);
  tsm::TileGeo dev_tg{};   // device-built tile arrays (dev_tiles)
  bool dev_tiles = false;
  if (c->tiled) {
    int tb = params->tile_vertices > 0 ? params->tile_vertices : 128;
    if (tb > 256) { delete c; return fail(TSM_E_ARG, "tile_vertices must be <= 256"); }
    if (params->graph_steps < 0 || params->graph_steps > 4096) {
      delete c;
      return fail(TSM_E_ARG, "graph_steps must be in [0, 4096]");
    }
    const bool grad = (params->flux == 0 || params->flux == 1) && params->reconstruct != 0;
    const bool ring2 = grad && params->tile_fused_gradient && params->model == 0 &&
                       params->time_scheme == 0;
    // on the device where the mesh allows it (no periodic sets, one halo
    // ring), else on the host; both give the same arrays
    int trc = 1;
    if (!params->host_preprocess) {
      auto dalloc = [c](void** p, size_t bytes) -> int {
        char* q = nullptr;
        const int r = c->alloc(&q, bytes);
        *p = q;
        return r;
      };
      trc = tsm::build_tiles_device(g, tb, ring2, c->tiling, dev_tg, dalloc, msg);
      if (trc < 0) { delete c; return fail(trc, "tiling: " + msg); }
    }
    dev_tiles = trc == 0;
    if (!dev_tiles &&
        (!tsm::build_tiles(g, tb, c->tiling, msg, ring2) || !tsm::validate_tiles(g, c->tiling, msg))) {
      delete c;
      return fail(TSM_E_MESH, "tiling: " + msg);
    }
    perm = c->tiling.perm;
    inv = c->tiling.inv;
  } else {
    for (int i = 0; i < nv; ++i) perm[i] = inv[i] = i;
  }
  c->perm = perm;
  // ---- host layout (device vertex ids) -----------------------------------
  std::vector<int> fv(2 * (size_t)nf);
  std::vector<double> fn(2 * (size_t)nf);
  for (int f = 0; f < nf; ++f) {
    fv[2 * f] = perm[g.face[f].v[0]]; fv[2 * f + 1] = perm[g.face[f].v[1]];
    fn[2 * f] = g.face[f].nx; fn[2 * f + 1] = g.face[f].ny;
  }
  // vertex -> face/bface CSR in reference order
  std::vector<int> vfe_ptr(nv + 1, 0), vfe;
  for (int f = 0; f < nf; ++f) { vfe_ptr[fv[2 * f] + 1]++; vfe_ptr[fv[2 * f + 1] + 1]++; }
  for (int b = 0; b < nb; ++b)
    if (!g.bface[b].periodic) { vfe_ptr[perm[g.bface[b].v[0]] + 1]++; vfe_ptr[perm[g.bface[b].v[1]] + 1]++; }
  for (int i = 0; i < nv; ++i) vfe_ptr[i + 1] += vfe_ptr[i];
  vfe.resize(vfe_ptr[nv]);
  {
    std::vector<int> pos(vfe_ptr.begin(), vfe_ptr.end() - 1);
    for (int f = 0; f < nf; ++f) {
      vfe[pos[fv[2 * f]]++] = (int)((0u << 30) | (unsigned)f);
      vfe[pos[fv[2 * f + 1]]++] = (int)((1u << 30) | (unsigned)f);
    }
    for (int b = 0; b < nb; ++b) {
      if (g.bface[b].periodic) continue;
      vfe[pos[perm[g.bface[b].v[0]]]++] = (int)((2u << 30) | (unsigned)b);
      vfe[pos[perm[g.bface[b].v[1]]]++] = (int)((2u << 30) | (unsigned)b);
    }
  }
  // cell -> non-periodic bfaces (bface order)
  std::vector<int> cbf(2 * (size_t)nc, -1);
  for (int b = 0; b < nb; ++b) {
    if (g.bface[b].periodic) continue;
    int cl = g.bface[b].lcell;
    if (cbf[2 * cl] < 0) cbf[2 * cl] = b;
    else if (cbf[2 * cl + 1] < 0) cbf[2 * cl + 1] = b;
    else { delete c; return fail(TSM_E_MESH, "cell with three boundary faces"); }
  }
  std::vector<int> cv(3 * (size_t)nc);
  for (size_t k = 0; k < cv.size(); ++k) cv[k] = perm[g.cell[k]];
  std::vector<int> vc_ptr(nv + 1, 0), vc;
  for (int k = 0; k < nc; ++k)
    for (int j = 0; j < 3; ++j) vc_ptr[cv[3 * k + j] + 1]++;
  for (int i = 0; i < nv; ++i) vc_ptr[i + 1] += vc_ptr[i];
  vc.resize(vc_ptr[nv]);
  {
    std::vector<int> pos(vc_ptr.begin(), vc_ptr.end() - 1);
    for (int k = 0; k < nc; ++k)
      for (int j = 0; j < 3; ++j)
        vc[pos[cv[3 * k + j]]++] = (cbf[2 * k] >= 0) ? (int)(0x80000000u | (unsigned)k) : k;
  }
  std::vector<int> cnb(nc);
  for (int k = 0; k < nc; ++k) cnb[k] = (cbf[2 * k] >= 0) ? 1 : 0;
  std::vector<int> bvv(2 * (size_t)nb), btype(nb);
  std::vector<double> bn(2 * (size_t)nb), bstate(8 * (size_t)nb, 0.0);
  for (int b = 0; b < nb; ++b) {
    bvv[2 * b] = perm[g.bface[b].v[0]]; bvv[2 * b + 1] = perm[g.bface[b].v[1]];
    bn[2 * b] = g.bface[b].nx; bn[2 * b + 1] = g.bface[b].ny;
    btype[b] = g.bface[b].bc;
    for (int i = 0; i < mesh->n_bc; ++i)
      if (mesh->bc_tag[i] == g.bface[b].tag) {
        for (int k = 0; k < 4; ++k)  // both end points take the tag's constant state
          bstate[8 * b + k] = bstate[8 * b + 4 + k] = mesh->bc_state ? mesh->bc_state[4 * i + k] : 0.0;
        if (btype[b] == TSM_BC_NOSLIP && mesh->bc_adiabatic && mesh->bc_adiabatic[i])
          btype[b] = tsm::BC_NOSLIP_ADIABATIC;
      }
  }
  std::vector<double> xs(nv), ys(nv), dca(nv), gfact(nv), wgt(nv);
  std::vector<int> pset(nv);
  for (int i = 0; i < nv; ++i) {
    int o = inv[i];
    xs[i] = g.x[o]; ys[i] = g.y[o]; dca[i] = g.dcarea[o]; wgt[i] = g.weight[o];
    gfact[i] = 1.0 / (6.0 * g.mcarea[o]);
    pset[i] = g.periodic_of[o];
  }
  std::vector<int> pptr(1, 0), pv;
  for (auto& P : g.periodic) { for (int v : P) pv.push_back(perm[v]); pptr.push_back((int)pv.size()); }

  // Navier-Stokes wall cells (src/fv.cc:1236-1312): each cell's boundary edges
  // in its slot order with their bfaces, and per vertex the bfaces, in bface
  // order, whose cell holds the vertex (the order the reference's bface loop
  // adds the cell and wall fluxes in)
  std::vector<int> blc(std::max(nb, 1), -1), cwf(2 * (size_t)nc, -1), cwb(2 * (size_t)nc, -1);
  std::vector<int> vbe_ptr(nv + 1, 0), vbe;
  if (wall_cells) {
    std::vector<int> face_bface(nf, -1);
    {
      std::vector<std::pair<long long, int>> key;
      for (int b = 0; b < nb; ++b) {
        if (g.bface[b].periodic) continue;
        int a0 = g.bface[b].v[0], a1 = g.bface[b].v[1];
        key.push_back({(long long)std::min(a0, a1) * nv + std::max(a0, a1), b});
      }
      std::sort(key.begin(), key.end());
      for (int f = 0; f < nf; ++f) {
        if (g.face[f].type < 0 || g.face[f].periodic) continue;
        int a0 = g.face[f].v[0], a1 = g.face[f].v[1];
        long long k = (long long)std::min(a0, a1) * nv + std::max(a0, a1);
        auto it = std::lower_bound(key.begin(), key.end(), std::make_pair(k, -1));
        if (it == key.end() || it->first != k) { delete c; return fail(TSM_E_MESH, "boundary edge without a bface"); }
        face_bface[f] = it->second;
      }
    }
    for (int k = 0; k < nc; ++k) {
      int t = 0;
      for (int j = 0; j < 3; ++j) {
        // slots keep their indices into the face list as it was before Voronoi
        // duals lost their empty faces (src/grid_preproc.cc:633-647, :811-831);
        // boundary faces come first and are never removed (build_geometry
        // refuses a dual that would lose one), so a slot past the shortened
        // list, like any slot past the boundary faces, is interior
        int f = g.cell_slot_face[3 * k + j];
        if (f < 0 || f >= nf || face_bface[f] < 0) continue;
        if (t == 2) { delete c; return fail(TSM_E_MESH, "cell with three boundary edges"); }
        cwf[2 * k + t] = f;
        cwb[2 * k + t] = face_bface[f];
        ++t;
      }
    }
    for (int b = 0; b < nb; ++b) {
      if (g.bface[b].periodic) continue;
      blc[b] = g.bface[b].lcell;
      for (int j = 0; j < 3; ++j) vbe_ptr[cv[3 * blc[b] + j] + 1]++;
    }
    for (int i = 0; i < nv; ++i) vbe_ptr[i + 1] += vbe_ptr[i];
    vbe.resize(vbe_ptr[nv]);
    std::vector<int> pos(vbe_ptr.begin(), vbe_ptr.end() - 1);
    for (int b = 0; b < nb; ++b) {
      if (g.bface[b].periodic) continue;
      for (int j = 0; j < 3; ++j) {
        int v = cv[3 * blc[b] + j];
        bool on = v == bvv[2 * b] || v == bvv[2 * b + 1];
        vbe[pos[v]++] = (int)((on ? 0x80000000u : 0u) | (unsigned)b);
      }
    }
  }
  if (vbe.empty()) vbe.push_back(0);
  // ---- device ------------------------------------------------------------
  if (hipGetDevice(&c->dev) != hipSuccess) { delete c; return fail(TSM_E_HIP, "hipGetDevice failed"); }
  hipError_t he = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
  if (he != hipSuccess) { delete c; return fail(TSM_E_HIP, "hipStreamCreate failed"); }
  tsm::DevGeo& G = c->G;
  G.nv = nv; G.nf = nf; G.nb = nb; G.nc = nc;
#define UP(field, vec) do { if ((rc = c->upload(&G.field, vec))) { delete c; return rc; } } while (0)
  UP(x, xs); UP(y, ys); UP(dcarea, dca); UP(gfact, gfact); UP(weight, wgt);
  UP(carea, g.cell_area);
  UP(fv, fv); UP(fn, fn); UP(vfe_ptr, vfe_ptr); UP(vfe, vfe); UP(vc_ptr, vc_ptr); UP(vc, vc);
  UP(cv, cv); UP(cbf, cbf); UP(cnb, cnb); UP(bv, bvv); UP(bn, bn); UP(btype, btype);
  UP(bstate, bstate); UP(pset, pset); UP(pptr, pptr); UP(pv, pv);
  G.bs_sample = 0;
  c->h_btype = btype;
  if (mesh->n_bface != nb) { delete c; return fail(TSM_E_MESH, "boundary face count changed in preprocessing"); }
  {
    std::vector<std::pair<long long, int>> key(nb);
    for (int b = 0; b < nb; ++b) {
      int a0 = g.bface[b].v[0], a1 = g.bface[b].v[1];
      key[b] = {(long long)std::min(a0, a1) * nv + std::max(a0, a1), b};
    }
    std::sort(key.begin(), key.end());
    c->bmap.assign(2 * (size_t)nb, 0);
    for (int bi = 0; bi < nb; ++bi) {
      int a0 = mesh->bface_v[2 * bi], a1 = mesh->bface_v[2 * bi + 1];
      long long k = (long long)std::min(a0, a1) * nv + std::max(a0, a1);
      auto it = std::lower_bound(key.begin(), key.end(), std::make_pair(k, -1));
      if (it == key.end() || it->first != k) { delete c; return fail(TSM_E_MESH, "bface lost in preprocessing"); }
      const int bd = it->second;
      for (int e = 0; e < 2; ++e)
        c->bmap[2 * bi + e] = 8LL * bd + 4 * (g.bface[bd].v[1] == mesh->bface_v[2 * bi + e] ? 1 : 0);
    }
  }
  UP(blc, blc); UP(cwf, cwf); UP(cwb, cwb); UP(vbe_ptr, vbe_ptr); UP(vbe, vbe);
  if (params->strong_bc) {  // boundary vertices with their bfaces in bface order
    std::vector<std::vector<int>> vb(nv);
    std::vector<unsigned char> noslip(nv, 0);
    for (int b = 0; b < nb; ++b)
      for (int e = 0; e < 2; ++e) {
        const int v = bvv[2 * b + e];
        vb[v].push_back(b);
        if (btype[b] == TSM_BC_NOSLIP || btype[b] == tsm::BC_NOSLIP_ADIABATIC) noslip[v] = 1;
      }
    std::vector<int> sbv, sb_ptr(1, 0), sb;
    for (int v = 0; v < nv; ++v) {
      if (vb[v].empty()) continue;
      sbv.push_back(v);
      sb.insert(sb.end(), vb[v].begin(), vb[v].end());
      sb_ptr.push_back((int)sb.size());
    }
    G.strong = 1;
    G.nsbv = (int)sbv.size();
    if (sbv.empty()) { sbv.push_back(0); sb.push_back(0); }
    UP(sbv, sbv); UP(sb_ptr, sb_ptr); UP(sb, sb); UP(snoslip, noslip);
  }
#undef UP
  if (c->tiled) {
    const tsm::Tiling& T = c->tiling;
    tsm::TileGeo& TG = c->TG;
    TG.ntiles = T.ntiles; TG.tb = T.tb; TG.max_slots = T.max_slots; TG.max_faces = T.max_faces;
    TG.n_periodic_tiles = T.n_periodic_tiles;
    TG.n_gbv = dev_tiles ? dev_tg.n_gbv : (int)T.gbv.size();
    TG.gbv = nullptr;
    TG.gb_ptr = nullptr;
    TG.gb_ent = nullptr;
    TG.fused = T.fused ? 1 : 0;
    TG.max_slots1 = T.max_slots1;
    TG.h2beg = TG.halo2 = TG.g1beg = TG.g1ptr = TG.g1cgid = nullptr;
    TG.g1ent = nullptr;
    TG.grad_block = std::max(0, params->block_grad);
    TG.fu_block = std::max(0, params->block_fu);
    const int bf = tsm::tile_fu_block(TG), bg = tsm::tile_grad_block(TG);
    if (bf > tsm::tile_fu_block_max() || bf % 64 || bg > 512 || bg % 64 ||
        (params->model == 1 ? tsm::tile_fu_lds_ns(TG, params->flux, params->reconstruct)
                            : tsm::tile_fu_lds(TG, params->flux, params->reconstruct)) > 160 * 1024 ||
        tsm::tile_grad_lds(TG) > 160 * 1024) {
      delete c;
      return fail(TSM_E_ARG, "tiled kernels: block size or LDS out of range (blocks " +
                                 std::to_string(bg) + "/" + std::to_string(bf) + ")");
    }
#define UT(field, vec) do { if ((rc = c->upload(&TG.field, vec))) { delete c; return rc; } } while (0)
    if (dev_tiles) {
      TG.vbeg = dev_tg.vbeg; TG.hbeg = dev_tg.hbeg; TG.halo = dev_tg.halo; TG.fbeg = dev_tg.fbeg;
      TG.bbeg = dev_tg.bbeg; TG.tbf = dev_tg.tbf; TG.gptr = dev_tg.gptr; TG.cptr = dev_tg.cptr;
      TG.cgid = dev_tg.cgid; TG.tper = dev_tg.tper; TG.pinfo = dev_tg.pinfo; TG.f_lr = dev_tg.f_lr;
      TG.cent = dev_tg.cent; TG.f_n = dev_tg.f_n; TG.gent = dev_tg.gent;
      if (TG.n_gbv > 0) TG.gbv = dev_tg.gbv;
    } else {
      UT(vbeg, T.vbeg); UT(hbeg, T.hbeg); UT(halo, T.halo); UT(fbeg, T.fbeg); UT(bbeg, T.bbeg);
      UT(tbf, T.tbf); UT(gptr, T.gptr); UT(cptr, T.cptr); UT(cgid, T.cgid); UT(tper, T.tper);
      UT(pinfo, T.pinfo); UT(f_lr, T.f_lr); UT(cent, T.cent); UT(f_n, T.f_n); UT(gent, T.gent);
      if (!T.gbv.empty()) UT(gbv, T.gbv);
    }
    if (T.fused) {
      UT(h2beg, T.h2beg); UT(halo2, T.halo2); UT(g1beg, T.g1beg); UT(g1ptr, T.g1ptr);
      UT(g1cgid, T.g1cgid); UT(g1ent, T.g1ent);
    }
#undef UT
    auto dalloc = [c](void** p, size_t bytes) -> int {
      char* q = nullptr;
      const int r = c->alloc(&q, bytes);
      *p = q;
      return r;
    };
    if ((rc = tsm::build_grad_bnd_table(c->G, TG, dalloc, c->st, msg))) {
      delete c;
      return fail(rc, msg);
    }
  }
  tsm::StageArgs& A = c->A;
  const int S = n_samples;
  A.cs = (size_t)S * nv;
  A.fs = (size_t)S * nf;
  A.off = 0;
  c->nblk_v = (nv + 255) / 256;
  const size_t cs = A.cs;
#define AL(ptr, n) do { if ((rc = c->alloc(&(ptr), (n)))) { delete c; return rc; } } while (0)
  AL(c->W[0], 4 * cs); AL(c->W[1], 4 * cs); AL(c->W[2], 4 * cs);
  AL(c->E[0], 4 * cs); AL(c->E[1], 4 * cs);
  AL(A.Uold, 4 * cs); AL(A.dE, 8 * cs); AL(A.F, 4 * A.fs); AL(A.R, 4 * cs);
  if (params->time_scheme == 1) AL(A.R2, 4 * cs);
  if (!params->unsteady) AL(A.dtl, cs);  // local time steps (src/fv.cc:1591-1618)
  if (params->liou_fix) AL(A.ssw, cs);
  AL(A.partial, (size_t)S * std::max(c->nblk_v, c->tiling.ntiles) * 4);
  if (c->tiled) AL(c->dtpart, (size_t)S * c->tiling.ntiles);
  AL(A.ctl, (size_t)S);
  double* dg; double* dr; double* ts;
  AL(dg, (size_t)S); AL(dr, (size_t)S);
  // time stamps of an unsteady run (src/fv.cc:50-60); a steady run has none
  // and no final time (src/parameter.cc:250-252)
  const int n_stamps = params->unsteady ? params->n_time_stamps : 1;
  const double final_time = params->unsteady ? params->final_time : 1.0e20;
  std::vector<double> tst(n_stamps + 1);
  double tdt = final_time / n_stamps;
  for (size_t i = 0; i < tst.size(); ++i) tst[i] = i * tdt;
  AL(ts, tst.size());
#undef AL
  HIP_TRY(hipMemcpy(ts, tst.data(), tst.size() * sizeof(double), hipMemcpyHostToDevice));
  c->h_gamma.assign(S, params->gamma);
  c->h_R.assign(S, params->gas_const);
  HIP_TRY(hipMemcpy(dg, c->h_gamma.data(), S * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dr, c->h_R.data(), S * sizeof(double), hipMemcpyHostToDevice));
  A.gamma = dg; A.gasR = dr; A.tstamps = ts;
  A.log = nullptr; A.log_cap = 0;
  A.fo = tsm::FluxOpts{params->ev_alpha, params->ev_beta, params->tvb_M, params->nd_active};
  A.vp = tsm::ViscParams{params->mu_ref, params->T_ref, params->T_0, params->omega,
                         params->prandtl, params->mu_model, params->visc_active,
                         params->heat_active};
  A.ns = params->model == 1;
  A.ns_walls = wall_cells;
  A.heun = params->time_scheme == 2;
  A.scheme = params->time_scheme;
  A.ns_samples = S;
  A.ignore_done = 0;
  c->SP = tsm::StepParams{params->min_residue, final_time, params->time_step,
                          params->cfl, (long long)params->max_iter, params->unsteady, n_stamps};
  std::vector<tsm::Ctl> ctl(S);
  for (auto& q : ctl) { q = tsm::Ctl{}; q.res_total = 1e20; q.counter = 1; q.done = 1; }
  HIP_TRY(hipMemcpy(A.ctl, ctl.data(), S * sizeof(tsm::Ctl), hipMemcpyHostToDevice));
  c->bind_state();
  *out = c;
  return TSM_OK;
}

extern "C" void tsm_destroy(tsm_ctx* c) { delete c; }

namespace {
// Gives every sample a block of boundary values of its own (they share one
// until a sample's values differ).
int own_bstate(tsm_ctx* c) {
  if (c->G.bs_sample != 0 || c->S == 1) return TSM_OK;
  const int nb = c->G.nb;
  double* all = nullptr;
  int rc = c->alloc(&all, (size_t)c->S * 8 * nb);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(c->st));
  for (int k = 0; k < c->S; ++k)
    HIP_TRY(hipMemcpy(all + (size_t)k * 8 * nb, c->G.bstate, 8 * (size_t)nb * sizeof(double),
                      hipMemcpyDeviceToDevice));
  c->G.bstate = all;
  c->G.bs_sample = 8LL * nb;
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }  // captured the old array
  return TSM_OK;
}
}  // namespace

extern "C" int tsm_set_bc_state(tsm_ctx* c, int s, const double* state) {
  if (!c || !state || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  const int nb = c->G.nb;
  std::vector<double> h(8 * (size_t)nb, 0.0);
  for (int bi = 0; bi < nb; ++bi)
    for (int e = 0; e < 2; ++e) {
      const double* q = state + 8 * (size_t)bi + 4 * e;
      const long long o = c->bmap[2 * bi + e];
      const int type = c->h_btype[o / 8];
      for (int k = 0; k < 4; ++k)
        if (!std::isfinite(q[k])) return fail(TSM_E_STATE, "non-finite boundary value at bface " + std::to_string(bi));
      const bool needT = type == TSM_BC_INLET || type == TSM_BC_FARFIELD || type == TSM_BC_NOSLIP;
      const bool needp = type == TSM_BC_INLET || type == TSM_BC_FARFIELD || type == TSM_BC_PRESSURE;
      if ((needT && !(q[0] > 0.0)) || (needp && !(q[3] > 0.0)))
        return fail(TSM_E_STATE, "non-positive boundary T or p at bface " + std::to_string(bi));
      for (int k = 0; k < 4; ++k) h[o + k] = q[k];
    }
  if (c->S > 1) {  // samples stop sharing one block
    int rc = own_bstate(c);
    if (rc) return rc;
  }
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(const_cast<double*>(c->G.bstate) + (size_t)s * c->G.bs_sample, h.data(),
                    h.size() * sizeof(double), hipMemcpyHostToDevice));
  return TSM_OK;
}

extern "C" int tsm_set_bc_time_expr(tsm_ctx* c, const int* bface_group, int n_groups,
                                    const char* const* expr, int n_const,
                                    const char* const* const_name, const double* const_value) {
  if (!c || n_groups < 0 || (n_groups > 0 && (!bface_group || !expr)) || n_const < 0 ||
      (n_const > 0 && (!const_name || !const_value)))
    return fail(TSM_E_ARG, "bad argument");
  if (n_groups == 0) {
    c->bt = tsm::BcTime{};
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    return TSM_OK;
  }
  // the reference advances the boundaries' time only in Heun's scheme; the
  // other schemes leave int_step_time unset (src/fv.cc:2294-2299)
  if (c->P.time_scheme != 2 || !c->P.unsteady)
    return fail(TSM_E_ARG, "time-dependent boundary values need an unsteady run with time_scheme heuns");
  std::map<std::string, double> consts;
  for (int k = 0; k < n_const; ++k) {
    if (!const_name[k]) return fail(TSM_E_ARG, "null constant name");
    consts[const_name[k]] = const_value[k];
  }
  // one program per group and field (T, u, v, p), none where no text is given
  std::vector<tsm::Expr> ex(4 * (size_t)n_groups);
  std::vector<int> code, pbeg(4 * (size_t)n_groups, -1);
  for (int g = 0; g < 4 * n_groups; ++g) {
    if (!expr[g] || !expr[g][0]) continue;
    std::string err;
    std::vector<int> pc;
    int depth = 0;
    if (!ex[g].parse(expr[g], consts, err) || !ex[g].time_program(pc, depth, err))
      return fail(TSM_E_ARG, "boundary group " + std::to_string(g / 4) + ": " + err);
    pbeg[g] = (int)code.size();
    code.insert(code.end(), pc.begin(), pc.end());
  }
  // entries: every end point of every bface in a group, each given field
  const int nb = c->G.nb;
  std::vector<int> target, prog, slot;
  std::vector<double> slots;
  for (int bi = 0; bi < nb; ++bi) {
    const int g = bface_group[bi];
    if (g < 0) continue;
    if (g >= n_groups) return fail(TSM_E_ARG, "bface " + std::to_string(bi) + ": group out of range");
    for (int e = 0; e < 2; ++e) {
      const long long o = c->bmap[2 * bi + e];
      const int v = c->geo.bface[o / 8].v[(o % 8) / 4];
      for (int q = 0; q < 4; ++q) {
        const int k = 4 * g + q;
        if (pbeg[k] < 0) continue;
        target.push_back((int)o + q);
        prog.push_back(pbeg[k]);
        slot.push_back((int)slots.size());
        for (int j = 0; j < ex[k].n_time_slots(); ++j)
          slots.push_back(ex[k].time_slot(j, c->geo.x[v], c->geo.y[v]));
      }
    }
  }
  int rc = own_bstate(c);
  if (rc) return rc;
  tsm::BcTime B{};
  B.n = (int)target.size();
  if ((rc = c->upload(&B.target, target)) || (rc = c->upload(&B.prog, prog)) ||
      (rc = c->upload(&B.slot, slot)) || (rc = c->upload(&B.code, code)) ||
      (rc = c->upload(&B.slots, slots)))
    return rc;
  c->bt = B;
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }  // captured without them
  // the values at the samples' current time
  tsm::StageArgs A = c->A;
  A.ignore_done = 1;
  tsm::launch_bc_time(c->G, c->bt, A, c->S, 0, c->st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->st));
  return TSM_OK;
}

extern "C" int tsm_set_sample_material(tsm_ctx* c, int s, double gamma, double R) {
  if (!c || s < 0 || s >= c->S || !(gamma > 1.0) || !(R > 0.0)) return fail(TSM_E_ARG, "bad argument");
  c->h_gamma[s] = gamma; c->h_R[s] = R;
  HIP_TRY(hipMemcpy((void*)(c->A.gamma + s), &gamma, sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy((void*)(c->A.gasR + s), &R, sizeof(double), hipMemcpyHostToDevice));
  return TSM_OK;
}

extern "C" int tsm_set_state(tsm_ctx* c, int s, const double* prim) {
  if (!c || !prim || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  const int nv = c->G.nv;
  for (int i = 0; i < nv; ++i) {
    if (!(prim[4 * i] > 0.0) || !(prim[4 * i + 3] > 0.0) || !std::isfinite(prim[4 * i]) ||
        !std::isfinite(prim[4 * i + 3]))
      return fail(TSM_E_STATE, "initial temperature and pressure must be positive and finite "
                               "(vertex " + std::to_string(i) + ")");
    if (!std::isfinite(prim[4 * i + 1]) || !std::isfinite(prim[4 * i + 2]))
      return fail(TSM_E_STATE, "initial velocity is not finite (vertex " + std::to_string(i) + ")");
  }
  std::vector<double> aos(4 * (size_t)nv);
  for (int i = 0; i < nv; ++i)
    for (int k = 0; k < 4; ++k) aos[4 * (size_t)c->perm[i] + k] = prim[4 * i + k];
  const size_t off = (size_t)s * nv;
  HIP_TRY(hipStreamSynchronize(c->st));
  c->csr_ready();
  HIP_TRY(hipMemcpy(c->W[c->cur] + 4 * off, aos.data(), aos.size() * sizeof(double),
                    hipMemcpyHostToDevice));
  c->bind_state();
  tsm::launch_prim2ent(c->G, c->A, s, c->st);
  HIP_TRY(hipGetLastError());
  if (c->A.W_keep)  // before any step the kept state is the current one
    HIP_TRY(hipMemcpyAsync(c->A.W_keep + 4 * off, c->W[c->cur] + 4 * off, aos.size() * sizeof(double),
                           hipMemcpyDeviceToDevice, c->st));
  tsm::Ctl q{};
  q.res_total = 1e20; q.counter = 1; q.done = 0;
  HIP_TRY(hipMemcpyAsync(c->A.ctl + s, &q, sizeof(q), hipMemcpyHostToDevice, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  return TSM_OK;
}

namespace {
// Vertex arrays are array-of-structs with ncomp doubles per vertex.
int download_vec(tsm_ctx* c, const double* base, int ncomp, int s, double* out) {
  const int nv = c->G.nv;
  const size_t off = (size_t)s * nv;
  std::vector<double> aos((size_t)ncomp * nv);
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(aos.data(), base + (size_t)ncomp * off, aos.size() * sizeof(double),
                    hipMemcpyDeviceToHost));
  for (int i = 0; i < nv; ++i)
    for (int k = 0; k < ncomp; ++k) out[(size_t)ncomp * i + k] = aos[(size_t)ncomp * c->perm[i] + k];
  return TSM_OK;
}
}  // namespace

extern "C" int tsm_get_state(tsm_ctx* c, int s, double* prim) {
  if (!c || !prim || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  return download_vec(c, c->W[c->cur], 4, s, prim);
}

extern "C" int tsm_gradients(tsm_ctx* c, int s, double* dE) {
  if (!c || !dE || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  c->sample_ready(s);
  c->A.ignore_done = 1;
  tsm::launch_gradient_one(c->G, c->A, s, c->st);
  c->A.ignore_done = 0;
  HIP_TRY(hipGetLastError());
  return download_vec(c, c->A.dE, 8, s, dE);
}

extern "C" int tsm_forces(tsm_ctx* c, int s, int force_index, const int* tags, int ntags,
                          double* out) {
  const int nb = c ? c->G.nb : 0;
  if (!c || !out || s < 0 || s >= c->S || ntags < 1 || !tags || force_index < 0 ||
      force_index >= nb)
    return fail(TSM_E_ARG, "bad argument");
  std::vector<int> sel;
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < ntags; ++k)
      if (c->geo.bface[b].tag == tags[k]) { sel.push_back(b); break; }
  if (sel.empty()) return fail(TSM_E_ARG, "no boundary face has the force's tags");
  if (c->force_cap < sel.size()) {
    int rc = c->alloc(&c->force_buf, 4 * sel.size() + 4);
    if (rc) return rc;
    rc = c->alloc(&c->force_sel, sel.size());
    if (rc) return rc;
    c->force_cap = sel.size();
  }
  c->sample_ready(s);
  HIP_TRY(hipMemcpyAsync(c->force_sel, sel.data(), sel.size() * sizeof(int), hipMemcpyHostToDevice,
                         c->st));
  c->A.ignore_done = 1;
  if (c->P.model == 1) tsm::launch_gradient_one(c->G, c->A, s, c->st);
  c->A.ignore_done = 0;
  tsm::launch_forces(c->G, c->A, s, force_index, c->force_sel, (int)sel.size(), c->P.model == 1,
                     c->force_buf + 4, c->force_buf, c->st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, c->force_buf, 4 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  return TSM_OK;
}

extern "C" int tsm_keep_stage_state(tsm_ctx* c, int on) {
  if (!c) return fail(TSM_E_ARG, "bad argument");
  if (on && !c->A.W_keep) {
    if (!c->W_keep) {
      int rc = c->alloc(&c->W_keep, 4 * c->A.cs);
      if (rc) return rc;
    }
    c->state_ready();
    HIP_TRY(hipMemcpyAsync(c->W_keep, c->W[c->cur], 4 * c->A.cs * sizeof(double),
                           hipMemcpyDeviceToDevice, c->st));
    c->A.W_keep = c->W_keep;
  } else if (!on) {
    c->A.W_keep = nullptr;
  }
  // a captured step graph holds the stage kernels' arguments
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  HIP_TRY(hipStreamSynchronize(c->st));
  return TSM_OK;
}

namespace {
// Vertex gradients of sample s into its slice of A.dE: of the kept last-stage
// input state (stage) or of the current state.
int sample_gradients(tsm_ctx* c, int s, bool stage) {
  if (stage && !c->A.W_keep)
    return fail(TSM_E_ARG, "the last stage's state is not kept (tsm_keep_stage_state)");
  if (stage) c->state_ready();
  else c->sample_ready(s);
  tsm::StageArgs K = c->A;
  K.ignore_done = 1;
  if (stage) {  // entropy variables of the kept state in the scratch entropy buffer
    K.W_in = c->A.W_keep;
    K.E_in = c->E[c->cur ^ 1];
    tsm::launch_prim2ent(c->G, K, s, c->st);
  }
  tsm::launch_gradient_one(c->G, K, s, c->st);
  HIP_TRY(hipGetLastError());
  return TSM_OK;
}
}  // namespace

extern "C" int tsm_stage_gradients(tsm_ctx* c, int s, double* dE) {
  if (!c || !dE || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  int rc = sample_gradients(c, s, true);
  if (rc) return rc;
  return download_vec(c, c->A.dE, 8, s, dE);
}

extern "C" int tsm_surface(tsm_ctx* c, int s, int stage, const int* tags, int ntags, double* out) {
  const int nb = c ? c->G.nb : 0;
  if (!c || !out || s < 0 || s >= c->S || ntags < 0 || (ntags > 0 && !tags))
    return fail(TSM_E_ARG, "bad argument");
  std::vector<int> sel;
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < ntags; ++k)
      if (c->geo.bface[b].tag == tags[k]) { sel.push_back(b); break; }
  if (sel.empty()) return TSM_OK;
  if (c->force_cap < sel.size()) {
    int rc = c->alloc(&c->force_buf, 4 * sel.size() + 4);
    if (rc) return rc;
    rc = c->alloc(&c->force_sel, sel.size());
    if (rc) return rc;
    c->force_cap = sel.size();
  }
  int rc = sample_gradients(c, s, stage != 0);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(c->force_sel, sel.data(), sel.size() * sizeof(int), hipMemcpyHostToDevice,
                         c->st));
  tsm::launch_surface(c->G, c->A, s, c->force_sel, (int)sel.size(), c->force_buf, c->st);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, c->force_buf, 2 * sel.size() * sizeof(double), hipMemcpyDeviceToHost,
                         c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  return TSM_OK;
}

extern "C" int tsm_residual(tsm_ctx* c, int s, double* R) {
  if (!c || !R || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  c->csr_ready();
  c->A.ignore_done = 1;
  if (c->P.liou_fix) tsm::launch_ssw(c->G, c->A, c->S, c->st);
  tsm::launch_gradient(c->G, c->A, c->S, c->st);
  tsm::launch_flux_update(c->P.flux, c->P.reconstruct, c->G, c->A, c->S, 0, 0, 0, c->st, true, true);
  c->A.ignore_done = 0;
  HIP_TRY(hipGetLastError());
  return download_vec(c, c->A.R, 4, s, R);
}

extern "C" int tsm_compute_dt(tsm_ctx* c, int s, double* dt) {
  if (!c || !dt || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  c->bind_state();
  c->A.ignore_done = 1;
  tsm::launch_dt(c->G, c->A, c->SP, c->S, c->st);
  c->A.ignore_done = 0;
  HIP_TRY(hipGetLastError());
  tsm::Ctl q;
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(&q, c->A.ctl + s, sizeof(q), hipMemcpyDeviceToHost));
  *dt = q.dt;
  return TSM_OK;
}

namespace {
// One full time step for every live sample with the CSR-gather kernels
// (src/fv.cc:2277-2316).
// dt_ready: the caller already ran the time-step pass (which also stores
// U_old) and may have adjusted dt.
void enqueue_step(tsm_ctx* c, int log_slot, bool dt_ready = false) {
  const int nrk = c->P.time_scheme == 0 ? 3 : (c->P.time_scheme == 1 ? 4 : 2);
  const int mode = c->P.time_scheme == 1 ? 2 : 1;
  c->bind_state();
  if (!dt_ready) tsm::launch_dt(c->G, c->A, c->SP, c->S, c->st);
  if (c->P.liou_fix) tsm::launch_ssw(c->G, c->A, c->S, c->st);
  for (int r = 0; r < nrk; ++r) {
    c->bind_state();
    tsm::launch_bc_time(c->G, c->bt, c->A, c->S, r, c->st);
    tsm::launch_gradient(c->G, c->A, c->S, c->st);
    tsm::launch_flux_update(c->P.flux, c->P.reconstruct, c->G, c->A, c->S, r, r == nrk - 1, mode,
                            c->st, true, true);
    if (c->P.strong_bc) tsm::launch_strong_bc(c->G, c->A, c->S, c->st);
    c->cur ^= 1;
  }
  c->bind_state();
  tsm::launch_step_final(c->G, c->A, c->SP, c->S, log_slot, c->st);
}

// One SSP-RK3 step with the tiled kernels.  W[cur] holds W^n throughout; the
// stages write the two other buffers and the last one writes W^{n+1} back
// in place (each vertex reads its own W^n for U_old before overwriting it).
void enqueue_tiled_step(tsm_ctx* c, int log_slot, hipEvent_t* ev = nullptr) {
  double* Wn = c->W[c->cur];
  double* Wa = c->W[(c->cur + 1) % 3];
  double* Wb = c->W[(c->cur + 2) % 3];
  // stage buffers: every stage reads W^n for U_old, the last one writes W^{n+1}
  // over it in place; SSP-RK3 and RK4 rotate through two scratch buffers
  const int nrk = c->P.time_scheme == 0 ? 3 : (c->P.time_scheme == 1 ? 4 : 2);
  const double* ins[4] = {Wn, Wa, Wb, Wa};
  double* outs[4] = {Wa, Wb, Wa, Wn};
  if (nrk == 3) outs[2] = Wn;
  if (nrk == 2) outs[1] = Wn;
  int e = 0;
  if (c->P.liou_fix) {  // the step's Liou sensor from W^n (src/fv.cc:2281-2285)
    c->A.W_in = Wn;
    tsm::launch_ssw(c->G, c->A, c->S, c->st);
  }
  for (int r = 0; r < nrk; ++r) {
    c->A.W_in = ins[r];
    c->A.W_out = outs[r];
    if (ev) (void)hipEventRecord(ev[e++], c->st);
    tsm::launch_bc_time(c->G, c->bt, c->A, c->S, r, c->st);
    if (tsm::tile_needs_gradient_pass(c->TG, c->P.flux, c->P.reconstruct)) {
      // boundary vertices, then the tiles (measured: on a second stream beside the
      // tile kernel the two compete for CU slots and the stage takes longer)
      tsm::launch_tile_gradient_bnd(c->G, c->TG, c->A, c->S, c->st);
      tsm::launch_tile_gradient(c->G, c->TG, c->A, c->S, c->st);
    }
    if (ev) (void)hipEventRecord(ev[e++], c->st);
    tsm::launch_tile_stage(c->P.flux, c->P.reconstruct, c->G, c->TG, c->A, c->SP, Wn, c->dtpart,
                           c->S, r, r == nrk - 1, c->st);
  }
  if (ev) (void)hipEventRecord(ev[e++], c->st);
  tsm::launch_step_final_tiled(c->G, c->A, c->SP, c->S, c->TG.ntiles, c->dtpart, log_slot, c->st);
  if (ev) (void)hipEventRecord(ev[e++], c->st);
  c->bind_state();
  c->e_valid = false;
}

// Steps of the current path; the tiled path first forms dt and the loop
// test of the first step with the CSR time-step kernel.
// One HIP graph of graph_steps tiled steps: every kernel argument of a step is
// fixed (the last stage writes W^n in place) and each sample's stop flag lives
// on the device, so the graph replays as it was captured.  Recaptured if the
// state buffer changes.
hipError_t tiled_graph(tsm_ctx* c) {
  if (c->gexec && c->gcur == c->cur) return hipSuccess;
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  for (int k = 0; k < c->P.graph_steps; ++k) enqueue_tiled_step(c, -1);
  e = hipStreamEndCapture(c->st, &g);
  if (e != hipSuccess) return e;
  e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e == hipSuccess) c->gcur = c->cur;
  return e;
}

void enqueue_steps(tsm_ctx* c, long nsteps, bool with_log) {
  if (c->tiled) {
    c->bind_state();
    if (nsteps > 0) tsm::launch_dt(c->G, c->A, c->SP, c->S, c->st);
    long k = 0;
    if (!with_log && c->P.graph_steps > 0 && nsteps >= c->P.graph_steps &&
        tiled_graph(c) == hipSuccess) {
      for (; k + c->P.graph_steps <= nsteps; k += c->P.graph_steps) (void)hipGraphLaunch(c->gexec, c->st);
      c->bind_state();
    }
    for (; k < nsteps; ++k) enqueue_tiled_step(c, with_log ? (int)k : -1);
  } else {
    c->csr_ready();
    for (long k = 0; k < nsteps; ++k) enqueue_step(c, with_log ? (int)k : -1);
  }
}
}  // namespace

extern "C" int tsm_run(tsm_ctx* c, long nsteps, long* done, double* log) {
  if (!c || nsteps < 0) return fail(TSM_E_ARG, "bad argument");
  const int S = c->S;
  std::vector<tsm::Ctl> before(S), after(S);
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(before.data(), c->A.ctl, S * sizeof(tsm::Ctl), hipMemcpyDeviceToHost));
  if (log && nsteps > 0) {
    if (c->A.log_cap < nsteps) {
      double* p = nullptr;
      int rc = c->alloc(&p, (size_t)S * nsteps * 4);
      if (rc) return rc;
      c->A.log = p;
      c->A.log_cap = (int)nsteps;
    }
  }
  enqueue_steps(c, nsteps, log != nullptr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(after.data(), c->A.ctl, S * sizeof(tsm::Ctl), hipMemcpyDeviceToHost));
  if (done)
    for (int s = 0; s < S; ++s) done[s] = (long)(after[s].iter - before[s].iter);
  if (log && nsteps > 0) {
    const size_t cap = (size_t)c->A.log_cap;
    std::vector<double> buf((size_t)S * cap * 4);
    HIP_TRY(hipMemcpy(buf.data(), c->A.log, buf.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int s = 0; s < S; ++s) {
      const long ns = (long)(after[s].iter - before[s].iter);
      for (long k = 0; k < nsteps * 4; ++k)
        log[(size_t)s * nsteps * 4 + k] = (k < ns * 4) ? buf[(size_t)s * cap * 4 + k] : 0.0;
    }
  }
  c->A.log = (log ? c->A.log : c->A.log);
  return TSM_OK;
}

extern "C" int tsm_step(tsm_ctx* c, double t_target, double* dt_out, double* res_out) {
  if (!c) return fail(TSM_E_ARG, "bad argument");
  if (t_target > 0.0 && !c->P.unsteady)
    return fail(TSM_E_ARG, "t_target needs an unsteady run (steady steps use local time steps)");
  const int S = c->S;
  std::vector<tsm::Ctl> before(S), after(S);
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(before.data(), c->A.ctl, S * sizeof(tsm::Ctl), hipMemcpyDeviceToHost));
  if (c->A.log_cap < 1) {  // the step's (iter, dt, t, res) rows
    double* p = nullptr;
    int rc = c->alloc(&p, (size_t)S * 4);
    if (rc) return rc;
    c->A.log = p;
    c->A.log_cap = 1;
  }
  if (c->tiled) {
    c->bind_state();
    tsm::launch_dt(c->G, c->A, c->SP, S, c->st);
    if (t_target > 0.0) tsm::launch_clamp_dt(c->A, S, t_target, c->st);
    enqueue_tiled_step(c, 0);
  } else {
    c->csr_ready();
    c->bind_state();
    tsm::launch_dt(c->G, c->A, c->SP, S, c->st);
    if (t_target > 0.0) tsm::launch_clamp_dt(c->A, S, t_target, c->st);
    enqueue_step(c, 0, /*dt_ready=*/true);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(after.data(), c->A.ctl, S * sizeof(tsm::Ctl), hipMemcpyDeviceToHost));
  std::vector<double> rows((size_t)S * c->A.log_cap * 4);
  HIP_TRY(hipMemcpy(rows.data(), c->A.log, rows.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int s = 0; s < S; ++s) {
    const bool ran = after[s].iter > before[s].iter;
    const double* r = &rows[(size_t)s * c->A.log_cap * 4];
    if (dt_out) dt_out[s] = ran ? r[1] : 0.0;
    if (res_out) res_out[s] = after[s].res_total;
  }
  return TSM_OK;
}

extern "C" int tsm_run_group(tsm_ctx** cs, int n, long nsteps, long* done) {
  if (!cs || n < 1 || nsteps < 0) return fail(TSM_E_ARG, "bad argument");
  std::vector<std::vector<tsm::Ctl>> before(n);
  for (int i = 0; i < n; ++i) {
    tsm_ctx* c = cs[i];
    if (!c) return fail(TSM_E_ARG, "null context in group");
    before[i].resize(c->S);
    HIP_TRY(hipStreamSynchronize(c->st));
    HIP_TRY(hipMemcpy(before[i].data(), c->A.ctl, c->S * sizeof(tsm::Ctl), hipMemcpyDeviceToHost));
  }
  // Interleave the contexts' launches in chunks so every stream has work queued
  // while the others run: levels of different sizes share the GPU.
  const long chunk = 8;
  for (long k0 = 0; k0 < nsteps; k0 += chunk) {
    const long m = std::min(chunk, nsteps - k0);
    for (int i = 0; i < n; ++i) {
      tsm_ctx* c = cs[i];
      if (c->tiled) {
        c->bind_state();
        if (k0 == 0) tsm::launch_dt(c->G, c->A, c->SP, c->S, c->st);
        for (long k = 0; k < m; ++k) enqueue_tiled_step(c, -1);
      } else {
        c->csr_ready();
        for (long k = 0; k < m; ++k) enqueue_step(c, -1);
      }
    }
  }
  HIP_TRY(hipGetLastError());
  for (int i = 0; i < n; ++i) {
    tsm_ctx* c = cs[i];
    std::vector<tsm::Ctl> after(c->S);
    HIP_TRY(hipStreamSynchronize(c->st));
    HIP_TRY(hipMemcpy(after.data(), c->A.ctl, c->S * sizeof(tsm::Ctl), hipMemcpyDeviceToHost));
    long tot = 0;
    for (int s = 0; s < c->S; ++s) tot += (long)(after[s].iter - before[i][s].iter);
    if (done) done[i] = tot;
  }
  return TSM_OK;
}

extern "C" int tsm_run_packed(tsm_ctx** cs, int n, long nsteps, long* done) {
  if (!cs || n < 1 || nsteps < 0) return fail(TSM_E_ARG, "bad argument");
  for (int i = 0; i < n; ++i) {
    const tsm_ctx* c = cs[i];
    if (!c) return fail(TSM_E_ARG, "null context in group");
    if (!c->tiled || c->TG.fused || c->P.time_scheme != 0 || c->A.ns || c->P.liou_fix)
      return fail(TSM_E_ARG, "packed steps need Euler SSP-RK3 contexts without the Liou fix on "
                             "the tiled kernels with a gradient pass of their own");
    if (c->P.flux != cs[0]->P.flux || c->P.reconstruct != cs[0]->P.reconstruct)
      return fail(TSM_E_ARG, "packed contexts must share the flux and the reconstruction");
    for (int j = 0; j < i; ++j)
      if (cs[j] == c) return fail(TSM_E_ARG, "a context appears twice in the group");
  }
  tsm_ctx* c0 = cs[0];
  hipStream_t st = c0->st;
  std::vector<std::vector<tsm::Ctl>> before(n);
  for (int i = 0; i < n; ++i) {
    tsm_ctx* c = cs[i];
    before[i].resize(c->S);
    HIP_TRY(hipStreamSynchronize(c->st));
    HIP_TRY(hipMemcpy(before[i].data(), c->A.ctl, c->S * sizeof(tsm::Ctl), hipMemcpyDeviceToHost));
  }
  // level descriptors of the three stages: every stage reads W^n for U_old and
  // the last writes W^{n+1} over it (the buffer rotation of enqueue_tiled_step)
  std::vector<tsm::PackLevel> lv(3 * (size_t)n);
  tsm::PackPlan plan{};
  for (int r = 0; r < 3; ++r) {
    for (int i = 0; i < n; ++i) {
      tsm_ctx* c = cs[i];
      double* Wn = c->W[c->cur];
      double* Wa = c->W[(c->cur + 1) % 3];
      double* Wb = c->W[(c->cur + 2) % 3];
      const double* ins[3] = {Wn, Wa, Wb};
      double* outs[3] = {Wa, Wb, Wn};
      tsm::PackLevel& L = lv[(size_t)r * n + i];
      L.G = c->G;
      L.TG = c->TG;
      L.A = c->A;
      L.A.W_in = ins[r];
      L.A.W_out = outs[r];
      L.P = c->SP;
      L.Wn = Wn;
      L.dtpart = c->dtpart;
      L.S = c->S;
    }
    tsm::pack_plan(&lv[(size_t)r * n], n, c0->P.flux, c0->P.reconstruct, plan);
  }
  if (c0->pack_cap < lv.size()) {
    int rc = c0->alloc(&c0->pack, lv.size());
    if (rc) return rc;
    c0->pack_cap = lv.size();
  }
  HIP_TRY(hipMemcpyAsync(c0->pack, lv.data(), lv.size() * sizeof(tsm::PackLevel),
                         hipMemcpyHostToDevice, st));
  for (int i = 0; i < n; ++i) {  // the first step's dt and loop test, per context
    tsm_ctx* c = cs[i];
    c->bind_state();
    if (nsteps > 0) tsm::launch_dt(c->G, c->A, c->SP, c->S, st);
  }
  for (long k = 0; k < nsteps; ++k) {
    for (int r = 0; r < 3; ++r) {
      const tsm::PackLevel* d = c0->pack + (size_t)r * n;
      tsm::launch_packed_gradient(d, plan, st);
      tsm::launch_packed_stage(c0->P.flux, c0->P.reconstruct, d, plan, r, r == 2, st);
    }
    tsm::launch_packed_final(c0->pack, plan, st);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  for (int i = 0; i < n; ++i) {
    tsm_ctx* c = cs[i];
    c->bind_state();
    c->e_valid = false;
    std::vector<tsm::Ctl> after(c->S);
    HIP_TRY(hipMemcpy(after.data(), c->A.ctl, c->S * sizeof(tsm::Ctl), hipMemcpyDeviceToHost));
    long tot = 0;
    for (int s = 0; s < c->S; ++s) tot += (long)(after[s].iter - before[i][s].iter);
    if (done) done[i] = tot;
  }
  return TSM_OK;
}

extern "C" int tsm_time(tsm_ctx* c, int s, double* t, long* iter) {
  if (!c || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  tsm::Ctl q;
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(&q, c->A.ctl + s, sizeof(q), hipMemcpyDeviceToHost));
  if (t) *t = q.t;
  if (iter) *iter = (long)q.iter;
  return TSM_OK;
}

extern "C" int tsm_bounds(tsm_ctx* c, int s, double* out) {
  if (!c || !out || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  const int nblk = std::max(1, std::min(1024, (c->G.nv + 255) / 256));
  if (!c->bounds_buf) {
    int rc = c->alloc(&c->bounds_buf, 8 * 1024);
    if (rc) return rc;
  }
  c->state_ready();
  tsm::launch_bounds(c->G, c->A, s, c->bounds_buf, nblk, c->st);
  HIP_TRY(hipGetLastError());
  std::vector<double> part(8 * (size_t)nblk);
  HIP_TRY(hipMemcpyAsync(part.data(), c->bounds_buf, part.size() * sizeof(double),
                         hipMemcpyDeviceToHost, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  // T, p, u, v as (min, max); w is zero in two dimensions
  double r[8] = {1.0e20, -1.0e20, 1.0e20, -1.0e20, 1.0e20, -1.0e20, 1.0e20, -1.0e20};
  for (int b = 0; b < nblk; ++b)
    for (int k = 0; k < 4; ++k) {
      r[2 * k] = std::min(r[2 * k], part[8 * (size_t)b + 2 * k]);
      r[2 * k + 1] = std::max(r[2 * k + 1], part[8 * (size_t)b + 2 * k + 1]);
    }
  for (int k = 0; k < 8; ++k) out[k] = r[k];
  out[8] = out[9] = 0.0;
  return TSM_OK;
}

extern "C" int tsm_get_clock(tsm_ctx* c, int s, long* iter, int* counter, double* t, double* res0) {
  if (!c || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  tsm::Ctl q;
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(&q, c->A.ctl + s, sizeof(q), hipMemcpyDeviceToHost));
  if (iter) *iter = (long)q.iter;
  if (counter) *counter = q.counter;
  if (t) *t = q.t;
  if (res0) *res0 = q.res0;
  return TSM_OK;
}

extern "C" int tsm_set_clock(tsm_ctx* c, int s, long iter, int counter, double t, double res0) {
  if (!c || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  if (iter < 0 || counter < 0 || counter > c->P.n_time_stamps + 1 || !(t >= 0.0) || !std::isfinite(t) ||
      !(res0 > 0.0) || !std::isfinite(res0))
    return fail(TSM_E_ARG, "clock: iter and counter must be >= 0 (counter <= time stamps + 1), t finite "
                           "and >= 0, res0 finite and > 0");
  tsm::Ctl q;
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(&q, c->A.ctl + s, sizeof(q), hipMemcpyDeviceToHost));
  q.iter = iter; q.counter = counter; q.t = t; q.res0 = res0;
  q.res_total = 1e20; q.reached = 0; q.done = 0;
  HIP_TRY(hipMemcpy(c->A.ctl + s, &q, sizeof(q), hipMemcpyHostToDevice));
  return TSM_OK;
}

extern "C" int tsm_status(tsm_ctx* c, int s, double* v, long* iter, int* counter, int* done) {
  if (!c || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  tsm::Ctl q;
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(&q, c->A.ctl + s, sizeof(q), hipMemcpyDeviceToHost));
  if (v) { v[0] = q.t; v[1] = q.dt; v[2] = q.res_total; }
  if (iter) *iter = (long)q.iter;
  if (counter) *counter = q.counter;
  if (done) *done = q.done;
  return TSM_OK;
}

extern "C" int tsm_status_all(tsm_ctx* c, double* v, long* iter, int* counter, int* done,
                              double* norms) {
  if (!c) return fail(TSM_E_ARG, "bad argument");
  std::vector<tsm::Ctl> q(c->S);
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(q.data(), c->A.ctl, q.size() * sizeof(tsm::Ctl), hipMemcpyDeviceToHost));
  for (int s = 0; s < c->S; ++s) {
    if (v) { v[3 * s] = q[s].t; v[3 * s + 1] = q[s].dt; v[3 * s + 2] = q[s].res_total; }
    if (iter) iter[s] = (long)q[s].iter;
    if (counter) counter[s] = q[s].counter;
    if (done) done[s] = q[s].done;
    if (norms)
      for (int k = 0; k < 4; ++k) norms[4 * s + k] = q[s].rn[k];
  }
  return TSM_OK;
}

extern "C" int tsm_residual_norms(tsm_ctx* c, int s, double* norms) {
  if (!c || s < 0 || s >= c->S || !norms) return fail(TSM_E_ARG, "bad argument");
  tsm::Ctl q;
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(&q, c->A.ctl + s, sizeof(q), hipMemcpyDeviceToHost));
  for (int k = 0; k < 4; ++k) norms[k] = q.rn[k];
  return TSM_OK;
}

namespace {
int ensure_red(tsm_ctx* c, size_t n) {
  if (c->red_cap >= n) return TSM_OK;
  int rc = c->alloc(&c->red, n);
  if (rc) return rc;
  c->red_cap = n;
  return TSM_OK;
}
}  // namespace

extern "C" int tsm_global(tsm_ctx* c, int s, double* ke, double* ent) {
  if (!c || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  int nb = c->nblk_v, rc = ensure_red(c, (size_t)nb * 24);
  if (rc) return rc;
  c->bind_state();
  tsm::launch_global(c->G, c->A, s, c->red, c->st);
  HIP_TRY(hipGetLastError());
  std::vector<double> h((size_t)nb * 2);
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(h.data(), c->red, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  double K = 0, E = 0;
  for (int b = 0; b < nb; ++b) { K += h[2 * b]; E += h[2 * b + 1]; }
  if (ke) *ke = K;
  if (ent) *ent = E;
  return TSM_OK;
}

extern "C" int tsm_enstrophy(tsm_ctx* c, int s, double* ens) {
  if (!c || !ens || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  const int nb = c->nblk_v;
  int rc = ensure_red(c, (size_t)nb * 24);
  if (rc) return rc;
  c->sample_ready(s);  // gradients of the current state of sample s
  c->A.ignore_done = 1;
  tsm::launch_gradient_one(c->G, c->A, s, c->st);
  c->A.ignore_done = 0;
  tsm::launch_enstrophy(c->G, c->A, s, c->red, c->st);
  HIP_TRY(hipGetLastError());
  std::vector<double> h((size_t)nb);
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(h.data(), c->red, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  double Z = 0;
  for (int b = 0; b < nb; ++b) Z += h[b];
  *ens = Z;
  return TSM_OK;
}

extern "C" int tsm_error_norms(tsm_ctx* c, int s, const double* exact, double* out) {
  if (!c || !exact || !out || s < 0 || s >= c->S) return fail(TSM_E_ARG, "bad argument");
  const int nv = c->G.nv, nb = c->nblk_v;
  int rc = ensure_red(c, (size_t)nb * 24);
  if (rc) return rc;
  std::vector<double> ex(4 * (size_t)nv);
  for (int i = 0; i < nv; ++i)
    for (int k = 0; k < 4; ++k) ex[4 * (size_t)c->perm[i] + k] = exact[4 * (size_t)i + k];
  double* dex = nullptr;
  HIP_TRY(hipMalloc(&dex, 4 * (size_t)nv * sizeof(double)));
  hipError_t e = hipMemcpy(dex, ex.data(), 4 * (size_t)nv * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) { (void)hipFree(dex); return fail(TSM_E_HIP, "upload exact"); }
  c->bind_state();
  tsm::launch_errnorm(c->G, c->A, s, dex, c->red, c->st);
  e = hipGetLastError();
  std::vector<double> h((size_t)nb * 24);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  if (e == hipSuccess) e = hipMemcpy(h.data(), c->red, h.size() * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(dex);
  if (e != hipSuccess) return fail(TSM_E_HIP, hipGetErrorString(e));
  double L1[8] = {0}, L2[8] = {0}, Li[8] = {0};
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 8; ++k) {
      L1[k] += h[24 * b + k];
      L2[k] += h[24 * b + 8 + k];
      Li[k] = std::fmax(Li[k], h[24 * b + 16 + k]);
    }
  for (int k = 0; k < 8; ++k) {
    out[3 * k] = L1[k]; out[3 * k + 1] = std::sqrt(L2[k]); out[3 * k + 2] = Li[k];
  }
  return TSM_OK;
}

extern "C" int tsm_geometry_sizes(tsm_ctx* c, int* nf, int* nb) {
  if (!c) return fail(TSM_E_ARG, "bad argument");
  if (nf) *nf = (int)c->geo.face.size();
  if (nb) *nb = (int)c->geo.bface.size();
  return TSM_OK;
}

extern "C" int tsm_geometry(tsm_ctx* c, int* fv, double* fn, double* dc, double* mc) {
  if (!c) return fail(TSM_E_ARG, "bad argument");
  const auto& g = c->geo;
  for (size_t f = 0; f < g.face.size(); ++f) {
    if (fv) { fv[2 * f] = g.face[f].v[0]; fv[2 * f + 1] = g.face[f].v[1]; }
    if (fn) { fn[2 * f] = g.face[f].nx; fn[2 * f + 1] = g.face[f].ny; }
  }
  for (int i = 0; i < g.nv; ++i) {
    if (dc) dc[i] = g.dcarea[i];
    if (mc) mc[i] = g.mcarea[i];
  }
  return TSM_OK;
}

extern "C" int tsm_bench(tsm_ctx* c, int nsteps, double* us) {
  if (!c || nsteps < 1 || !us) return fail(TSM_E_ARG, "bad argument");
  hipEvent_t ev[16];
  for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
  double acc[4] = {0, 0, 0, 0};
  float ms = 0;
  HIP_TRY(hipStreamSynchronize(c->st));
  if (c->tiled) {
    c->bind_state();
    tsm::launch_dt(c->G, c->A, c->SP, c->S, c->st);
    const int nrk = c->P.time_scheme == 0 ? 3 : (c->P.time_scheme == 1 ? 4 : 2);
    for (int k = 0; k < nsteps; ++k) {
      enqueue_tiled_step(c, -1, ev);
      HIP_TRY(hipEventSynchronize(ev[2 * nrk + 1]));
      for (int r = 0; r < nrk; ++r) {
        HIP_TRY(hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1])); acc[1] += ms;
        HIP_TRY(hipEventElapsedTime(&ms, ev[2 * r + 1], ev[2 * r + 2])); acc[2] += ms;
      }
      HIP_TRY(hipEventElapsedTime(&ms, ev[2 * nrk], ev[2 * nrk + 1])); acc[3] += ms;
    }
    us[0] = 0.0;
    us[1] = 1e3 * acc[1] / ((double)nrk * nsteps);
    us[2] = 1e3 * acc[2] / ((double)nrk * nsteps);
    us[3] = 1e3 * acc[3] / nsteps;
  } else {
    c->csr_ready();
    const int nrk = c->P.time_scheme == 0 ? 3 : (c->P.time_scheme == 1 ? 4 : 2);
    const int mode = c->P.time_scheme == 1 ? 2 : 1;
    for (int k = 0; k < nsteps; ++k) {
      c->bind_state();
      HIP_TRY(hipEventRecord(ev[0], c->st));
      tsm::launch_dt(c->G, c->A, c->SP, c->S, c->st);
      HIP_TRY(hipEventRecord(ev[1], c->st));
      if (c->P.liou_fix) tsm::launch_ssw(c->G, c->A, c->S, c->st);
      int e = 2;
      for (int r = 0; r < nrk; ++r) {
        c->bind_state();
        HIP_TRY(hipEventRecord(ev[e++], c->st));
        tsm::launch_gradient(c->G, c->A, c->S, c->st);
        HIP_TRY(hipEventRecord(ev[e++], c->st));
        tsm::launch_flux_update(c->P.flux, c->P.reconstruct, c->G, c->A, c->S, r, r == nrk - 1,
                                mode, c->st, true, false);
        HIP_TRY(hipEventRecord(ev[e++], c->st));
        tsm::launch_flux_update(c->P.flux, c->P.reconstruct, c->G, c->A, c->S, r, r == nrk - 1,
                                mode, c->st, false, true);
        if (c->P.strong_bc) tsm::launch_strong_bc(c->G, c->A, c->S, c->st);
        c->cur ^= 1;
      }
      c->bind_state();
      tsm::launch_step_final(c->G, c->A, c->SP, c->S, -1, c->st);
      HIP_TRY(hipEventRecord(ev[e++], c->st));
      HIP_TRY(hipEventSynchronize(ev[e - 1]));
      HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1])); acc[0] += ms;
      for (int r = 0; r < nrk; ++r) {
        HIP_TRY(hipEventElapsedTime(&ms, ev[2 + 3 * r], ev[3 + 3 * r])); acc[1] += ms;
        HIP_TRY(hipEventElapsedTime(&ms, ev[3 + 3 * r], ev[4 + 3 * r])); acc[2] += ms;
        HIP_TRY(hipEventElapsedTime(&ms, ev[4 + 3 * r], ev[(r == nrk - 1) ? e - 1 : 5 + 3 * r])); acc[3] += ms;
      }
    }
    us[0] = 1e3 * acc[0] / nsteps;
    us[1] = 1e3 * acc[1] / (nsteps * (double)nrk);
    us[2] = 1e3 * acc[2] / (nsteps * (double)nrk);
    us[3] = 1e3 * acc[3] / (nsteps * (double)nrk);
  }
  // whole steps without interleaved events
  HIP_TRY(hipEventRecord(ev[14], c->st));
  enqueue_steps(c, nsteps, false);
  HIP_TRY(hipEventRecord(ev[15], c->st));
  HIP_TRY(hipEventSynchronize(ev[15]));
  HIP_TRY(hipEventElapsedTime(&ms, ev[14], ev[15]));
  us[4] = 1e3 * ms / nsteps;
  for (auto& e : ev) (void)hipEventDestroy(e);
  return TSM_OK;
}

extern "C" int tsm_moments(tsm_ctx* c, double* mean, double* m2, double* s2) {
  if (!c) return fail(TSM_E_ARG, "bad argument");
  const int nv = c->G.nv;
  double *dm = nullptr, *dq = nullptr, *dz = nullptr;
  const size_t n = 5 * (size_t)nv;
  HIP_TRY(hipMalloc(&dm, 3 * n * sizeof(double)));
  dq = dm + n;
  dz = dq + n;
  c->bind_state();
  tsm::launch_moments(c->G, c->A, c->S, dm, dq, dz, c->st);
  hipError_t e = hipGetLastError();
  std::vector<double> h(3 * n);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  if (e == hipSuccess) e = hipMemcpy(h.data(), dm, h.size() * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(dm);
  if (e != hipSuccess) return fail(TSM_E_HIP, hipGetErrorString(e));
  double* outs[3] = {mean, m2, s2};
  for (int a = 0; a < 3; ++a) {
    if (!outs[a]) continue;
    for (int i = 0; i < nv; ++i)
      for (int q = 0; q < 5; ++q) outs[a][5 * (size_t)i + q] = h[a * n + 5 * (size_t)c->perm[i] + q];
  }
  return TSM_OK;
}

extern "C" int tsm_probe_density(tsm_ctx* c, int np, const double* px, const double* py,
                                 double* out) {
  if (!c || np < 1 || !px || !py || !out) return fail(TSM_E_ARG, "bad argument");
  // device point location (k_probe_locate: the first cell in cell order that
  // holds the point), then the interpolation of every sample
  const size_t n = (size_t)np;
  double *dpx = nullptr, *dpy = nullptr, *dpw = nullptr, *dout = nullptr;
  int *dpc = nullptr, *dcell = nullptr;
  hipError_t e = hipMalloc(&dpx, 8 * n);
  if (e == hipSuccess) e = hipMalloc(&dpy, 8 * n);
  if (e == hipSuccess) e = hipMalloc(&dpw, 24 * n);
  if (e == hipSuccess) e = hipMalloc(&dout, 8 * n * (size_t)c->S);
  if (e == hipSuccess) e = hipMalloc(&dpc, 12 * n);
  if (e == hipSuccess) e = hipMalloc(&dcell, 4 * n);
  if (e == hipSuccess) e = hipMemcpyAsync(dpx, px, 8 * n, hipMemcpyHostToDevice, c->st);
  if (e == hipSuccess) e = hipMemcpyAsync(dpy, py, 8 * n, hipMemcpyHostToDevice, c->st);
  if (e == hipSuccess) {
    tsm::launch_probe_locate(c->G, np, dpx, dpy, dpc, dpw, dcell, c->st);
    c->bind_state();
    tsm::launch_probes(c->A, c->G.nv, np, dpc, dpw, c->S, dout, c->st);
    e = hipGetLastError();
  }
  std::vector<int> cells(n);
  if (e == hipSuccess) e = hipMemcpyAsync(cells.data(), dcell, 4 * n, hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, 8 * n * (size_t)c->S, hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  (void)hipFree(dpx); (void)hipFree(dpy); (void)hipFree(dpw); (void)hipFree(dout);
  (void)hipFree(dpc); (void)hipFree(dcell);
  if (e != hipSuccess) return fail(TSM_E_HIP, hipGetErrorString(e));
  for (int p = 0; p < np; ++p)
    if (cells[p] < 0) return fail(TSM_E_ARG, "probe point " + std::to_string(p) + " outside the mesh");
  return TSM_OK;
}

extern "C" int tsm_probe_density_noalloc(tsm_ctx* c, int np, const double* px, const double* py,
                                         double* out) {
  if (!c || np < 1 || !px || !py || !out) return fail(TSM_E_ARG, "bad argument");
  // scratch layout: px, py, weights (3 per probe), densities (one per probe and
  // sample), corner ids (3 int32 per probe), containing cell (one int32 per
  // probe, -1 outside the mesh); each part 256-byte aligned
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t n = (size_t)np;
  const size_t o_py = up(8 * n), o_pw = o_py + up(8 * n), o_out = o_pw + up(24 * n);
  const size_t o_pc = o_out + up(8 * n * (size_t)c->S), o_cell = o_pc + up(12 * n);
  const size_t need = o_cell + up(4 * n);
  if (need > tsm::noalloc_scratch_bytes())
    return fail(TSM_E_ARG, "tsm_probe_density_noalloc: " + std::to_string(np) + " probes x " +
                               std::to_string(c->S) + " samples need " + std::to_string(need) +
                               " bytes of scratch, more than its fixed " +
                               std::to_string(tsm::noalloc_scratch_bytes()));
  int prev = -1;
  HIP_TRY(hipGetDevice(&prev));
  if (prev != c->dev) HIP_TRY(hipSetDevice(c->dev));
  struct Restore {  // the caller's device again on every return
    int d, p;
    ~Restore() { if (d != p) (void)hipSetDevice(p); }
  } restore{c->dev, prev};
  void* base = nullptr;
  int he = tsm::noalloc_scratch(&base);
  if (he) return fail(TSM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString((hipError_t)he));
  char* b = static_cast<char*>(base);
  double* dpx = reinterpret_cast<double*>(b);
  double* dpy = reinterpret_cast<double*>(b + o_py);
  double* dpw = reinterpret_cast<double*>(b + o_pw);
  double* dout = reinterpret_cast<double*>(b + o_out);
  int* dpc = reinterpret_cast<int*>(b + o_pc);
  int* dcell = reinterpret_cast<int*>(b + o_cell);
  HIP_TRY(hipMemcpyAsync(dpx, px, 8 * n, hipMemcpyHostToDevice, c->st));
  HIP_TRY(hipMemcpyAsync(dpy, py, 8 * n, hipMemcpyHostToDevice, c->st));
  tsm::launch_probe_locate(c->G, np, dpx, dpy, dpc, dpw, dcell, c->st);
  HIP_TRY(hipGetLastError());
  c->bind_state();
  tsm::launch_probes(c->A, c->G.nv, np, dpc, dpw, c->S, dout, c->st);
  HIP_TRY(hipGetLastError());
  // the containing cells first (into the caller's output, which is large
  // enough: 8 np S >= 4 np bytes), to report points outside the mesh
  HIP_TRY(hipMemcpyAsync(out, dcell, 4 * n, hipMemcpyDeviceToHost, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  const int* cells = reinterpret_cast<const int*>(out);
  for (int p = 0; p < np; ++p)
    if (cells[p] < 0)
      return fail(TSM_E_ARG, "probe point " + std::to_string(p) + " outside the mesh");
  HIP_TRY(hipMemcpyAsync(out, dout, 8 * n * (size_t)c->S, hipMemcpyDeviceToHost, c->st));
  HIP_TRY(hipStreamSynchronize(c->st));
  return TSM_OK;
}

extern "C" int tsm_check_arith(long n, const double* a, const double* b, double* out) {
  if (n < 1 || !a || !b || !out) return fail(TSM_E_ARG, "bad argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(TSM_E_HIP, "no HIP device visible");
  double* d = nullptr;
  HIP_TRY(hipMalloc(&d, 6 * (size_t)n * sizeof(double)));
  hipError_t e = hipMemcpy(d, a, 8 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + n, b, 8 * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    tsm::launch_check_arith(n, d, d + n, d + 2 * n, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, d + 2 * n, 32 * (size_t)n, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(TSM_E_HIP, hipGetErrorString(e));
  return TSM_OK;
}

extern "C" int tsm_global_all(tsm_ctx* c, double* ke, double* ent) {
  if (!c) return fail(TSM_E_ARG, "bad argument");
  const int nb = c->nblk_v, S = c->S;
  int rc = ensure_red(c, (size_t)nb * S * 2);  // the context's reduction buffer: no allocation per call
  if (rc) return rc;
  c->bind_state();
  tsm::launch_global_all(c->G, c->A, S, c->red, c->st);
  HIP_TRY(hipGetLastError());
  std::vector<double> h((size_t)nb * S * 2);
  HIP_TRY(hipStreamSynchronize(c->st));
  HIP_TRY(hipMemcpy(h.data(), c->red, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int s = 0; s < S; ++s) {
    double K = 0, E = 0;
    for (int b = 0; b < nb; ++b) { K += h[((size_t)s * nb + b) * 2]; E += h[((size_t)s * nb + b) * 2 + 1]; }
    if (ke) ke[s] = K;
    if (ent) ent[s] = E;
  }
  return TSM_OK;
}
