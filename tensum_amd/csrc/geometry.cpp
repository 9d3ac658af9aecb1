// geometry.cpp -- host-side dual-mesh preprocessing of libtsm.
//
// Produces the face list, normals, dual areas and periodic sets of one mesh
// partition in the order the reference's Grid::preproc produces them
// (src/grid_preproc.cc:1147-1194), because the device residual gather sums
// face contributions per vertex in that order.  Face discovery uses an edge
// hash instead of the reference's per-vertex linear search
// (src/grid_preproc.cc:396-454); the resulting order is the same: boundary
// faces first in file order, then interior faces in first-encounter order over
// cells with edges (v0,v1), (v1,v2), (v2,v0).
#include <cmath>
#include <unordered_map>

#include "tsm_host.h"

namespace tsm {

namespace {
inline uint64_t edge_key(int a, int b) {
  uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
  return ((uint64_t)hi << 32) | lo;
}
}  // namespace

bool build_geometry(const tsm_mesh& m, int cell_type, Geometry& g, std::string& msg) {
  g.nv = m.n_vertex;
  g.nc = m.n_cell;
  if (g.nv <= 0 || g.nc <= 0) { msg = "empty mesh"; return false; }
  g.x.assign(m.x, m.x + g.nv);
  g.y.assign(m.y, m.y + g.nv);
  g.cell.assign(m.cell_v, m.cell_v + 3 * (size_t)g.nc);
  for (int v : g.cell)
    if (v < 0 || v >= g.nv) { msg = "cell vertex out of range"; return false; }

  auto bc_of = [&](int tag, int* type) -> bool {
    for (int i = 0; i < m.n_bc; ++i)
      if (m.bc_tag[i] == tag) { *type = m.bc_type[i]; return true; }
    return false;
  };

  // ---- faces (src/grid_preproc.cc:461-590) --------------------------------
  std::vector<Face>& face = g.face;
  face.clear();
  face.reserve((size_t)m.n_bface + 2 * (size_t)g.nc);
  std::unordered_map<uint64_t, int> lookup;
  lookup.reserve((size_t)3 * g.nc);
  for (int i = 0; i < m.n_bface; ++i) {
    Face f{};
    f.type = m.bface_tag[i];
    f.v[0] = m.bface_v[2 * i];
    f.v[1] = m.bface_v[2 * i + 1];
    if (f.v[0] < 0 || f.v[0] >= g.nv || f.v[1] < 0 || f.v[1] >= g.nv) {
      msg = "boundary face vertex out of range"; return false;
    }
    int bt;
    if (!bc_of(f.type, &bt)) { msg = "boundary tag " + std::to_string(f.type) + " has no bc"; return false; }
    f.periodic = (bt == TSM_BC_PERIODIC);
    f.lcell = f.rcell = -1;
    lookup.emplace(edge_key(f.v[0], f.v[1]), i);
    face.push_back(f);
  }
  for (int c = 0; c < g.nc; ++c) {
    const int* v = &g.cell[3 * (size_t)c];
    const int e[3][2] = {{v[0], v[1]}, {v[1], v[2]}, {v[2], v[0]}};
    for (int k = 0; k < 3; ++k) {
      auto it = lookup.find(edge_key(e[k][0], e[k][1]));
      if (it != lookup.end()) {
        Face& f = face[it->second];
        if (f.lcell == -1) { f.lcell = c; continue; }
        if (f.rcell == -1) { f.rcell = c; continue; }
        msg = "edge shared by more than two cells"; return false;
      }
      Face f{};
      f.type = -1; f.v[0] = e[k][0]; f.v[1] = e[k][1];
      f.lcell = c; f.rcell = -1; f.periodic = false;
      lookup.emplace(edge_key(f.v[0], f.v[1]), (int)face.size());
      face.push_back(f);
    }
  }
  {  // interior-typed faces with one cell are ghost-domain faces: dropped
    std::vector<Face> kept;
    kept.reserve(face.size());
    for (auto& f : face)
      if (!(f.type == -1 && f.rcell == -1)) kept.push_back(f);
    face.swap(kept);
  }
  for (auto& f : face)
    if (f.lcell == -1) { msg = "boundary face not attached to any cell"; return false; }

  // cell face slots (src/grid_preproc.cc:600-623)
  g.cell_slot_face.assign(3 * (size_t)g.nc, -1);
  for (int i = 0; i < (int)face.size(); ++i) {
    int* L = &g.cell_slot_face[3 * (size_t)face[i].lcell];
    int j = 0; while (L[j] != -1) ++j; L[j] = i;
    if (face[i].rcell != -1) {
      int* R = &g.cell_slot_face[3 * (size_t)face[i].rcell];
      j = 0; while (R[j] != -1) ++j; R[j] = i;
    }
  }
  // boundary faces (src/grid_preproc.cc:580-589)
  g.bface.clear();
  for (auto& f : face) {
    if (f.type == -1) continue;
    BFace b{};
    b.v[0] = f.v[0]; b.v[1] = f.v[1];
    b.tag = f.type;
    bc_of(f.type, &b.bc);
    b.periodic = f.periodic;
    b.lcell = f.lcell;
    g.bface.push_back(b);
  }

  // ---- centroids (src/grid_preproc.cc:18-184) -----------------------------
  for (auto& f : face) {
    f.cx = g.x[f.v[0]] * 0.5 + g.x[f.v[1]] * 0.5;
    f.cy = g.y[f.v[0]] * 0.5 + g.y[f.v[1]] * 0.5;
  }
  const double third = 1.0 / 3.0;
  g.ccx.resize(g.nc); g.ccy.resize(g.nc);
  for (int c = 0; c < g.nc; ++c) {
    const int* v = &g.cell[3 * (size_t)c];
    double x1 = g.x[v[0]], y1 = g.y[v[0]], x2 = g.x[v[1]], y2 = g.y[v[1]];
    double x3 = g.x[v[2]], y3 = g.y[v[2]];
    if (cell_type == 0) {
      g.ccx[c] = x1 * third + x2 * third + x3 * third;
      g.ccy[c] = y1 * third + y2 * third + y3 * third;
      continue;
    }
    double l1 = std::pow(x2 - x3, 2) + std::pow(y2 - y3, 2);
    double l2 = std::pow(x3 - x1, 2) + std::pow(y3 - y1, 2);
    double l3 = std::pow(x1 - x2, 2) + std::pow(y1 - y2, 2);
    double b1 = std::fmax(0.0, l2 + l3 - l1);
    double b2 = std::fmax(0.0, l3 + l1 - l2);
    double b3 = std::fmax(0.0, l1 + l2 - l3);
    if (b1 < b2 / 2.0 && b1 < b3 / 2.0) b1 = 0.0;
    if (b2 < b3 / 2.0 && b2 < b1 / 2.0) b2 = 0.0;
    if (b3 < b1 / 2.0 && b3 < b2 / 2.0) b3 = 0.0;
    double det = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
    double q1 = 0.5 * ((x2 - x1) * (x2 + x1) + (y2 - y1) * (y2 + y1));
    double q2 = 0.5 * ((x3 - x1) * (x3 + x1) + (y3 - y1) * (y3 + y1));
    double xc = ((y3 - y1) * q1 - (y2 - y1) * q2) / det;
    double yc = (-(x3 - x1) * q1 + (x2 - x1) * q2) / det;
    if (b1 == 0.0) { g.ccx[c] = 0.5 * (x2 + x3); g.ccy[c] = 0.5 * (y2 + y3); }
    else if (b2 == 0.0) { g.ccx[c] = 0.5 * (x3 + x1); g.ccy[c] = 0.5 * (y3 + y1); }
    else if (b3 == 0.0) { g.ccx[c] = 0.5 * (x1 + x2); g.ccy[c] = 0.5 * (y1 + y2); }
    else { g.ccx[c] = xc; g.ccy[c] = yc; }
  }
  if (cell_type == 1) {
    for (auto& b : g.bface) {
      int c = b.lcell;
      double fx = g.x[b.v[0]] * 0.5 + g.x[b.v[1]] * 0.5;
      double fy = g.y[b.v[0]] * 0.5 + g.y[b.v[1]] * 0.5;
      double dx = fx * 1.0 + g.ccx[c] * -1.0, dy = fy * 1.0 + g.ccy[c] * -1.0;
      if (std::sqrt(dx * dx + dy * dy) < 1.0e-14) {
        const int* v = &g.cell[3 * (size_t)c];
        g.ccx[c] = g.x[v[0]] * third + g.x[v[1]] * third + g.x[v[2]] * third;
        g.ccy[c] = g.y[v[0]] * third + g.y[v[1]] * third + g.y[v[2]] * third;
      }
    }
  }

  // ---- areas (src/grid_preproc.cc:175-254, src/reduce_dual_area.cc) -------
  g.mcarea.assign(g.nv, 0.0);
  g.cell_area.resize(g.nc);
  for (int c = 0; c < g.nc; ++c) {
    const int* v = &g.cell[3 * (size_t)c];
    double l1x = g.x[v[1]] * 1.0 + g.x[v[0]] * -1.0, l1y = g.y[v[1]] * 1.0 + g.y[v[0]] * -1.0;
    double l2x = g.x[v[2]] * 1.0 + g.x[v[0]] * -1.0, l2y = g.y[v[2]] * 1.0 + g.y[v[0]] * -1.0;
    double a = 0.5 * (l1x * l2y - l1y * l2x);
    if (!(a > 0.0)) { msg = "cell " + std::to_string(c) + " has non-positive area"; return false; }
    g.cell_area[c] = a;
    for (int j = 0; j < 3; ++j) g.mcarea[v[j]] += a / 3.0;
  }
  if (cell_type == 0) {
    g.dcarea = g.mcarea;
  } else {
    g.dcarea.assign(g.nv, 0.0);
    for (int c = 0; c < g.nc; ++c) {
      const int* v = &g.cell[3 * (size_t)c];
      for (int j = 0; j < 3; ++j) {
        int n0 = v[j], n1 = (j == 0) ? v[2] : v[j - 1], n2 = (j == 2) ? v[0] : v[j + 1];
        double px[5], py[5];
        px[0] = g.x[n0]; py[0] = g.y[n0];
        px[1] = g.x[n0] * 0.5 + g.x[n2] * 0.5; py[1] = g.y[n0] * 0.5 + g.y[n2] * 0.5;
        px[2] = g.ccx[c]; py[2] = g.ccy[c];
        px[3] = g.x[n0] * 0.5 + g.x[n1] * 0.5; py[3] = g.y[n0] * 0.5 + g.y[n1] * 0.5;
        px[4] = px[0]; py[4] = py[0];
        double ar = 0;
        for (int k = 0; k < 4; ++k) ar += px[k] * py[k + 1] - px[k + 1] * py[k];
        if (ar < 0.0) { msg = "dual cell area is non-positive"; return false; }
        g.dcarea[n0] += ar * 0.5;
      }
    }
  }
  g.periodic.assign(m.n_periodic, {});
  g.periodic_of.assign(g.nv, -1);
  for (int i = 0; i < m.n_periodic; ++i) {
    for (int k = m.periodic_ptr[i]; k < m.periodic_ptr[i + 1]; ++k) {
      int v = m.periodic_v[k];
      if (v < 0 || v >= g.nv) { msg = "periodic vertex out of range"; return false; }
      if (g.periodic_of[v] != -1) { msg = "vertex in two periodic sets"; return false; }
      g.periodic_of[v] = i;
      g.periodic[i].push_back(v);
    }
    if (g.periodic[i].empty()) { msg = "empty periodic set"; return false; }
  }
  for (auto& P : g.periodic) {
    double mc = 0, dc = 0;
    for (int v : P) { mc += g.mcarea[v] * 1.0; dc += g.dcarea[v] * 1.0; }
    for (int v : P) { g.mcarea[v] = mc; g.dcarea[v] = dc; }
  }
  g.weight.assign(g.nv, 1.0);
  for (auto& P : g.periodic) {
    double w = 1.0 / ((double)P.size() * 1.0);
    for (int v : P) g.weight[v] = w;
  }

  // ---- normals (src/grid_preproc.cc:259-365) ------------------------------
  for (auto& f : face) {
    int v0 = f.v[0], v1 = f.v[1];
    double drx = g.x[v1] * 1.0 + g.x[v0] * -1.0, dry = g.y[v1] * 1.0 + g.y[v0] * -1.0;
    double dcx = f.cx * 1.0 + g.ccx[f.lcell] * -1.0, dcy = f.cy * 1.0 + g.ccy[f.lcell] * -1.0;
    if (dcx * dry + dcy * -drx < 0.0) { f.v[0] = v1; f.v[1] = v0; }
    double nx = f.cx * 1.0 + g.ccx[f.lcell] * -1.0, ny = f.cy * 1.0 + g.ccy[f.lcell] * -1.0;
    if (f.type == -1) {
      nx += g.ccx[f.rcell] * 1.0 + f.cx * -1.0;
      ny += g.ccy[f.rcell] * 1.0 + f.cy * -1.0;
    }
    f.nx = -ny;
    f.ny = nx;
    f.measure = std::sqrt(f.nx * f.nx + f.ny * f.ny);
  }
  for (auto& b : g.bface) {
    int v0 = b.v[0], v1 = b.v[1];
    double drx = g.x[v1] * 1.0 + g.x[v0] * -1.0, dry = g.y[v1] * 1.0 + g.y[v0] * -1.0;
    b.nx = dry; b.ny = -drx;
    b.measure = std::sqrt(b.nx * b.nx + b.ny * b.ny);
    double fx = g.x[v0] * 0.5 + g.x[v1] * 0.5, fy = g.y[v0] * 0.5 + g.y[v1] * 0.5;
    double dcx = fx * 1.0 + g.ccx[b.lcell] * -1.0, dcy = fy * 1.0 + g.ccy[b.lcell] * -1.0;
    if (dcx * b.nx + dcy * b.ny < 0.0) {
      b.v[0] = v1; b.v[1] = v0;
      b.nx *= -1.0; b.ny *= -1.0;
    }
  }
  g.removed_faces = 0;
  if (cell_type == 1) {  // src/grid_preproc.cc:640-655 remove_empty_faces
    // Boundary entries lead face[] and must all survive: the cells' slot ids
    // (and the reference's d_bnd_* lists, src/grid_preproc.cc:564-572) keep
    // their pre-removal values, which name the right boundary faces only while
    // nothing before them is removed.
    std::vector<Face> kept;
    for (auto& f : face) {
      if (f.measure > 0.0) kept.push_back(f);
      else if (f.type != -1) {
        msg = "a boundary face of the Voronoi dual has zero length";
        return false;
      }
    }
    g.removed_faces = (int)(face.size() - kept.size());
    face.swap(kept);
  }
  // src/grid_preproc.cc:811-831 (slot ids index the possibly shortened list)
  g.cell_has_nonper_bface.assign(g.nc, 0);
  for (int c = 0; c < g.nc; ++c)
    for (int j = 0; j < 3; ++j) {
      int f = g.cell_slot_face[3 * (size_t)c + j];
      if (f < 0 || f >= (int)face.size()) continue;
      if (face[f].type != -1) {
        int bt = 0;
        bc_of(face[f].type, &bt);
        if (bt != TSM_BC_PERIODIC) g.cell_has_nonper_bface[c] = 1;
      }
    }
  return true;
}

}  // namespace tsm

extern "C" int tsm_host_geometry(const tsm_mesh* mesh, int cell_type, int* fv, double* fn,
                                 int* bv, double* bn, double* dc, double* mc, int* nf, int* nb) {
  if (!mesh || !nf || !nb) return TSM_E_ARG;
  tsm::Geometry g;
  std::string msg;
  if (!tsm::build_geometry(*mesh, cell_type, g, msg)) return TSM_E_MESH;
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
