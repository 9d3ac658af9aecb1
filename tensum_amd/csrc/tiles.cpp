// tiles.cpp -- vertex tiling for the LDS-staged stage kernels.
//
// The device works on tiles: a tile owns up to `tb` vertices (consecutive new
// vertex ids) and sees a halo of every vertex adjacent to them.  Per tile the
// host prepares
//   * the halo list (new ids, ascending), so slot s < n_own is owned vertex
//     vbeg+s and slot s >= n_own is halo[hbeg + s - n_own];
//   * the local faces: every dual face with an owned endpoint, in reference
//     face order, as (left slot, right slot) + normal;
//   * per owned vertex, 16-bit gather entries into the local faces, in the
//     reference's face order (then its non-periodic boundary faces), which is
//     the summation order of the reference's face-loop scatter
//     (src/fv.cc:1071-1189);
//   * per owned vertex, its cells as three packed slot ids (cell vertex order)
//     for the Green-Gauss gradient (src/fv.cc:887-996).
// Periodic sets (src/reduce_residual_periodic.cc) are never split: their
// members are consecutive owned vertices of one tile, in set order, so the
// set sums become an in-tile exchange through LDS.  All other vertices are
// tiled in Morton order of their coordinates.
#include <algorithm>
#include <cstdint>
#include <unordered_map>

#include "tsm_host.h"

namespace tsm {

namespace {
inline uint32_t spread16(uint32_t v) {
  v &= 0xffff;
  v = (v | (v << 8)) & 0x00ff00ff;
  v = (v | (v << 4)) & 0x0f0f0f0f;
  v = (v | (v << 2)) & 0x33333333;
  v = (v | (v << 1)) & 0x55555555;
  return v;
}
}  // namespace

bool build_tiles(const Geometry& g, int tb, Tiling& T, std::string& msg, bool ring2) {
  const int nv = g.nv;
  T = Tiling{};
  T.tb = tb;
  T.fused = ring2 && g.periodic.empty();
  // ---- ownership and numbering -------------------------------------------
  std::vector<int> order;
  order.reserve(nv);
  std::vector<int> tile_start;  // in `order`
  // periodic tiles: whole sets, in set order.  Their members lie on opposite
  // sides of the domain, so each owned vertex brings its own halo; these
  // tiles own a quarter of a spatial tile to keep the largest slot count (and
  // with it the LDS size of every tile) close to that of a spatial tile.
  {
    const int cap = std::max(4, tb / 4);
    int cnt = 0;
    for (const auto& P : g.periodic) {
      if ((int)P.size() > cap) { msg = "periodic set larger than a tile"; return false; }
      if (cnt == 0 || cnt + (int)P.size() > cap) { tile_start.push_back((int)order.size()); cnt = 0; }
      for (int v : P) order.push_back(v);
      cnt += (int)P.size();
    }
  }
  T.n_periodic_tiles = (int)tile_start.size();
  {
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
    for (int i = 0; i < nv; ++i) {
      x0 = std::min(x0, g.x[i]); x1 = std::max(x1, g.x[i]);
      y0 = std::min(y0, g.y[i]); y1 = std::max(y1, g.y[i]);
    }
    double sx = (x1 > x0) ? 65535.0 / (x1 - x0) : 0.0, sy = (y1 > y0) ? 65535.0 / (y1 - y0) : 0.0;
    std::vector<std::pair<uint32_t, int>> key;
    key.reserve(nv);
    for (int i = 0; i < nv; ++i) {
      if (g.periodic_of[i] >= 0) continue;
      uint32_t qx = (uint32_t)((g.x[i] - x0) * sx + 0.5), qy = (uint32_t)((g.y[i] - y0) * sy + 0.5);
      key.push_back({spread16(qx) | (spread16(qy) << 1), i});
    }
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint32_t, int>& a, const std::pair<uint32_t, int>& b) {
                       return a.first < b.first;
                     });
    for (size_t k = 0; k < key.size(); ++k) {
      if (k % tb == 0) tile_start.push_back((int)order.size());
      order.push_back(key[k].second);
    }
  }
  if ((int)order.size() != nv) { msg = "tiling lost vertices"; return false; }
  const int nt = (int)tile_start.size();
  T.ntiles = nt;
  T.perm.assign(nv, -1);
  T.inv = order;
  for (int i = 0; i < nv; ++i) T.perm[order[i]] = i;
  T.vbeg.assign(tile_start.begin(), tile_start.end());
  T.vbeg.push_back(nv);
  std::vector<int> owner(nv);
  for (int t = 0; t < nt; ++t)
    for (int k = T.vbeg[t]; k < T.vbeg[t + 1]; ++k) owner[k] = t;   // new ids

  // ---- adjacency (new ids) -----------------------------------------------
  const int nf = (int)g.face.size();
  std::vector<std::vector<int>> vfaces(nv);  // faces per new vertex, ascending
  for (int f = 0; f < nf; ++f) {
    vfaces[T.perm[g.face[f].v[0]]].push_back(f);
    vfaces[T.perm[g.face[f].v[1]]].push_back(f);
  }
  std::vector<std::vector<int>> vcells(nv);
  for (int c = 0; c < g.nc; ++c)
    for (int j = 0; j < 3; ++j) vcells[T.perm[g.cell[3 * (size_t)c + j]]].push_back(c);
  std::vector<std::vector<int>> vbf(nv);
  for (int b = 0; b < (int)g.bface.size(); ++b) {
    if (g.bface[b].periodic) continue;
    vbf[T.perm[g.bface[b].v[0]]].push_back(b);
    vbf[T.perm[g.bface[b].v[1]]].push_back(b);
  }
  // cells with boundary corrections
  std::vector<uint8_t> ccorr(g.nc, 0);
  for (auto& b : g.bface)
    if (!b.periodic) ccorr[b.lcell] = 1;

  // ---- per tile ----------------------------------------------------------
  T.hbeg.assign(1, 0);
  T.fbeg.assign(1, 0);
  T.bbeg.assign(1, 0);
  T.gptr.assign(nv + 1, 0);
  T.cptr.assign(nv + 1, 0);
  T.tper.assign(nt, 0);
  T.pinfo.assign(nv, -1);
  std::vector<int> slot(nv, -1);
  for (int t = 0; t < nt; ++t) {
    const int b0 = T.vbeg[t], b1 = T.vbeg[t + 1], nown = b1 - b0;
    std::vector<int> slots;
    for (int v = b0; v < b1; ++v) { slot[v] = v - b0; slots.push_back(v); }
    auto add_slot = [&](int v) {
      if (slot[v] < 0) { slot[v] = (int)slots.size(); slots.push_back(v); T.halo.push_back(v); }
    };
    // halo: every cell corner of every owned vertex (covers face neighbours)
    for (int v = b0; v < b1; ++v)
      for (int c : vcells[v])
        for (int j = 0; j < 3; ++j) add_slot(T.perm[g.cell[3 * (size_t)c + j]]);
    for (int v = b0; v < b1; ++v)
      for (int f : vfaces[v]) { add_slot(T.perm[g.face[f].v[0]]); add_slot(T.perm[g.face[f].v[1]]); }
    // halo slots in ascending vertex id (the order the device tiling forms)
    std::sort(slots.begin() + nown, slots.end());
    std::sort(T.halo.begin() + T.hbeg.back(), T.halo.end());
    for (int k = nown; k < (int)slots.size(); ++k) slot[slots[k]] = k;
    T.hbeg.push_back((int)T.halo.size());
    const int n1 = (int)slots.size();
    if (T.fused) {
      // second ring: every cell corner of every first-ring vertex
      if (t == 0) { T.h2beg.assign(1, 0); T.g1beg.assign(1, 0); }
      for (int k = nown; k < n1; ++k)
        for (int c : vcells[slots[k]])
          for (int j = 0; j < 3; ++j) {
            const int v = T.perm[g.cell[3 * (size_t)c + j]];
            if (slot[v] < 0) { slot[v] = (int)slots.size(); slots.push_back(v); T.halo2.push_back(v); }
          }
      T.h2beg.push_back((int)T.halo2.size());
      T.max_slots1 = std::max(T.max_slots1, n1);
      for (int k = 0; k < n1; ++k) {
        T.g1ptr.push_back((int)T.g1ent.size());
        for (int c : vcells[slots[k]]) {
          uint32_t a = slot[T.perm[g.cell[3 * (size_t)c]]], bb = slot[T.perm[g.cell[3 * (size_t)c + 1]]],
                   d = slot[T.perm[g.cell[3 * (size_t)c + 2]]];
          T.g1ent.push_back(a | (bb << 10) | (d << 20) | ((uint32_t)ccorr[c] << 30));
          T.g1cgid.push_back(c);
        }
      }
      T.g1ptr.push_back((int)T.g1ent.size());
      T.g1beg.push_back((int)T.g1ptr.size());
    }
    if ((int)slots.size() > 1023) { msg = "tile exceeds 1023 slots"; return false; }
    T.max_slots = std::max(T.max_slots, (int)slots.size());
    // local faces in reference order
    std::vector<int> lf;
    for (int v = b0; v < b1; ++v)
      for (int f : vfaces[v]) lf.push_back(f);
    std::sort(lf.begin(), lf.end());
    lf.erase(std::unique(lf.begin(), lf.end()), lf.end());
    if (lf.size() >= (1u << 14)) { msg = "tile exceeds 16383 local faces"; return false; }
    T.max_faces = std::max(T.max_faces, (int)lf.size());
    std::unordered_map<int, int> lfi;
    for (int k = 0; k < (int)lf.size(); ++k) {
      int f = lf[k];
      lfi[f] = k;
      int l = slot[T.perm[g.face[f].v[0]]], r = slot[T.perm[g.face[f].v[1]]];
      T.f_lr.push_back((uint32_t)l | ((uint32_t)r << 16));
      T.f_n.push_back(g.face[f].nx);
      T.f_n.push_back(g.face[f].ny);
    }
    T.fbeg.push_back((int)T.f_lr.size());
    // boundary faces of the tile
    std::vector<int> lb;
    for (int v = b0; v < b1; ++v)
      for (int b : vbf[v]) lb.push_back(b);
    std::sort(lb.begin(), lb.end());
    lb.erase(std::unique(lb.begin(), lb.end()), lb.end());
    std::unordered_map<int, int> lbi;
    for (int k = 0; k < (int)lb.size(); ++k) { lbi[lb[k]] = k; T.tbf.push_back(lb[k]); }
    T.bbeg.push_back((int)T.tbf.size());
    // gather entries and cell entries per owned vertex
    for (int v = b0; v < b1; ++v) {
      int oldv = T.inv[v];
      for (int f : vfaces[v]) {
        uint16_t e = (uint16_t)lfi[f];
        if (g.face[f].v[1] == oldv) e |= 0x8000;   // right end: -F
        T.gent.push_back(e);
      }
      for (int b : vbf[v]) T.gent.push_back((uint16_t)(0x4000 | lbi[b]));
      T.gptr[v + 1] = (int)T.gent.size();
      for (int c : vcells[v]) {
        uint32_t a = slot[T.perm[g.cell[3 * (size_t)c]]], bb = slot[T.perm[g.cell[3 * (size_t)c + 1]]],
                 d = slot[T.perm[g.cell[3 * (size_t)c + 2]]];
        T.cent.push_back(a | (bb << 10) | (d << 20) | ((uint32_t)ccorr[c] << 30));
        T.cgid.push_back(c);
      }
      T.cptr[v + 1] = (int)T.cent.size();
      if (t >= T.n_periodic_tiles) {
        bool corr = false;
        for (int c : vcells[v]) corr = corr || ccorr[c];
        if (corr) T.gbv.push_back(v);
      }
    }
    // periodic sets of this tile: consecutive owned slots
    for (int v = b0; v < b1; ++v) {
      int p = g.periodic_of[T.inv[v]];
      if (p < 0) continue;
      T.tper[t] = 1;
      const auto& P = g.periodic[p];
      int first = slot[T.perm[P[0]]];
      for (size_t k = 0; k < P.size(); ++k)
        if (slot[T.perm[P[k]]] != first + (int)k) { msg = "periodic set not contiguous"; return false; }
      T.pinfo[v] = (first << 8) | (int)P.size();
    }
    for (int v : slots) slot[v] = -1;
    (void)nown;
  }
  return true;
}

}  // namespace tsm

namespace tsm {

// Checks every index of a tiling against the bounds the kernels assume.
bool validate_tiles(const Geometry& g, const Tiling& T, std::string& msg) {
  const int nv = g.nv;
  if ((int)T.vbeg.size() != T.ntiles + 1 || T.vbeg.back() != nv) { msg = "vbeg"; return false; }
  for (int t = 0; t < T.ntiles; ++t) {
    const int nown = T.vbeg[t + 1] - T.vbeg[t];
    if (nown <= 0 || nown > T.tb) { msg = "tile size"; return false; }
    const int ns = nown + T.hbeg[t + 1] - T.hbeg[t];
    if (ns > T.max_slots || ns > 1023) { msg = "slot count"; return false; }
    for (int h = T.hbeg[t]; h < T.hbeg[t + 1]; ++h)
      if (T.halo[h] < 0 || T.halo[h] >= nv) { msg = "halo id"; return false; }
    const int nfl = T.fbeg[t + 1] - T.fbeg[t];
    if (nfl > T.max_faces) { msg = "face count"; return false; }
    for (int j = T.fbeg[t]; j < T.fbeg[t + 1]; ++j) {
      const int l = T.f_lr[j] & 0xffff, r = T.f_lr[j] >> 16;
      if (l >= ns || r >= ns) { msg = "face slot"; return false; }
    }
    const int nbt = T.bbeg[t + 1] - T.bbeg[t];
    for (int j = T.bbeg[t]; j < T.bbeg[t + 1]; ++j)
      if (T.tbf[j] < 0 || T.tbf[j] >= (int)g.bface.size()) { msg = "bface id"; return false; }
    for (int v = T.vbeg[t]; v < T.vbeg[t + 1]; ++v) {
      bool seen_bface = false;  // the kernels take a vertex's faces before its boundary faces
      for (int k = T.gptr[v]; k < T.gptr[v + 1]; ++k) {
        const unsigned e = T.gent[k];
        if (e & 0x4000u) {
          if ((int)(e & 0x3fffu) >= nbt) { msg = "gather bface"; return false; }
          seen_bface = true;
        } else {
          if ((int)(e & 0x3fffu) >= nfl) { msg = "gather face"; return false; }
          if (seen_bface) { msg = "gather face after a boundary face"; return false; }
        }
      }
      for (int k = T.cptr[v]; k < T.cptr[v + 1]; ++k) {
        const unsigned e = T.cent[k];
        if ((int)(e & 1023) >= ns || (int)((e >> 10) & 1023) >= ns || (int)((e >> 20) & 1023) >= ns) {
          msg = "cell slot"; return false;
        }
        if (T.cgid[k] < 0 || T.cgid[k] >= g.nc) { msg = "cell id"; return false; }
      }
      if (T.pinfo[v] >= 0) {
        const int first = T.pinfo[v] >> 8, cnt = T.pinfo[v] & 255;
        if (!T.tper[t] || first + cnt > nown) { msg = "periodic slots"; return false; }
      }
    }
    if (T.fused) {
      const int n1 = ns;  // owned + first ring
      const int ns2 = n1 + T.h2beg[t + 1] - T.h2beg[t];
      if (n1 > T.max_slots1 || ns2 > T.max_slots || ns2 > 1023) { msg = "ring-2 slot count"; return false; }
      for (int h = T.h2beg[t]; h < T.h2beg[t + 1]; ++h)
        if (T.halo2[h] < 0 || T.halo2[h] >= nv) { msg = "ring-2 id"; return false; }
      if (T.g1beg[t + 1] - T.g1beg[t] != n1 + 1) { msg = "gradient slot pointers"; return false; }
      for (int k = T.g1beg[t]; k < T.g1beg[t + 1] - 1; ++k) {
        if (T.g1ptr[k] > T.g1ptr[k + 1] || T.g1ptr[k + 1] > (int)T.g1ent.size()) {
          msg = "gradient cell range"; return false;
        }
        for (int e = T.g1ptr[k]; e < T.g1ptr[k + 1]; ++e) {
          const unsigned c = T.g1ent[e];
          if ((int)(c & 1023) >= ns2 || (int)((c >> 10) & 1023) >= ns2 || (int)((c >> 20) & 1023) >= ns2) {
            msg = "gradient cell slot"; return false;
          }
          if (T.g1cgid[e] < 0 || T.g1cgid[e] >= g.nc) { msg = "gradient cell id"; return false; }
        }
      }
    }
  }
  for (int v : T.gbv)
    if (v < 0 || v >= nv || v < T.vbeg[T.n_periodic_tiles]) { msg = "boundary gradient vertex"; return false; }
  if (T.fused && ((int)T.h2beg.size() != T.ntiles + 1 || (int)T.g1beg.size() != T.ntiles + 1)) {
    msg = "ring-2 tables"; return false;
  }
  return true;
}

}  // namespace tsm

extern "C" int tsm_host_tiling(const tsm_mesh* mesh, int cell_type, int tile_vertices, int two_rings,
                               int* stats) {
  if (!mesh || !stats) return TSM_E_ARG;
  tsm::Geometry g;
  std::string msg;
  if (!tsm::build_geometry(*mesh, cell_type, g, msg)) return TSM_E_MESH;
  tsm::Tiling T;
  if (!tsm::build_tiles(g, tile_vertices, T, msg, two_rings != 0)) return TSM_E_MESH;
  if (!tsm::validate_tiles(g, T, msg)) return TSM_E_MESH;
  stats[0] = T.ntiles;
  stats[1] = T.n_periodic_tiles;
  stats[2] = T.max_slots;
  stats[3] = T.max_faces;
  stats[4] = (int)T.halo.size();
  stats[5] = (int)T.f_lr.size();
  stats[6] = T.fused ? 1 : 0;
  stats[7] = (int)T.halo2.size();
  stats[8] = (int)T.g1ent.size();
  stats[9] = T.max_slots1;
  return TSM_OK;
}
