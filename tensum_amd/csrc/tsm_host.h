// tsm_host.h -- host-side data structures of libtsm (geometry, layout).
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/tsm.h"

namespace tsm {

struct Face {
  int v[2];
  int type;        // boundary tag (>= 0) or -1 for an interior face
  bool periodic;
  int lcell, rcell;
  double cx, cy;   // edge midpoint
  double nx, ny;   // dual-face normal (boundary faces: half dual face)
  double measure;
};

struct BFace {
  int v[2];
  int tag;
  int bc;          // TSM_BC_*
  bool periodic;
  int lcell;
  double nx, ny, measure;
};

// Dual-mesh geometry of one NPART = 1 partition, in reference order.
struct Geometry {
  int nv = 0, nc = 0;
  std::vector<double> x, y;
  std::vector<int> cell;            // 3 per cell
  std::vector<double> cell_area;
  std::vector<double> ccx, ccy;     // cell centroids
  std::vector<int> cell_slot_face;  // 3 per cell: face ids (reference slot order)
  std::vector<uint8_t> cell_has_nonper_bface;
  int removed_faces = 0;            // empty Voronoi dual faces dropped in preprocessing
  std::vector<Face> face;
  std::vector<BFace> bface;
  std::vector<double> mcarea, dcarea;
  std::vector<std::vector<int>> periodic;  // P_ij sets, member order of the file
  std::vector<int> periodic_of;            // set id per vertex or -1
  std::vector<double> weight;              // 1/|P| for periodic vertices else 1
};

// Median-dual geometry built on the device (tsm_geom.hip), bit-identical to
// build_geometry: 0 = built, 1 = left to the host builder, else TSM_E_*.
int build_geometry_device(const tsm_mesh& m, Geometry& g, std::string& msg);

// Vertex tiling for the LDS-staged kernels (see tiles.cpp).  Vertex ids in
// every array are the new (tile-contiguous) ids.
struct Tiling {
  int tb = 0, ntiles = 0, n_periodic_tiles = 0, max_slots = 0, max_faces = 0;
  std::vector<int> perm, inv;        // old -> new, new -> old
  std::vector<int> vbeg;             // ntiles+1: owned range
  std::vector<int> hbeg, halo;       // halo slots
  std::vector<int> fbeg;             // ntiles+1: local faces
  std::vector<uint32_t> f_lr;        // left slot | right slot << 16
  std::vector<double> f_n;           // 2 per local face
  std::vector<int> bbeg, tbf;        // tile boundary faces (global bface ids)
  std::vector<int> gptr;             // nv+1
  std::vector<uint16_t> gent;        // bit15 right end, bit14 bface, low 14 bits local index
  std::vector<int> cptr;             // nv+1
  std::vector<uint32_t> cent;        // 3 x 10-bit slots | correction flag << 30
  std::vector<int> cgid;             // global cell id per cell entry
  std::vector<int> tper;             // tile holds periodic sets
  std::vector<int> pinfo;            // per new vertex: first member slot << 8 | set size, or -1
  std::vector<int> gbv;              // outside periodic tiles: vertices with a boundary-corrected cell
  // Second halo ring (meshes without periodic sets): the stage kernel then
  // forms the gradients of the owned and first-ring slots itself.  Slots are
  // owned, first ring (halo), second ring (halo2); the gradient slots are the
  // first n1 = n_own + first-ring slots.
  bool fused = false;
  int max_slots1 = 0;                // largest n1
  std::vector<int> h2beg, halo2;     // second-ring slots
  std::vector<int> g1beg;            // ntiles+1: start of the tile's n1+1 entries in g1ptr
  std::vector<int> g1ptr;            // per gradient slot: cell entries [g1ptr[k], g1ptr[k+1])
  std::vector<uint32_t> g1ent;       // 3 x 10-bit slots | correction flag << 30
  std::vector<int> g1cgid;           // global cell id per entry
};

bool build_tiles(const Geometry& g, int tb, Tiling& T, std::string& msg, bool ring2 = false);

// The same tiling built on the device (tsm_tiling.hip) for meshes without
// the second halo ring: the arrays go straight into
// device memory from `dalloc` (pointers in TG, with n_gbv); T receives the
// numbering (perm, inv, vbeg) and the sizes.  0 = built, 1 = left to
// build_tiles, TSM_E_* with msg.
struct TileGeo;
int build_tiles_device(const Geometry& g, int tb, bool ring2, Tiling& T, TileGeo& TG,
                       const std::function<int(void**, size_t)>& dalloc, std::string& msg);
bool validate_tiles(const Geometry& g, const Tiling& T, std::string& msg);

// Builds faces, normals and areas exactly as Grid::preproc does for one
// partition (src/grid_preproc.cc:1147-1194).  Returns false with msg set.
bool build_geometry(const tsm_mesh& m, int cell_type, Geometry& g, std::string& msg);

}  // namespace tsm
