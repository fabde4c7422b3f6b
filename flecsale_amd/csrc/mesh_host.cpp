// Host-side connectivity builder: the part of the burton mesh
// (flecsi-sp, un-vendored in the reference) that the hydro tasks consume:
// mesh.faces(c), mesh.cells(f), mesh.cell_vertex_neighbors(c),
// mesh.corners(v) (apps/hydro/tasks.h:190,355,414,629;
// apps/maire_hydro/tasks.h:256,480).  Linear time: every face is matched
// through the cells around its first vertex.
#include <algorithm>
#include <array>

#include "internal.h"

namespace fsale {

namespace {
// Exodus hex8 sides, outward for a positively oriented hex.
constexpr int kHexFaces[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                                 {0, 4, 7, 3}, {0, 3, 2, 1}, {4, 5, 6, 7}};

inline void local_face(int dim, int vpc, int lf, int *lv) {
  if (dim == 2) {
    lv[0] = lf;
    lv[1] = (lf + 1) % vpc;
  } else {
    for (int i = 0; i < 4; ++i) lv[i] = kHexFaces[lf][i];
  }
}

inline std::array<int32_t, 4> face_key(const HostTopology &t, int c, int lf) {
  int lv[4];
  local_face(t.dim, t.vpc, lf, lv);
  std::array<int32_t, 4> k{INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX};
  for (int i = 0; i < t.vpf; ++i) k[i] = t.cv[(size_t)c * t.vpc + lv[i]];
  std::sort(k.begin(), k.begin() + t.vpf);
  return k;
}
}  // namespace

int build_topology(const fsale_mesh_desc &mesh, HostTopology &t, std::string &err) {
  if (!mesh.coords || !mesh.cells || mesh.num_cells <= 0 || mesh.num_vertices <= 0) {
    err = "empty mesh";
    return FSALE_ERR_ARG;
  }
  if (!((mesh.dim == 2 && mesh.verts_per_cell >= 3 && mesh.verts_per_cell <= 8) ||
        (mesh.dim == 3 && mesh.verts_per_cell == 8))) {
    err = "unsupported cell type (quads/polygons in 2D, hexes in 3D)";
    return FSALE_ERR_MESH;
  }
  t.dim = mesh.dim;
  t.nc = mesh.num_cells;
  t.nv = mesh.num_vertices;
  t.vpc = mesh.verts_per_cell;
  t.fpc = (t.dim == 2) ? t.vpc : 6;
  t.vpf = (t.dim == 2) ? 2 : 4;
  const size_t nslots = (size_t)t.nc * t.vpc;
  t.cv.assign(mesh.cells, mesh.cells + nslots);
  for (size_t i = 0; i < nslots; ++i) {
    if (t.cv[i] < 0 || t.cv[i] >= t.nv) {
      err = "cell vertex id out of range";
      return FSALE_ERR_MESH;
    }
  }

  // vertex -> corner (c*vpc+k), ascending cell id by construction
  t.vc_off.assign(t.nv + 1, 0);
  for (size_t i = 0; i < nslots; ++i) t.vc_off[t.cv[i] + 1]++;
  for (int v = 0; v < t.nv; ++v) t.vc_off[v + 1] += t.vc_off[v];
  t.vc.resize(nslots);
  {
    std::vector<int32_t> fill(t.vc_off.begin(), t.vc_off.end() - 1);
    for (size_t i = 0; i < nslots; ++i) t.vc[fill[t.cv[i]]++] = (int32_t)i;
  }

  // faces by first visit
  t.cf.assign((size_t)t.nc * t.fpc, -1);
  t.fc.clear();
  t.fv.clear();
  t.fc.reserve((size_t)t.nc * t.fpc);
  t.fv.reserve((size_t)t.nc * t.fpc * t.vpf / 2 + 16);
  int nf = 0;
  for (int c = 0; c < t.nc; ++c) {
    for (int lf = 0; lf < t.fpc; ++lf) {
      const auto key = face_key(t, c, lf);
      int other = -1, other_lf = -1, matches = 0;
      const int v0 = key[0];
      for (int q = t.vc_off[v0]; q < t.vc_off[v0 + 1]; ++q) {
        const int c2 = t.vc[q] / t.vpc;
        if (c2 == c) continue;
        for (int lf2 = 0; lf2 < t.fpc; ++lf2) {
          if (face_key(t, c2, lf2) == key) {
            if (other != c2) ++matches;
            other = c2;
            other_lf = lf2;
            break;
          }
        }
      }
      if (matches > 1) {
        err = "non-manifold face";
        return FSALE_ERR_MESH;
      }
      if (other >= 0 && other < c) {
        const int f = t.cf[(size_t)other * t.fpc + other_lf];
        if (f < 0 || t.fc[2 * (size_t)f + 1] != -1) {
          err = "inconsistent face matching";
          return FSALE_ERR_MESH;
        }
        // 2D cells must share an edge in opposite directions (consistent
        // orientation): the fused flux kernel rebuilds face geometry from the
        // cell's own vertices and relies on it
        if (t.dim == 2 && !(t.cv[(size_t)c * t.vpc + lf] == t.fv[2 * (size_t)f + 1] &&
                            t.cv[(size_t)c * t.vpc + (lf + 1) % t.vpc] == t.fv[2 * (size_t)f])) {
          err = "inconsistently oriented cells (2D cells must all be counter-clockwise)";
          return FSALE_ERR_MESH;
        }
        t.fc[2 * (size_t)f + 1] = c;
        t.cf[(size_t)c * t.fpc + lf] = f;
      } else {
        const int f = nf++;
        t.fc.push_back(c);
        t.fc.push_back(-1);
        int lv[4];
        local_face(t.dim, t.vpc, lf, lv);
        for (int i = 0; i < t.vpf; ++i) t.fv.push_back(t.cv[(size_t)c * t.vpc + lv[i]]);
        t.cf[(size_t)c * t.fpc + lf] = f;
      }
    }
  }
  t.nf = nf;

  // cell-vertex neighbours
  t.nbr_off.assign(t.nc + 1, 0);
  t.nbr.clear();
  t.nbr.reserve((size_t)t.nc * (t.dim == 2 ? 8 : 26));
  std::vector<int32_t> tmp;
  for (int c = 0; c < t.nc; ++c) {
    tmp.clear();
    for (int k = 0; k < t.vpc; ++k) {
      const int v = t.cv[(size_t)c * t.vpc + k];
      for (int q = t.vc_off[v]; q < t.vc_off[v + 1]; ++q) {
        const int c2 = t.vc[q] / t.vpc;
        if (c2 != c) tmp.push_back(c2);
      }
    }
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    t.nbr.insert(t.nbr.end(), tmp.begin(), tmp.end());
    t.nbr_off[c + 1] = (int32_t)t.nbr.size();
  }
  return FSALE_OK;
}

}  // namespace fsale
