// Host-side connectivity builder: the part of the burton mesh
// (flecsi-sp, un-vendored in the reference) that the hydro tasks consume:
// mesh.faces(c), mesh.cells(f), mesh.cell_vertex_neighbors(c),
// mesh.corners(v) (apps/hydro/tasks.h:190,355,414,629;
// apps/maire_hydro/tasks.h:256,480).  Linear time: every face is matched
// through the cells around its first vertex.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <unordered_map>

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

// local faces holding local vertex kk: 2D edges (kk-1 -> kk) and (kk -> kk+1);
// hex: the three sides of kHexFaces that list kk
constexpr int kHexVertexFaces[8][3] = {{0, 3, 4}, {0, 1, 4}, {1, 2, 4}, {2, 3, 4},
                                       {0, 3, 5}, {0, 1, 5}, {1, 2, 5}, {2, 3, 5}};

inline int vertex_faces(const HostTopology &t, int kk, int j) {
  if (t.dim == 2) return j == 0 ? kk : (kk + t.vpc - 1) % t.vpc;
  return kHexVertexFaces[kk][j];
}

inline std::array<int32_t, 4> face_key(const HostTopology &t, int c, int lf) {
  const int32_t *cv = &t.cv[(size_t)c * t.vpc];
  if (t.dim == 2) {
    const int32_t a = cv[lf], b = cv[lf + 1 == t.vpc ? 0 : lf + 1];
    return {std::min(a, b), std::max(a, b), INT32_MAX, INT32_MAX};
  }
  int lv[4];
  local_face(t.dim, t.vpc, lf, lv);
  std::array<int32_t, 4> k{cv[lv[0]], cv[lv[1]], cv[lv[2]], cv[lv[3]]};
  // sorting network for 4 keys
  auto cs = [&](int i, int j) {
    if (k[j] < k[i]) std::swap(k[i], k[j]);
  };
  cs(0, 1);
  cs(2, 3);
  cs(0, 2);
  cs(1, 3);
  cs(1, 2);
  return k;
}
}  // namespace

namespace {
// Spread the low 21 bits of x over every third bit (3D) or the low 31 bits
// over every second bit (2D).
inline uint64_t spread3(uint64_t x) {
  x &= 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffull;
  x = (x | x << 16) & 0x1f0000ff0000ffull;
  x = (x | x << 8) & 0x100f00f00f00f00full;
  x = (x | x << 4) & 0x10c30c30c30c30c3ull;
  x = (x | x << 2) & 0x1249249249249249ull;
  return x;
}
inline uint64_t spread2(uint64_t x) {
  x &= 0x7fffffff;
  x = (x | x << 16) & 0x0000ffff0000ffffull;
  x = (x | x << 8) & 0x00ff00ff00ff00ffull;
  x = (x | x << 4) & 0x0f0f0f0f0f0f0f0full;
  x = (x | x << 2) & 0x3333333333333333ull;
  x = (x | x << 1) & 0x5555555555555555ull;
  return x;
}

// Storage order mode, FSALE_STORAGE_ORDER = identity | morton | blocks.  By
// default: blocks for 2D meshes of more than 512^2 cells, else identity.
// Blocks of ~256 cells make a large 2D mesh's Eulerian cell tiles compact,
// so they are staged in LDS with small halos (1024^2 four-quadrant MUSCL:
// 4.49e9 vs 3.75e9 zone-updates/s).  Smaller meshes keep their own order:
// their blocks rarely align with the 256-cell tiles, and the served stream
// of 32^2..512^2 meshes ran at 2.48e9 zone-updates/s in identity order and
// 2.07e9 in blocks.  3D: the given order coalesces the gathers best on
// meshes numbered with locality (structured, Exodus).
int storage_order_mode(int dim, int num_cells) {
  static const int mode = [] {
    const char *e = std::getenv("FSALE_STORAGE_ORDER");
    if (!e) return -1;
    if (!std::strcmp(e, "identity")) return 0;
    if (!std::strcmp(e, "morton")) return 1;
    if (!std::strcmp(e, "blocks")) return 2;
    return -1;
  }();
  return mode >= 0 ? mode : (dim == 2 && num_cells > 512 * 512 ? 2 : 0);
}

// Order of n points: identity; Z-order (Morton) of the points; or Z-order of
// spatial blocks of ~256 points with the original order inside a block.
// idx[s] = the point at position s, pos its inverse; ties keep id order.
// box (optional, 2*dim values lo then hi): the region the blocks divide,
// e.g. the mesh's extent for cell centroids, so that a structured mesh's
// blocks hold whole 16 x 16 (x 16 ...) cell groups instead of being cut at
// the centroids' own, half a cell smaller, extent.
void morton_order(int n, int dim, const std::vector<double> &pts, std::vector<int32_t> &idx,
                  std::vector<int32_t> &pos, int mode, const double *box = nullptr) {
  if (mode == 0) {
    idx.resize(n);
    pos.resize(n);
    for (int i = 0; i < n; ++i) idx[i] = pos[i] = i;
    return;
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < dim; ++d) {
      lo[d] = std::min(lo[d], pts[(size_t)i * dim + d]);
      hi[d] = std::max(hi[d], pts[(size_t)i * dim + d]);
    }
  if (box && mode == 2)
    for (int d = 0; d < dim; ++d) {
      lo[d] = std::min(lo[d], box[d]);
      hi[d] = std::max(hi[d], box[dim + d]);
    }
  double levels[3];
  for (int d = 0; d < dim; ++d) levels[d] = dim == 3 ? 2097151.0 : 2147483647.0;
  if (mode == 2) {
    // blocks of ~256 points: edge h with (prod extents / h^dim) = n / 256
    double vol = 1;
    int live = 0;
    for (int d = 0; d < dim; ++d)
      if (hi[d] > lo[d]) {
        vol *= hi[d] - lo[d];
        ++live;
      }
    const double nb = std::max(1.0, n / 256.0);
    const double h = live ? std::pow(vol / nb, 1.0 / live) : 1.0;
    for (int d = 0; d < dim; ++d) levels[d] = std::max(0.0, std::floor((hi[d] - lo[d]) / h));
  }
  std::vector<std::pair<uint64_t, int32_t>> key(n);
  for (int i = 0; i < n; ++i) {
    uint64_t code = 0;
    for (int d = 0; d < dim; ++d) {
      const double span = hi[d] - lo[d];
      const double u = span > 0 ? (pts[(size_t)i * dim + d] - lo[d]) / span : 0.0;
      const double q0 = std::min(std::max(u, 0.0), 1.0) * levels[d];
      const uint64_t q = (uint64_t)(mode == 2 ? std::floor(q0 - 1e-9 * levels[d]) : std::llround(q0));
      code |= (dim == 3 ? spread3(q) : spread2(q)) << d;
    }
    key[i] = {code, i};
  }
  std::sort(key.begin(), key.end());
  idx.resize(n);
  pos.resize(n);
  for (int s = 0; s < n; ++s) {
    idx[s] = key[s].second;
    pos[key[s].second] = s;
  }
}
}  // namespace

// LDS staging plan of a 2D quad mesh's cell tiles (HostTopology::halo_cnt,
// halo), in storage positions.
void build_halo_plan(HostTopology &t) {
  const int T = (t.nc + kBlock - 1) / kBlock;
  t.halo_cnt.assign(T, -1);
  t.halo.assign((size_t)T * kHalo2, 0);
  std::vector<int32_t> h;
  for (int j = 0; j < T; ++j) {
    const int s0 = j * kBlock, s1 = std::min(t.nc, s0 + kBlock);
    h.clear();
    bool ok = true;
    for (int s = s0; s < s1 && ok; ++s) {
      const int c = t.cidx[s];
      if (t.nbr_off[c + 1] - t.nbr_off[c] > kNbk2) ok = false;
      for (int q = t.nbr_off[c]; q < t.nbr_off[c + 1]; ++q) {
        const int p = t.cpos[t.nbr[q]];
        if (p < s0 || p >= s1) h.push_back(p);
      }
    }
    std::sort(h.begin(), h.end());
    h.erase(std::unique(h.begin(), h.end()), h.end());
    if (!ok || (int)h.size() > kHalo2) continue;
    t.halo_cnt[j] = (int32_t)h.size();
    std::copy(h.begin(), h.end(), t.halo.begin() + (size_t)j * kHalo2);
  }
}

// The bricks of a structured grid (HostTopology::brick), dealt to the tiles
// in dispatch order (tile_order, or storage order).
void brick_plan(HostTopology &t, int nx, int ny, int nz) {
  t.brick.clear();
  if (nx % kBrickX || ny % kBrickY || nz % kBrickZ) return;
  const int BX = nx / kBrickX, BY = ny / kBrickY, BZ = nz / kBrickZ;
  const int T = BX * BY * BZ;
  if (T * kBlock != t.nc) return;
  // sweep: XCD part (a slab of brick rows along y), then z, then y, then x
  std::vector<int64_t> key(T);
  std::vector<int32_t> ord(T);
  for (int q = 0; q < T; ++q) {
    const int bx = q % BX, by = (q / BX) % BY, bz = q / (BX * BY);
    const int part = (int)((int64_t)by * 8 / BY);
    key[q] = (((int64_t)part * BZ + bz) * BY + by) * BX + bx;
    ord[q] = q;
  }
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
  t.brick.assign(T, 0);
  for (int i = 0; i < T; ++i) {
    const int j = t.tile_order.empty() ? i : t.tile_order[i];
    const int q = ord[i];
    const int bx = q % BX, by = (q / BX) % BY, bz = q / (BX * BY);
    t.brick[j] = bx * kBrickX + by * kBrickY * nx + bz * kBrickZ * nx * ny;
  }
}

// A 3D hex mesh whose cells form a structured nx x ny x nz grid numbered x
// fastest (HostTopology::s_nx, s_np): cell 0's neighbours name nx and nx*ny,
// and every cell's ascending neighbour list must be the grid's (dz, dy, dx)
// lexicographic list.  Off with FSALE_STRUCTURED=0, or when a tile's plane
// range would not fit the staging (nx > (kLds3 - kBlock - 2) / 2).
void detect_structured(HostTopology &t) {
  t.s_nx = t.s_np = t.s_fp = 0;
  static const bool on = [] {
    const char *e = std::getenv("FSALE_STRUCTURED");
    return !(e && !std::strcmp(e, "0"));
  }();
  if (!on || t.dim != 3 || t.nc < 8) return;
  for (int s = 0; s < t.nc; ++s)
    if (t.cidx[s] != s) return;   // stored in another order
  if (t.nbr_off[1] - t.nbr_off[0] != 7) return;
  const int *n0 = t.nbr.data() + t.nbr_off[0];
  const int nx = n0[1], np = n0[3];
  if (nx < 2 || np < 2 * nx || np % nx || t.nc % np || t.nc / np < 2) return;
  if (2 * nx + 2 + kBlock > kLds3) return;
  const int ny = np / nx, nz = t.nc / np;
  int want[26];
  for (int c = 0; c < t.nc; ++c) {
    const int i = c % nx, j = (c / nx) % ny, k = c / np;
    int n = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if ((dz | dy | dx) == 0) continue;
          if (i + dx < 0 || i + dx >= nx || j + dy < 0 || j + dy >= ny || k + dz < 0 || k + dz >= nz) continue;
          want[n++] = c + dz * np + dy * nx + dx;
        }
    if (t.nbr_off[c + 1] - t.nbr_off[c] != n) return;
    if (!std::equal(want, want + n, t.nbr.data() + t.nbr_off[c])) return;
  }
  t.s_nx = nx;
  t.s_np = np;
  // the -x, -y, -z sides of a cell against the +x, +y, +z sides of the
  // neighbour across them (local faces of kHexFaces: 3/1, 0/2, 4/5)
  static constexpr int kLow[3] = {3, 0, 4}, kHigh[3] = {1, 2, 5};
  const int stride[3] = {1, nx, np};
  for (int c = 0; c < t.nc; ++c) {
    const int at[3] = {c % nx, (c / nx) % ny, c / np};
    for (int a = 0; a < 3; ++a) {
      if (at[a] == 0) continue;
      const int n = c - stride[a];
      if (t.cf[(size_t)c * 6 + kLow[a]] != t.cf[(size_t)n * 6 + kHigh[a]]) return;
    }
  }
  t.s_fp = 1;
  brick_plan(t, nx, ny, nz);
}

// Pencil order of a 3D mesh's cell tiles (HostTopology::tile_order) from the
// tiles' centroids; cellsum holds each cell's vertex-coordinate sum.
void tile_pencils(HostTopology &t, const std::vector<double> &cellsum) {
  const int T = (t.nc + kBlock - 1) / kBlock;
  t.tile_order.clear();
  if (T < 64) return;
  std::vector<double> tc((size_t)T * 3, 0.0);
  for (int s = 0; s < t.nc; ++s) {
    const int c = t.cidx.empty() ? s : t.cidx[s];
    for (int d = 0; d < 3; ++d) tc[(size_t)(s / kBlock) * 3 + d] += cellsum[(size_t)c * 3 + d];
  }
  for (int j = 0; j < T; ++j) {
    const double n = (double)std::min(kBlock, t.nc - j * kBlock);
    for (int d = 0; d < 3; ++d) tc[(size_t)j * 3 + d] /= n;
  }
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) lo[d] = hi[d] = tc[d];
  for (int j = 1; j < T; ++j)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], tc[(size_t)j * 3 + d]);
      hi[d] = std::max(hi[d], tc[(size_t)j * 3 + d]);
    }
  int ax[3] = {0, 1, 2};
  std::sort(ax, ax + 3, [&](int a, int b) { return hi[a] - lo[a] > hi[b] - lo[b]; });
#ifdef FSALE_TORD_EXP
  if (std::getenv("FSALE_NO_PENCIL")) return;
  const bool swp = std::getenv("FSALE_PENCIL_SWAP") != nullptr;
  const int sweep = swp ? ax[1] : ax[0], slab = swp ? ax[0] : ax[1];
#else
  const int sweep = ax[0], slab = ax[1];
#endif
  std::vector<int32_t> &ord = t.tile_order;
  ord.resize(T);
  for (int j = 0; j < T; ++j) ord[j] = j;
  auto key = [&](int j, int d) { return tc[(size_t)j * 3 + d]; };
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key(a, slab) < key(b, slab); });
  for (int k = 0; k < 8; ++k) {
    const auto b = ord.begin() + (size_t)T * k / 8, e = ord.begin() + (size_t)T * (k + 1) / 8;
    std::stable_sort(b, e, [&](int x, int y) {
      return key(x, sweep) < key(y, sweep) || (key(x, sweep) == key(y, sweep) && key(x, slab) < key(y, slab));
    });
  }
}

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
  for (int v = 0; v < t.nv; ++v) {
    t.max_corners = std::max(t.max_corners, t.vc_off[v + 1]);
    t.vc_off[v + 1] += t.vc_off[v];
  }
  t.vc.resize(nslots);
  {
    std::vector<int32_t> fill(t.vc_off.begin(), t.vc_off.end() - 1);
    for (size_t i = 0; i < nslots; ++i) t.vc[fill[t.cv[i]]++] = (int32_t)i;
  }
  if (t.max_corners > kMaxCornersPerVertex) {
    err = "more than 8 cells around a vertex";
    return FSALE_ERR_MESH;
  }
  // cell and local vertex of each vertex-corner entry
  std::vector<int32_t> vcell(nslots), vlocal(nslots);
  for (size_t q = 0; q < nslots; ++q) {
    vcell[q] = t.vc[q] / t.vpc;
    vlocal[q] = t.vc[q] - vcell[q] * t.vpc;
  }

  // faces by first visit; each face is matched among the faces (with the
  // same sorted vertex key) of the cells around its lowest vertex
  const size_t nfk = (size_t)t.nc * t.fpc;
  std::vector<std::array<int32_t, 4>> keys(nfk);
  for (int c = 0; c < t.nc; ++c)
    for (int lf = 0; lf < t.fpc; ++lf) keys[(size_t)c * t.fpc + lf] = face_key(t, c, lf);
  t.cf.assign(nfk, -1);
  t.fc.clear();
  t.fv.clear();
  t.fc.reserve(nfk);
  t.fv.reserve(nfk * t.vpf / 2 + 16);
  int nf = 0;
  const int nvf = t.dim == 2 ? 2 : 3;   // faces of a cell holding one of its vertices
  for (int c = 0; c < t.nc; ++c) {
    for (int lf = 0; lf < t.fpc; ++lf) {
      const auto &key = keys[(size_t)c * t.fpc + lf];
      int other = -1, other_lf = -1, matches = 0;
      const int v0 = key[0];
      for (int q = t.vc_off[v0]; q < t.vc_off[v0 + 1]; ++q) {
        const int c2 = vcell[q], kk = vlocal[q];
        if (c2 == c) continue;
        // only the faces of c2 that hold its local vertex kk can match
        const auto *k2 = &keys[(size_t)c2 * t.fpc];
        for (int j = 0; j < nvf; ++j) {
          const int lf2 = vertex_faces(t, kk, j);
          if (k2[lf2] == key) {
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
  // gather the cells around the cell's vertices, then an insertion sort and
  // duplicate removal (a few dozen entries)
  int32_t tmp[8 * kMaxCornersPerVertex + 8];
  for (int c = 0; c < t.nc; ++c) {
    int n = 0;
    for (int k = 0; k < t.vpc; ++k) {
      const int v = t.cv[(size_t)c * t.vpc + k];
      for (int q = t.vc_off[v]; q < t.vc_off[v + 1]; ++q) {
        const int c2 = vcell[q];
        if (c2 != c && n < (int)(sizeof(tmp) / sizeof(tmp[0]))) tmp[n++] = c2;
      }
    }
    for (int i = 1; i < n; ++i) {
      const int32_t x = tmp[i];
      int j = i - 1;
      while (j >= 0 && tmp[j] > x) {
        tmp[j + 1] = tmp[j];
        --j;
      }
      tmp[j + 1] = x;
    }
    int count = 0;
    for (int i = 0; i < n; ++i)
      if (i == 0 || tmp[i] != tmp[i - 1]) {
        t.nbr.push_back(tmp[i]);
        ++count;
      }
    t.nbr_off[c + 1] = (int32_t)t.nbr.size();
    t.max_neighbours = std::max(t.max_neighbours, count);
  }

  // device storage order of cells, vertices and faces (vertex averages)
  {
    const int D = t.dim;
    // the mesh's extent, scaled like the vertex sums the cells and faces are
    // placed by
    double ext[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (int v = 0; v < t.nv; ++v)
      for (int d = 0; d < D; ++d) {
        ext[d] = std::min(ext[d], mesh.coords[(size_t)v * D + d]);
        ext[3 + d] = std::max(ext[3 + d], mesh.coords[(size_t)v * D + d]);
      }
    double cbox[6], fbox[6];
    for (int d = 0; d < D; ++d) {
      cbox[d] = t.vpc * ext[d];
      cbox[D + d] = t.vpc * ext[3 + d];
      fbox[d] = t.vpf * ext[d];
      fbox[D + d] = t.vpf * ext[3 + d];
    }
    std::vector<double> pts((size_t)t.nc * D, 0.0);
    for (int c = 0; c < t.nc; ++c)
      for (int k = 0; k < t.vpc; ++k)
        for (int d = 0; d < D; ++d) pts[(size_t)c * D + d] += mesh.coords[(size_t)t.cv[(size_t)c * t.vpc + k] * D + d];
    const int mode = storage_order_mode(D, t.nc);
    morton_order(t.nc, D, pts, t.cidx, t.cpos, mode, cbox);
    if (D == 3) tile_pencils(t, pts);
    pts.assign(mesh.coords, mesh.coords + (size_t)t.nv * D);
    morton_order(t.nv, D, pts, t.vidx, t.vpos, mode);
    pts.assign((size_t)t.nf * D, 0.0);
    for (int f = 0; f < t.nf; ++f)
      for (int k = 0; k < t.vpf; ++k)
        for (int d = 0; d < D; ++d) pts[(size_t)f * D + d] += mesh.coords[(size_t)t.fv[(size_t)f * t.vpf + k] * D + d];
    morton_order(t.nf, D, pts, t.fidx, t.fpos, mode, fbox);
  }
  if (t.dim == 2 && t.vpc == 4) build_halo_plan(t);
  detect_structured(t);
  return FSALE_OK;
}

// Process-wide connectivity cache: meshes are identified by their cell table
// (a hash to find candidates, the table itself to confirm), so every context
// and batch that sees the same mesh again skips the build.  Oldest entries
// are dropped past the budget.
namespace {
// budget: FSALE_TOPOLOGY_CACHE_MB in the environment, default 8 GiB
size_t topology_cache_budget() {
  static const size_t b = [] {
    const char *e = std::getenv("FSALE_TOPOLOGY_CACHE_MB");
    const long long mb = e ? std::atoll(e) : 8192;
    return (size_t)(mb > 0 ? mb : 0) << 20;
  }();
  return b;
}

struct TopologyCache {
  std::mutex mu;
  std::unordered_multimap<uint64_t, std::shared_ptr<const HostTopology>> map;
  std::deque<std::pair<uint64_t, const HostTopology *>> order;
  size_t bytes = 0;
};

TopologyCache &topology_cache() {
  static TopologyCache cache;
  return cache;
}

size_t topology_bytes(const HostTopology &t) {
  return sizeof(int32_t) * (t.cv.size() + t.cf.size() + t.fc.size() + t.fv.size() + t.nbr_off.size() +
                            t.nbr.size() + t.vc_off.size() + t.vc.size() + t.tile_order.size() +
                            2 * (t.cidx.size() + t.vidx.size() +
                            t.fidx.size()) + t.halo_cnt.size() + t.halo.size());
}

uint64_t mesh_hash(const fsale_mesh_desc &m) {
  uint64_t h = 0x9E3779B97F4A7C15ull ^ ((uint64_t)m.dim << 48) ^ ((uint64_t)m.verts_per_cell << 40) ^
               ((uint64_t)(uint32_t)m.num_vertices << 8) ^ (uint64_t)(uint32_t)m.num_cells;
  const size_t n = (size_t)m.num_cells * m.verts_per_cell;
  for (size_t i = 0; i < n; ++i) {
    h ^= (uint64_t)(uint32_t)m.cells[i] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  }
  return h;
}
}  // namespace

std::shared_ptr<const HostTopology> cached_topology(const fsale_mesh_desc &mesh, int &rc, std::string &err) {
  rc = FSALE_OK;
  if (!mesh.cells || mesh.num_cells <= 0 || mesh.verts_per_cell <= 0) {
    rc = FSALE_ERR_ARG;
    err = "empty mesh";
    return nullptr;
  }
  TopologyCache &C = topology_cache();
  const uint64_t h = mesh_hash(mesh);
  const size_t n = (size_t)mesh.num_cells * mesh.verts_per_cell;
  {
    std::lock_guard<std::mutex> lock(C.mu);
    auto range = C.map.equal_range(h);
    for (auto it = range.first; it != range.second; ++it) {
      const HostTopology &t = *it->second;
      if (t.dim == mesh.dim && t.nc == mesh.num_cells && t.nv == mesh.num_vertices && t.vpc == mesh.verts_per_cell &&
          std::memcmp(t.cv.data(), mesh.cells, n * sizeof(int32_t)) == 0)
        return it->second;
    }
  }
  auto t = std::make_shared<HostTopology>();
  rc = build_topology(mesh, *t, err);
  if (rc) return nullptr;
  std::lock_guard<std::mutex> lock(C.mu);
  const size_t b = topology_bytes(*t);
  C.map.emplace(h, t);
  C.order.emplace_back(h, t.get());
  C.bytes += b;
  while (C.bytes > topology_cache_budget() && C.order.size() > 1) {
    const auto victim = C.order.front();
    C.order.pop_front();
    auto range = C.map.equal_range(victim.first);
    for (auto it = range.first; it != range.second; ++it)
      if (it->second.get() == victim.second) {
        C.bytes -= topology_bytes(*it->second);
        C.map.erase(it);
        break;
      }
  }
  return t;
}

}  // namespace fsale
