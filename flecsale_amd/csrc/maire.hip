// Cell-centred Lagrangian (Maire) step of apps/maire_hydro on the device, 2D
// quads and 3D hexes.
//
//   k_maire_nodal     evaluate_nodal_state      apps/maire_hydro/tasks.h:221-453
//   k_maire_residual  evaluate_residual         apps/maire_hydro/tasks.h:461-489
//   k_maire_dt_*      evaluate_time_step        apps/maire_hydro/tasks.h:142-190
//   k_maire_move      apply_update (mesh move)  apps/maire_hydro/tasks.h:517-534
//   k_maire_update    apply_update (geometry + state) + update_state_from_energy
//                     apps/maire_hydro/tasks.h:536-546, :110-133
//   k_maire_copy      save_/restore_coordinates, save_/restore_solution :557-654
//                     (task-level entry points; the driver step saves inside
//                     the predictor's move and update and needs no restore)
//
// Loops over the batch's vertex tiles (nodal, move) and cell tiles (the
// rest; batch_dev.h) with per-problem scalars.  Wedge facets (half edges: burton's
// wedge facet_normal / facet_area) are recomputed from the vertex coordinates
// where needed instead of being stored: the mesh moves every half step.
// Corner arrays are k-major [vpc][cap_c] so the residual gather of a cell's
// corners is coalesced.  The kernels are templated on the dimension N; a 3D
// corner has six wedge facets, two on each hex side holding its vertex.
#include "batch_dev.h"

namespace fsale {
namespace {

constexpr double kMinSoundSpeed = 1.e-6;   // flecsale/eqns/lagrange_eqns.h:54

// 3D: the packed copy of what the nodal solve gathers per cell, (rho a, p,
// v), written by every kernel that writes one of them
__device__ inline void put_cs(const MaireFields &u, int c, double d, double a, double p, double v0, double v1,
                              double v2) {
  double2 *cs = reinterpret_cast<double2 *>(u.cs) + 3 * (size_t)c;
  cs[0] = make_double2(d * a, p);
  cs[1] = make_double2(v0, v1);
  cs[2] = make_double2(v2, 0.0);
}
__device__ inline void put_xa(const MaireFields &u, int v, double x, double y, double z) {
  reinterpret_cast<double4 *>(u.xa)[v] = make_double4(x, y, z, 0.0);
}
// 2D: (rho a, p) and (v_x, v_y) planes; coordinate pairs
__device__ inline void put_cs2(const MaireFields &u, int c, double d, double a, double p, double v0, double v1) {
  u.cs2[0][c] = make_double2(d * a, p);
  u.cs2[1][c] = make_double2(v0, v1);
}
__device__ inline void put_xy(const DevMesh &m, int v, double x, double y) { m.xy[v] = make_double2(x, y); }

// ------------------------------------------------------------ state init
__global__ void __launch_bounds__(kBlock) k_maire_ic(DevMesh m, MaireFields u, const ProbParams *pp, int prob) {
  const ProbParams &P = pp[prob];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.cell_count) return;
  const int c = P.cell_first + i;
  // initial_conditions apps/maire_hydro/tasks.h:86-100 (u.d holds rho on entry)
  const double vol = m.cvol[c];
  const double M = u.d[c] * vol;
  u.M[c] = M;
  u.V[c] = vol;
  // update_state_from_pressure lagrange_eqns.h:245-271
  const double d = M / vol;
  const double g = u.g[c];
  const double ie = eos_e_dp(g, d, u.p[c]);
  u.d[c] = d;
  u.e[c] = ie;
  u.T[c] = eos_T_de(P.cv, ie);
  const double a = smax(eos_a_de(g, ie), kMinSoundSpeed);
  u.a[c] = a;
  if (m.dim == 3)
    put_cs(u, c, d, a, u.p[c], u.v[0][c], u.v[1][c], u.v[2][c]);
  else
    put_cs2(u, c, d, a, u.p[c], u.v[0][c], u.v[1][c]);
}

// wedge facets of corner (c, k): wedge 0 on edge v_{k-1} -> v_k, wedge 1 on
// edge v_k -> v_{k+1}; outward unit edge normal, facet_area = |edge
// midpoint - vertex| (burton wedge geometry; pinned by the sedov_2d golden).
__device__ inline void wedge_facets(const DevMesh &m, int c, int k, double n[2][2], double l[2]) {
  const int km = (k + 3) & 3, kp = (k + 1) & 3;
  const int im = m.cv[(size_t)km * m.cap_c + c], ik = m.cv[(size_t)k * m.cap_c + c],
            ip = m.cv[(size_t)kp * m.cap_c + c];
  const double2 wm = m.xy[im], wk = m.xy[ik], wp = m.xy[ip];
  const double xm = wm.x, ym = wm.y;
  const double xk = wk.x, yk = wk.y;
  const double xp = wp.x, yp = wp.y;
  const double ax[2] = {xm, xk}, ay[2] = {ym, yk}, bx[2] = {xk, xp}, by[2] = {yk, yp};
#pragma unroll
  for (int w = 0; w < 2; ++w) {
    const double tx = bx[w] - ax[w], ty = by[w] - ay[w];
    const double len = sqrt(tx * tx + ty * ty);
    n[w][0] = ty / len;
    n[w][1] = -tx / len;
    const double mx = 0.5 * (ax[w] + bx[w]) - xk;
    const double my = 0.5 * (ay[w] + by[w]) - yk;
    l[w] = sqrt(mx * mx + my * my);
  }
}

// corner matrix M_pc = sum_w z_c l_w n_w n_w^T (ristra outer_product with
// factor z_c l_w) and (l n)_pc = sum_w l_w n_w.
__device__ inline void corner_matrix_from(const MaireFields &u, int c, const double n[2][2], const double l[2],
                                          double M[2][2], double np[2]) {
  const double zc = u.cs2[0][c].x;   // rho a, as the producers of rho and a multiplied it
  M[0][0] = M[0][1] = M[1][0] = M[1][1] = 0;
  np[0] = np[1] = 0;
#pragma unroll
  for (int w = 0; w < 2; ++w) {
    const double fact = zc * l[w];
    for (int r = 0; r < 2; ++r)
      for (int q = 0; q < 2; ++q) M[r][q] += fact * n[w][r] * n[w][q];
    for (int d = 0; d < 2; ++d) np[d] += l[w] * n[w][d];
  }
}

__device__ inline void corner_matrix(const DevMesh &m, const MaireFields &u, int c, int k, double M[2][2],
                                     double np[2]) {
  double n[2][2], l[2];
  wedge_facets(m, c, k, n, l);
  corner_matrix_from(u, c, n, l, M, np);
}

// Nodal velocity of boundary vertex vt (apps/maire_hydro/tasks.h:317-420)
// under the problem's boundary conditions: the first of the vertex's tags
// (ascending) whose condition prescribes a velocity sets u_p, with no solve
// and no scatter (the task's `continue`); otherwise each boundary wedge, in
// corner order, subtracts l p n from the right-hand side for a prescribed
// pressure, or adds its l n to its tag's symmetry normal, and the augmented
// system [M_p N^T; N 0] is solved by flecsale::linalg::qr (:383-418) --
// Cramer's rule when no wedge carries a symmetry tag.
constexpr int kNodalKeep = 4;
// boundary vertices of quad meshes have one or two corners: the boundary
// path keeps two (its registers then fit beside the interior path's)
constexpr int kNodalKeepBnd = 2;

// the first tag (ascending) of tag mask `vtags` whose condition prescribes
// the velocity, or -1
__device__ inline int velocity_tag(const ProbParams &P, int vtags, int dim) {
  for (int t = 0; t < dim; ++t)
    if ((vtags & (1 << t)) && P.bc_kind[t] == FSALE_BC_VELOCITY) return t;
  return -1;
}

// keep: wedge normals, facet areas and tags of the first kNodalKeep corners,
// already loaded by the caller
template <int K>
struct KeptWedges {
  double n[K][2][2], l[K][2];
  int tags[K];
};

// returns false for a prescribed velocity (the vertex's corners keep
// F_pc = M_pc u_c + p_c (l n)_pc)
template <int K>
__device__ inline bool boundary_solve(const DevMesh &m, const ProbParams &P, int vt, int cnt, const double Mp[2][2],
                                      double rhs[2], const KeptWedges<K> &kw, double un[2]) {
  {
    int vtags = 0;
#pragma unroll
    for (int q = 0; q < K; ++q)
      if (q < cnt) vtags |= kw.tags[q];
    for (int q = K; q < cnt; ++q) {
      const int cn = m.vck[(size_t)q * m.cap_v + vt];
      vtags |= m.cwtag[(size_t)(cn & 3) * m.cap_c + (cn >> 2)];
    }
    // nibble w: bit 2 boundary, bits 0-1 its tags
    const int vel = velocity_tag(P, (vtags | (vtags >> 4)) & 3, 2);
    if (vel >= 0) {
      un[0] = P.bc_val[vel][0];
      un[1] = P.bc_val[vel][1];
      return false;
    }
  }
  int has[2] = {0, 0};
  double symn[2][2] = {{0, 0}, {0, 0}};
  auto add_wedges = [&](int tags, const double n[2][2], const double l[2]) {
#pragma unroll
    for (int w = 0; w < 2; ++w) {
      const int t = (tags >> (4 * w)) & 0xf;
      if (!(t & 4)) continue;
#pragma unroll
      for (int tag = 0; tag < 2; ++tag) {
        if (!(t & (1 << tag))) continue;
        const int kind = P.bc_kind[tag];
        if (kind == FSALE_BC_PRESSURE) {
          const double fact = l[w] * P.bc_val[tag][0];
          for (int d = 0; d < 2; ++d) rhs[d] -= fact * n[w][d];
        } else if (kind == FSALE_BC_SYMMETRY) {
          if (has[tag]) {
            for (int d = 0; d < 2; ++d) symn[tag][d] += l[w] * n[w][d];
          } else {
            has[tag] = 1;
            for (int d = 0; d < 2; ++d) symn[tag][d] = l[w] * n[w][d];
          }
        }
      }
    }
  };
#pragma unroll
  for (int q = 0; q < K; ++q)
    if (q < cnt && (kw.tags[q] & 0x44)) add_wedges(kw.tags[q], kw.n[q], kw.l[q]);
  for (int q = K; q < cnt; ++q) {
    const int cn = m.vck[(size_t)q * m.cap_v + vt];
    const int c = cn >> 2, k = cn & 3;
    const int tags = m.cwtag[(size_t)k * m.cap_c + c];
    if (!(tags & 0x44)) continue;
    double n[2][2], l[2];
    wedge_facets(m, c, k, n, l);
    add_wedges(tags, n, l);
  }
  // augmented system [M_p N^T; N 0] solved by flecsale::linalg::qr
  // (apps/maire_hydro/tasks.h:383-418), rows of the tags in order
  const int nsym = has[0] + has[1];
  if (nsym == 0) {
    const double det = Mp[0][0] * Mp[1][1] - Mp[0][1] * Mp[1][0];
    un[0] = (rhs[0] * Mp[1][1] - Mp[0][1] * rhs[1]) / det;
    un[1] = (Mp[0][0] * rhs[1] - rhs[0] * Mp[1][0]) / det;
    return true;
  }
  if (nsym == 1) {
    const double sn0 = has[0] ? symn[0][0] : symn[1][0];
    const double sn1 = has[0] ? symn[0][1] : symn[1][1];
    double A[3][3] = {{Mp[0][0], Mp[0][1], sn0}, {Mp[1][0], Mp[1][1], sn1}, {sn0, sn1, 0}};
    double bb[3] = {rhs[0], rhs[1], 0};
    qr_solve_fixed<3>(A, bb);
    un[0] = bb[0];
    un[1] = bb[1];
  } else {
    double A[4][4] = {{Mp[0][0], Mp[0][1], symn[0][0], symn[1][0]},
                      {Mp[1][0], Mp[1][1], symn[0][1], symn[1][1]},
                      {symn[0][0], symn[0][1], 0, 0},
                      {symn[1][0], symn[1][1], 0, 0}};
    double bb[4] = {rhs[0], rhs[1], 0, 0};
    qr_solve_fixed<4>(A, bb);
    un[0] = bb[0];
    un[1] = bb[1];
  }
  return true;
}

// ------------------------------------------------------------ 3D corners
// For local vertex k of a hex, its three sides j in side order (Exodus hex8
// sides {0,1,5,4} {1,2,6,5} {2,3,7,6} {0,4,7,3} {0,3,2,1} {4,5,6,7}): the
// local vertices after, opposite and before k in side j's outward list,
// nibble k of kRingNext[j], kRingOpp[j], kRingPrev[j].
constexpr uint32_t kRingNext[3] = {0x65407651u, 0x32170324u, 0x47652103u};
constexpr uint32_t kRingOpp[3] = {0x21016545u, 0x03234767u, 0x54761032u};
constexpr uint32_t kRingPrev[3] = {0x32152104u, 0x47607653u, 0x65470321u};

// The six wedge facets of 3D corner (c, k) at vertex ik, in order: for each
// side j holding k, the triangles (x_k, m+, x_s) and (x_k, x_s, m-), m+ / m-
// the midpoints of the side's edges after / before k and x_s the side's
// vertex average summed as ((a0 + a2) + (a1 + a3)) (the same bits from both
// cells of an interior side).  f(j, n, l) receives each facet's unit normal
// and area (burton's wedge facet_normal / facet_area; the normal is the area
// vector times the reciprocal of its length).
template <class F>
__device__ inline void for_each_facet3(const MaireFields &u, int c, int k, int ik, F &&f) {
  const int32_t *row = u.cv8 + 8 * (size_t)c;
  auto coords = [&](int v, double x[3]) {
    const double4 t = reinterpret_cast<const double4 *>(u.xa)[v];
    x[0] = t.x;
    x[1] = t.y;
    x[2] = t.z;
  };
  double xk[3];
  coords(ik, xk);
  const int sh = 4 * k;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double xn[3], xo[3], xp[3];
    coords(row[(kRingNext[j] >> sh) & 0xf], xn);
    coords(row[(kRingOpp[j] >> sh) & 0xf], xo);
    coords(row[(kRingPrev[j] >> sh) & 0xf], xp);
    double xs[3], mp[3], mm[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      xs[d] = ((xk[d] + xo[d]) + (xn[d] + xp[d])) * 0.25;
      mp[d] = 0.5 * (xk[d] + xn[d]);
      mm[d] = 0.5 * (xp[d] + xk[d]);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double a[3], b[3], t[3], n[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        a[d] = (h == 0 ? mp[d] : xs[d]) - xk[d];
        b[d] = (h == 0 ? xs[d] : mm[d]) - xk[d];
      }
      cross3(a, b, t);
#pragma unroll
      for (int d = 0; d < 3; ++d) t[d] = 0.5 * t[d];
      const double len = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
      const double inv = 1.0 / len;
#pragma unroll
      for (int d = 0; d < 3; ++d) n[d] = t[d] * inv;
      f(j, n, len);
    }
  }
}

// corner matrix M_pc = sum_w z_c l_w n_w n_w^T, (l n)_pc = sum_w l_w n_w and
// the corner force before the nodal velocity, F_pc = M_pc u_c + p_c (l n)_pc
// (apps/maire_hydro/tasks.h:263-303)
// KEEP: also each facet's unit normal and area, nl[w] = (n, l), in facet
// order (the boundary conditions' pressure and symmetry terms)
template <bool KEEP = false>
__device__ inline void corner3(const MaireFields &u, int c, int k, int vt, double M[3][3], double np[3],
                               double F[3], double (*nl)[4] = nullptr) {
  const double2 s0 = reinterpret_cast<const double2 *>(u.cs)[3 * (size_t)c];
  const double2 s1 = reinterpret_cast<const double2 *>(u.cs)[3 * (size_t)c + 1];
  const double2 s2 = reinterpret_cast<const double2 *>(u.cs)[3 * (size_t)c + 2];
  const double zc = s0.x;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    np[r] = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) M[r][q] = 0;
  }
  int w = 0;
  for_each_facet3(u, c, k, vt, [&](int, const double n[3], double l) {
    const double fact = zc * l;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) M[r][q] += fact * n[r] * n[q];
#pragma unroll
    for (int d = 0; d < 3; ++d) np[d] += l * n[d];
    if constexpr (KEEP) {
#pragma unroll
      for (int d = 0; d < 3; ++d) nl[w][d] = n[d];
      nl[w][3] = l;
    }
    ++w;
  });
  const double uc[3] = {s1.x, s1.y, s2.x};
  const double pc = s0.y;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    F[r] = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) F[r] += M[r][q] * uc[q];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) F[d] += pc * np[d];
}

// the packed gathers of one problem's cells [c0, c0 + nc) and vertices
// [v0, v0 + nv) from the cell and vertex arrays (after host writes; the
// step's kernels keep them current themselves)
__global__ void __launch_bounds__(kBlock) k_maire3_pack(DevMesh m, MaireFields u, int c0, int nc, int v0, int nv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) {
    const int c = c0 + i;
    put_cs(u, c, u.d[c], u.a[c], u.p[c], u.v[0][c], u.v[1][c], u.v[2][c]);
  }
  if (i < nv) {
    const int v = v0 + i;
    put_xa(u, v, m.x[0][v], m.x[1][v], m.x[2][v]);
  }
}

__global__ void __launch_bounds__(kBlock) k_maire2_pack(DevMesh m, MaireFields u, int c0, int nc, int v0, int nv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) {
    const int c = c0 + i;
    put_cs2(u, c, u.d[c], u.a[c], u.p[c], u.v[0][c], u.v[1][c]);
  }
  if (i < nv) {
    const int v = v0 + i;
    put_xy(m, v, m.x[0][v], m.x[1][v]);
  }
}

// ristra::math::solve (un-vendored) of the 3x3 nodal system: adjugate over
// the determinant
__device__ inline void solve3(const double M[3][3], const double r[3], double x[3]) {
  const double c00 = M[1][1] * M[2][2] - M[1][2] * M[2][1];
  const double c01 = M[1][2] * M[2][0] - M[1][0] * M[2][2];
  const double c02 = M[1][0] * M[2][1] - M[1][1] * M[2][0];
  const double det = (M[0][0] * c00 + M[0][1] * c01) + M[0][2] * c02;
  x[0] = ((c00 * r[0] + (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * r[1]) + (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * r[2]) /
         det;
  x[1] = ((c01 * r[0] + (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * r[1]) + (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * r[2]) /
         det;
  x[2] = ((c02 * r[0] + (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * r[1]) + (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * r[2]) /
         det;
}

// the augmented system [M_p N^T; N 0] of NS symmetry rows (tags in order)
template <int NS>
__device__ inline void sym_solve3(const double Mp[3][3], const double rhs[3], const double (&sn)[3][3],
                                  double un[3]) {
  constexpr int NR = 3 + NS;
  double A[NR][NR], b[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    b[r] = r < 3 ? rhs[r] : 0.0;
#pragma unroll
    for (int q = 0; q < NR; ++q)
      A[r][q] = r < 3 ? (q < 3 ? Mp[r][q] : sn[q - 3][r]) : (q < 3 ? sn[r - 3][q] : 0.0);
  }
  qr_solve_fixed<NR>(A, b);
#pragma unroll
  for (int d = 0; d < 3; ++d) un[d] = b[d];
}

// evaluate_nodal_state for 3D vertices (apps/maire_hydro/tasks.h:249-448),
// eight lanes per vertex: lane q of a group builds corner q's matrix, normal
// and force once and posts the matrix and force in LDS; lane q then sums
// entries q and q + 8 of the twelve over the vertex's corners in corner
// order (the sequential sums, bit for bit), every lane of the group reads
// the twelve sums, solves (the 3x3 system inside; on the boundary the QR of
// [M_p N^T; N 0] with the rows of the x / y / z tags present, normals summed
// over the vertex's boundary wedges in corner and facet order, :338-418),
// and finishes its own corner, F_pc -= M_pc u_p.  cnt = 0: no vertex; such
// lanes still take part in the exchanges (every lane of the block calls this
// together: it holds block barriers).
constexpr int kGroups3 = kBlock / kMaxCornersPerVertex;   // vertices per pass of a block

struct Nodal3Lds {
  double corner[kGroups3][kMaxCornersPerVertex][12];   // M (row-major), F
  double sum[kGroups3][12];
};
struct Nodal3BndLds {
  double facet[kGroups3][kMaxCornersPerVertex][6][4];  // n, l of the six facets
  int tags[kGroups3][kMaxCornersPerVertex];
  double sn[kGroups3][9];                              // symmetry normals [tag][d]
  int has[kGroups3][3];
};

// KIND 0: interior vertices; 1: boundary vertices with a prescribed velocity
// or at most one symmetry row; 2: the other boundary vertices (two or three
// symmetry rows: the 5x5 and 6x6 solves, whose registers would otherwise
// hold down the first boundary launch's occupancy)
template <int KIND>
__device__ inline void nodal3_group(const DevMesh &m, const MaireFields &u, const ProbParams &P, int vt, int cnt,
                                    int g, int q, Nodal3Lds &L, Nodal3BndLds *B) {
  constexpr bool BOUNDARY = KIND != 0;
  const bool mine = q < cnt;
  double M[3][3], np[3], F[3], ln[6][4];
  int c = 0, k = 0;
  if (mine) {
    const int cn = m.vck[(size_t)q * m.cap_v + vt];
    c = cn >> 3;
    k = cn & 7;
    corner3<BOUNDARY>(u, c, k, vt, M, np, F, ln);
  } else {
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      np[r] = F[r] = 0;
#pragma unroll
      for (int s2 = 0; s2 < 3; ++s2) M[r][s2] = 0;
    }
#pragma unroll
    for (int w = 0; w < 6; ++w)
#pragma unroll
      for (int d = 0; d < 4; ++d) ln[w][d] = 0;
  }
  {
    double *row = L.corner[g][q];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
      for (int s2 = 0; s2 < 3; ++s2) row[3 * r + s2] = M[r][s2];
      row[9 + r] = F[r];
    }
  }
  if constexpr (BOUNDARY) {
#pragma unroll
    for (int w = 0; w < 6; ++w)
#pragma unroll
      for (int d = 0; d < 4; ++d) B->facet[g][q][w][d] = ln[w][d];
    B->tags[g][q] = mine ? (int)m.cwtag[(size_t)k * m.cap_c + c] : 0;
  }
  __syncthreads();
  // entries q and q + 8 summed over the corners in order; the sums start at
  // +0 and never become -0, so adding +0 for a missing corner is exact
  {
    double s0 = 0, s1 = 0;
#pragma unroll
    for (int j = 0; j < kMaxCornersPerVertex; ++j) {
      const double a0 = L.corner[g][j][q];
      const double a1 = L.corner[g][j][q < 4 ? q + 8 : q];   // lanes 4-7: unused
      s0 += j < cnt ? a0 : 0.0;
      s1 += j < cnt ? a1 : 0.0;
    }
    L.sum[g][q] = s0;
    if (q < 4) L.sum[g][q + 8] = s1;
  }
  if constexpr (BOUNDARY) {
    // symmetry normal entry (t, d) = P / 3, P % 3 for P = q and, on lane 0,
    // P = 8: first boundary wedge of tag t assigns, later ones add
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
      const int Pe = pass == 0 ? q : 8;
      if (pass == 1 && q != 0) break;
      const int t = Pe / 3, d = Pe - 3 * t;
      double sum = 0;
      int has = 0;
#pragma unroll 1
      for (int j = 0; j < kMaxCornersPerVertex; ++j) {
        if (j >= cnt || P.bc_kind[t] != FSALE_BC_SYMMETRY) break;
        const int tj = B->tags[g][j];
        if (!(tj & 0x888)) continue;
#pragma unroll
        for (int w = 0; w < 6; ++w) {
          const int tg = (tj >> (4 * (w >> 1))) & 0xf;
          if (!(tg & 8) || !(tg & (1 << t))) continue;
          const double v = B->facet[g][j][w][3] * B->facet[g][j][w][d];
          if (has) {
            sum += v;
          } else {
            sum = v;
            has = 1;
          }
        }
      }
      B->sn[g][Pe] = sum;
      if (d == 0) B->has[g][t] = has;
    }
  }
  __syncthreads();
  double Mp[3][3], rhs[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
#pragma unroll
    for (int s2 = 0; s2 < 3; ++s2) Mp[r][s2] = L.sum[g][3 * r + s2];
    rhs[r] = L.sum[g][9 + r];
  }
  double un[3] = {0, 0, 0};
  bool scatter = true;
  if constexpr (BOUNDARY) {
    // the vertex's tags; a prescribed velocity ends the solve here
    int vtags = 0;
    for (int j = 0; j < cnt; ++j) {
      const int tj = B->tags[g][j];
#pragma unroll
      for (int s2 = 0; s2 < 3; ++s2)
        if ((tj >> (4 * s2)) & 8) vtags |= (tj >> (4 * s2)) & 7;
    }
    const int vel = velocity_tag(P, vtags, 3);
    // prescribed pressures: rhs -= (l p) n over the boundary wedges in order
    if (vel < 0 && (P.bc_kind[0] == FSALE_BC_PRESSURE || P.bc_kind[1] == FSALE_BC_PRESSURE ||
                    P.bc_kind[2] == FSALE_BC_PRESSURE)) {
#pragma unroll 1
      for (int j = 0; j < cnt; ++j) {
        const int tj = B->tags[g][j];
        for (int w = 0; w < 6; ++w) {
          const int tg = (tj >> (4 * (w >> 1))) & 0xf;
          if (!(tg & 8)) continue;
          for (int t = 0; t < 3; ++t) {
            if (!(tg & (1 << t)) || P.bc_kind[t] != FSALE_BC_PRESSURE) continue;
            const double fact = B->facet[g][j][w][3] * P.bc_val[t][0];
#pragma unroll
            for (int d = 0; d < 3; ++d) rhs[d] -= fact * B->facet[g][j][w][d];
          }
        }
      }
    }
    double sn[3][3];
#pragma unroll
    for (int t = 0; t < 3; ++t)
#pragma unroll
      for (int d = 0; d < 3; ++d) sn[t][d] = B->sn[g][3 * t + d];
    const bool h0 = B->has[g][0], h1 = B->has[g][1], h2 = B->has[g][2];
    const int ns = (int)h0 + (int)h1 + (int)h2;
    // the rows of the tags present, in tag order
    double rows[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      rows[0][d] = h0 ? sn[0][d] : (h1 ? sn[1][d] : sn[2][d]);
      rows[1][d] = (h0 && h1) ? sn[1][d] : sn[2][d];
      rows[2][d] = sn[2][d];
    }
    if (vel >= 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) un[d] = P.bc_val[vel][d];
      scatter = false;
    } else if (cnt > 0) {
      if constexpr (KIND == 1) {
        if (ns == 0)
          solve3(Mp, rhs, un);
        else
          sym_solve3<1>(Mp, rhs, rows, un);
      } else {
        if (ns == 2)
          sym_solve3<2>(Mp, rhs, rows, un);
        else
          sym_solve3<3>(Mp, rhs, rows, un);
      }
    }
  } else {
    if (cnt > 0) solve3(Mp, rhs, un);
  }
  if (cnt > 0 && q == 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) u.un[d][vt] = un[d];
  }
  // F_pc -= M_pc u_p  (matrix_vector(-1, Mpc, un, 1, Fpc), tasks.h:435-448)
  if (mine) {
    const size_t ci = (size_t)k * m.cap_c + c;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      u.npc[r][ci] = np[r];
      double t = M[r][0] * un[0];
      t += M[r][1] * un[1];
      t += M[r][2] * un[2];
      u.Fpc[r][ci] = scatter ? -1.0 * t + 1.0 * F[r] : F[r];
    }
  }
  // the next pass of the block overwrites the exchange arrays
  __syncthreads();
}

// one tile of 3D vertices, 32 per pass of the block's 256 lanes
template <int KIND>
__device__ inline void nodal3_tile(const DevMesh &m, const MaireFields &u, const Batch &b, const Tile &tile) {
  if (!wave_live(b.s, tile.prob, 0)) return;
  constexpr bool BOUNDARY = KIND != 0;
  __shared__ Nodal3Lds lds;
  __shared__ typename std::conditional<BOUNDARY, Nodal3BndLds, char>::type blds;
  constexpr int G = kMaxCornersPerVertex;
  const int q = threadIdx.x & (G - 1), g = threadIdx.x / G;
  const ProbParams &P = b.pp[tile.prob];
  for (int base = 0; base < tile.count; base += kGroups3) {
    const int i = base + g;
    int vt = tile.first + i, cnt = 0;
    if (i < tile.count) {
      if constexpr (BOUNDARY) {
        vt = m.bv[vt];
        cnt = m.vc_cnt[vt];
      } else {
        cnt = m.vbnd[vt] ? 0 : m.vc_cnt[vt];
      }
    }
    Nodal3BndLds *bp = nullptr;
    if constexpr (BOUNDARY) {
      bp = &blds;
      // the vertex's class from its corners' tags: the symmetry rows its
      // solve will have (tags whose condition is symmetry), or a velocity
      int tq = 0;
      if (q < cnt) {
        const int cn = m.vck[(size_t)q * m.cap_v + vt];
        tq = m.cwtag[(size_t)(cn & 7) * m.cap_c + (cn >> 3)];
      }
      blds.tags[g][q] = tq;
      __syncthreads();
      int vtags = 0;
      for (int j = 0; j < cnt; ++j) {
        const int tj = blds.tags[g][j];
#pragma unroll
        for (int s2 = 0; s2 < 3; ++s2)
          if ((tj >> (4 * s2)) & 8) vtags |= (tj >> (4 * s2)) & 7;
      }
      int ns = 0;
#pragma unroll
      for (int t = 0; t < 3; ++t) ns += ((vtags >> t) & 1) && P.bc_kind[t] == FSALE_BC_SYMMETRY;
      const bool small = velocity_tag(P, vtags, 3) >= 0 || ns <= 1;
      if (small != (KIND == 1)) cnt = 0;
      __syncthreads();
    }
    nodal3_group<KIND>(m, u, P, vt, cnt, g, q, lds, bp);
  }
}

// evaluate_nodal_state, one work-item per vertex.  The vertex's corners
// (ascending cell id, k-major table) are visited in order; the first four
// corner matrices stay in registers for the final F_pc -= M_pc u_p (a vertex
// with more corners recomputes the rest), and the first four corners' loads
// do not depend on each other.

__device__ inline void corner_force(const MaireFields &u, int c, const double M[2][2], const double np[2],
                                    double F[2]) {
  const double2 zp = u.cs2[0][c], vc = u.cs2[1][c];
  const double uc[2] = {vc.x, vc.y};
  const double pc = zp.y;
  for (int r = 0; r < 2; ++r) {
    F[r] = 0;
    F[r] += M[r][0] * uc[0];
    F[r] += M[r][1] * uc[1];
  }
  for (int d = 0; d < 2; ++d) F[d] += pc * np[d];
}

// BOUNDARY = false: the interior vertices (Cramer's rule); true: the vertices
// on a boundary face, whose symmetry QR solve would otherwise hold the
// registers (and the waves) of the whole interior launch.
// one tile of interior (Cramer) or boundary (symmetry QR) vertices
// Whether a problem's waves work in a 2D driver step's first nodal solve and
// residual.  PEND: the previous step's close (k_maire_end's update of t,
// steps and the step flag) has not run yet -- the predictor's update
// (k_maire_move_update FOLD) writes it right after these two launches -- so
// the step flag it would set is computed here, from the same values, which
// nothing writes meanwhile.
template <bool PEND>
__device__ inline bool step_live(const Batch &b, int prob) {
  if (!wave_live(b.s, prob, 0)) return false;
  if constexpr (PEND) {
    const DevScalars &S = b.s[prob];
    const ProbParams &P = b.pp[prob];
    return S.err_code == 0 && S.steps + 1 < P.max_steps && S.t + S.dt < P.final_time;
  }
  return true;
}

// SAVE (2D, a driver step's first nodal solve): each vertex also saves its
// coordinates as the step's start (x0p; save_coordinates), which no kernel
// of this launch reads.
template <bool BOUNDARY, int N, bool SAVE = false, bool PEND = false>
__device__ inline void nodal_tile(const DevMesh &m, const MaireFields &u, const Batch &b, const Tile &tile) {
  if constexpr (N == 3) {
    nodal3_tile<BOUNDARY ? 1 : 0>(m, u, b, tile);
  } else {
    if (tile.count <= 0) return;
    // the vertex, its first corners and its boundary flag are read ahead of
    // the step check, so that the loads overlap (discarded when the tile
    // does not step; lanes past the tile read its last vertex)
    int vt = tile.first + min((int)threadIdx.x, tile.count - 1);
    // interior: a boundary vertex (the boundary path's) computes like the
    // others with a corner count of 0, so that nothing it stores and no
    // gather waits on the boundary flag (a branch on it would)
    int cn0[BOUNDARY ? 1 : kNodalKeep];
    bool own = true;
    if constexpr (BOUNDARY) {
      vt = m.bv[vt];
    } else {
#pragma unroll
      for (int q = 0; q < kNodalKeep; ++q) cn0[q] = m.vck[(size_t)q * m.cap_v + vt];
      own = !m.vbnd[vt];
    }
    if (!step_live<PEND>(b, tile.prob)) return;
    if ((int)threadIdx.x >= tile.count) return;
    if (SAVE && own) u.x0p[vt] = m.xy[vt];   // u.x0 follows in k_maire_settle
    const int cnt = m.vc_cnt[vt] & -(int)own;
    constexpr int KEEP = BOUNDARY ? kNodalKeepBnd : kNodalKeep;
    double M[KEEP][2][2], np[KEEP][2], F[KEEP][2];
    int cc[KEEP], kk[KEEP];
    KeptWedges<KEEP> kw;
    if constexpr (BOUNDARY) {
#pragma unroll
      for (int q = 0; q < KEEP; ++q) {
        const int cn = q < cnt ? m.vck[(size_t)q * m.cap_v + vt] : -1;
        cc[q] = cn >> 2;
        kk[q] = cn & 3;
        if (cn >= 0) {
          wedge_facets(m, cc[q], kk[q], kw.n[q], kw.l[q]);
          kw.tags[q] = m.cwtag[(size_t)kk[q] * m.cap_c + cc[q]];
          corner_matrix_from(u, cc[q], kw.n[q], kw.l[q], M[q], np[q]);
          corner_force(u, cc[q], M[q], np[q], F[q]);
          // n_pc is final here: stored now, its registers free for the solve
          const size_t ci = (size_t)kk[q] * m.cap_c + cc[q];
          u.npc[0][ci] = np[q][0];
          u.npc[1][ci] = np[q][1];
        }
      }
    } else {
      // every kept corner's gathers without a branch, so that the four
      // corners' chains (corner -> cell vertices -> coordinates) overlap: a
      // missing corner (-1 past the vertex's count) computes the first
      // corner again and is not used
#pragma unroll
      for (int q = 0; q < KEEP; ++q) {
        const int cn = cn0[q] >= 0 ? cn0[q] : cn0[0];
        cc[q] = cn >> 2;
        kk[q] = cn & 3;
        corner_matrix(m, u, cc[q], kk[q], M[q], np[q]);
        corner_force(u, cc[q], M[q], np[q], F[q]);
      }
#pragma unroll
      for (int q = 0; q < KEEP; ++q)
        if (q < cnt) {
          const size_t ci = (size_t)kk[q] * m.cap_c + cc[q];
          u.npc[0][ci] = np[q][0];
          u.npc[1][ci] = np[q][1];
        }
    }
    // M_p = sum_c M_pc and rhs = sum_c F_pc in corner order
    double Mp[2][2] = {{0, 0}, {0, 0}}, rhs[2] = {0, 0};
#pragma unroll
    for (int q = 0; q < KEEP; ++q) {
      if (q >= cnt) break;
      for (int r = 0; r < 2; ++r) {
        Mp[r][0] += M[q][r][0];
        Mp[r][1] += M[q][r][1];
        rhs[r] += F[q][r];
      }
    }
    for (int q = KEEP; q < cnt; ++q) {
      const int cn = m.vck[(size_t)q * m.cap_v + vt];
      double Mq[2][2], npq[2], Fq[2];
      corner_matrix(m, u, cn >> 2, cn & 3, Mq, npq);
      corner_force(u, cn >> 2, Mq, npq, Fq);
      for (int r = 0; r < 2; ++r) {
        Mp[r][0] += Mq[r][0];
        Mp[r][1] += Mq[r][1];
        rhs[r] += Fq[r];
      }
    }
    double un[2];
    bool scatter = true;
    if constexpr (BOUNDARY) {
      scatter = boundary_solve(m, b.pp[tile.prob], vt, cnt, Mp, rhs, kw, un);
    } else {
      // ristra::math::solve (un-vendored): Cramer's rule
      const double det = Mp[0][0] * Mp[1][1] - Mp[0][1] * Mp[1][0];
      un[0] = (rhs[0] * Mp[1][1] - Mp[0][1] * rhs[1]) / det;
      un[1] = (Mp[0][0] * rhs[1] - rhs[0] * Mp[1][0]) / det;
    }
    if (own) {
      u.un[0][vt] = un[0];
      u.un[1][vt] = un[1];
    }
    // F_pc -= M_pc u_p  (matrix_vector(-1, Mpc, un, 1, Fpc), tasks.h:435-448)
    auto finish = [&](int c, int k, const double Mq[2][2], const double Fq[2]) {
      const size_t ci = (size_t)k * m.cap_c + c;
      for (int r = 0; r < 2; ++r) {
        double t = Mq[r][0] * un[0];
        t += Mq[r][1] * un[1];
        u.Fpc[r][ci] = scatter ? -1.0 * t + 1.0 * Fq[r] : Fq[r];
      }
    };
#pragma unroll
    for (int q = 0; q < KEEP; ++q)
      if (q < cnt) finish(cc[q], kk[q], M[q], F[q]);
    for (int q = KEEP; q < cnt; ++q) {
      const int cn = m.vck[(size_t)q * m.cap_v + vt];
      double Mq[2][2], npq[2], Fq[2];
      corner_matrix(m, u, cn >> 2, cn & 3, Mq, npq);
      corner_force(u, cn >> 2, Mq, npq, Fq);
      const size_t ci = (size_t)(cn & 3) * m.cap_c + (cn >> 2);
      u.npc[0][ci] = npq[0];
      u.npc[1][ci] = npq[1];
      finish(cn >> 2, cn & 3, Mq, Fq);
    }
  }
}

template <bool BOUNDARY, int N, bool SAVE = false, bool PEND = false>
__global__ void __launch_bounds__(kBlock) k_maire_nodal(DevMesh m, MaireFields u, Batch b) {
  for_each_tile(BOUNDARY ? b.bt : b.vt,
                [&](const Tile &tile) { nodal_tile<BOUNDARY, N, SAVE, PEND>(m, u, b, tile); });
}

// 3D boundary vertices with two or three symmetry rows
__global__ void __launch_bounds__(kBlock) k_maire3_nodal_bnd_big(DevMesh m, MaireFields u, Batch b) {
  for_each_tile(b.bt, [&](const Tile &tile) { nodal3_tile<2>(m, u, b, tile); });
}

// Both kinds in one launch over the one-block-per-tile lists: blocks
// [0, nb) take the boundary tiles (nb a multiple of 8, so the interior list
// keeps its XCD arrangement), the rest the interior tiles.  The boundary
// vertices' solves are long serial chains in few waves; issued first, they
// run while the interior waves fill the machine around them instead of
// forming a launch of their own.  Held at 3 waves per SIMD for the interior
// tiles (the boundary path then keeps part of its state in scratch).
#ifndef FSALE_NODAL_WAVES
#define FSALE_NODAL_WAVES 3
#endif
template <int N, bool SAVE = false, bool PEND = false>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(FSALE_NODAL_WAVES)))
k_maire_nodal_both(DevMesh m, MaireFields u, Batch b, int nb) {
  if ((int)blockIdx.x < nb) {
    if ((int)blockIdx.x < b.bt.ndirect) {
      const Tile t = b.bt.direct[blockIdx.x];
      if (t.count > 0) nodal_tile<true, N, SAVE, PEND>(m, u, b, t);
    }
  } else {
    const int i = (int)blockIdx.x - nb;
    if (i < b.vt.ndirect) {
      const Tile t = b.vt.direct[i];
      if (t.count > 0) nodal_tile<false, N, SAVE, PEND>(m, u, b, t);
    }
  }
}

// evaluate_residual of cell c (compute_update lagrange_eqns.h:374-389): its
// corners' normals and forces against its vertices' velocities, in local
// vertex order; vt / uv return the vertex ids and velocities it read.
template <int N, int V, bool PRE = false>
__device__ __forceinline__ void cell_residual(const DevMesh &m, const MaireFields &u, int c, double r[N + 2],
                                              int vt[V], double uv[V][N]) {
#pragma unroll
  for (int i = 0; i < N + 2; ++i) r[i] = 0;
#pragma unroll
  for (int k = 0; k < V; ++k) {
    const size_t ci = (size_t)k * m.cap_c + c;
    if (!PRE) vt[k] = m.cv[ci];   // PRE: the caller read them
    double F[N], n[N];
#pragma unroll
    for (int d = 0; d < N; ++d) {
      uv[k][d] = u.un[d][vt[k]];
      F[d] = u.Fpc[d][ci];
      n[d] = u.npc[d][ci];
    }
#pragma unroll
    for (int d = 0; d < N; ++d) {
      r[0] += n[d] * uv[k][d];
      r[N + 1] -= F[d] * uv[k][d];
      r[1 + d] -= F[d];
    }
  }
}

// DT: also the time step's per-wave partials of the new residual (the
// driver step's predictor, whose time step follows), as k_maire_dt_partial
// computes them.  TILE (2D driver step, with DT): one partial pair per cell
// tile, at 2 slot (a block-wide max), and the block of each problem's first
// tile records the step as its predictor update will see it (DevScalars::
// snap_*: with PEND the previous step closed as close_step closes it).
template <int N, bool DT, bool PEND = false, bool TILE = false>
__global__ void __launch_bounds__(kBlock) k_maire_residual(DevMesh m, MaireFields u, Batch b) {
  constexpr int V = N == 2 ? 4 : 8;
  static_assert(!TILE || DT, "tile partials are time-step partials");
  for_each_tile(b.ct, [&](const Tile &tile) {
    // the cell's vertex ids are read ahead of the step check, so that the
    // loads overlap (discarded when the tile does not step; lanes past the
    // tile read its last cell's)
    int vt[V];
    {
      const int cp = tile.first + max(min((int)threadIdx.x, tile.count - 1), 0);
#pragma unroll
      for (int k = 0; k < V; ++k) vt[k] = m.cv[(size_t)k * m.cap_c + cp];
    }
    if (TILE && threadIdx.x == 0 && tile.first == b.pp[tile.prob].cell_first) {
      DevScalars &S = b.s[tile.prob];
      const ProbParams &P = b.pp[tile.prob];
      double t = S.t;
      long long steps = S.steps;
      int active = S.active;
      if (PEND) {
        if (S.active && !S.err_code) {
          t += S.dt;
          steps = steps + 1;
        }
        active = (S.err_code == 0) && (steps < P.max_steps) && (t < P.final_time);
      }
      S.snap_t = t;
      S.snap_dt = S.dt;
      S.snap_steps = steps;
      S.snap_active = active;
    }
    if (!step_live<PEND>(b, tile.prob)) return;
    double acc = 0, vol = 0;
    if ((int)threadIdx.x < tile.count) {
    const int c = tile.first + threadIdx.x;
    double r[N + 2], uv[V][N];
    cell_residual<N, V, true>(m, u, c, r, vt, uv);
#pragma unroll
    for (int i = 0; i < N + 2; ++i) u.dudt[i][c] = r[i];
    if (DT) {
      acc = u.a[c] / m.cminlen[c];
      vol = fabs(r[0]) / m.cvol[c];
    }
    }
    if (DT) {
      acc = wave_max(acc);
      vol = wave_max(vol);
      if (TILE) {
        __shared__ double red[2][kWaves];
        if ((threadIdx.x & 63) == 0) {
          red[0][threadIdx.x >> 6] = acc;
          red[1][threadIdx.x >> 6] = vol;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
          for (int w = 1; w < kWaves; ++w) {
            acc = smax(red[0][w], acc);
            vol = smax(red[1][w], vol);
          }
          u.partials[2 * (size_t)tile.slot] = acc;
          u.partials[2 * (size_t)tile.slot + 1] = vol;
        }
        __syncthreads();   // red is free again (a persistent block's next tile)
      } else if ((threadIdx.x & 63) == 0) {
        const size_t i = 2 * ((size_t)tile.slot * kWaves + (threadIdx.x >> 6));
        u.partials[i] = acc;
        u.partials[i + 1] = vol;
      }
    }
  });
}

// per cell tile: max a/min_length and max |dV/dt|/V
__global__ void __launch_bounds__(kBlock) k_maire_dt_partial(DevMesh m, MaireFields u, Batch b) {
  for_each_tile(b.ct, [&](const Tile &tile) {
    if (!wave_live(b.s, tile.prob, 0)) return;
    double acc = 0, vol = 0;
    if ((int)threadIdx.x < tile.count) {
      const int c = tile.first + threadIdx.x;
      acc = u.a[c] / m.cminlen[c];
      vol = fabs(u.dudt[0][c]) / m.cvol[c];
    }
    acc = wave_max(acc);
    vol = wave_max(vol);
    if ((threadIdx.x & 63) == 0) {
      const size_t i = 2 * ((size_t)tile.slot * kWaves + (threadIdx.x >> 6));
      u.partials[i] = acc;
      u.partials[i + 1] = vol;
    }
  });
}

// the driver step's close (apps/maire_hydro/driver.h:514-515): t += dt,
// one step more, and the next step's flag (k_maire_begin's)
__device__ inline void close_step(DevScalars &S, const ProbParams &P) {
  if (S.active && !S.err_code) {
    S.t += S.dt;
    S.steps = S.steps + 1;
  }
  S.active = (S.err_code == 0) && (S.steps < P.max_steps) && (S.t < P.final_time);
}

// evaluate_time_step :179-188 (first minimum of the three limits) and the
// driver's cap min(dt, t_final - t) (driver.h:391), one block per problem.
__device__ inline void dt_finalize(const Batch &b, const double *partials) {
  const int nprob = *b.np;
  for (int ip = blockIdx.x; ip < nprob; ip += gridDim.x) {
    const int p = b.plist[ip];
    const ProbParams &P = b.pp[p];
    if (P.cell_count == 0) continue;
    if (!block_active(b.s, p)) continue;
    double mx[2];
    final_max<2>(partials + 2 * (size_t)P.tile_first * kWaves, P.tile_count * kWaves, mx);
    const double acc_inv = mx[0];
    const double vol_inv = mx[1];
    if (threadIdx.x != 0) continue;
    DevScalars &S = b.s[p];
    const double dts[3] = {P.cfl_acc / acc_inv, P.cfl_vol / vol_inv, P.cfl_growth * S.dt};
    int it = 0;
    for (int i = 1; i < 3; ++i)
      if (dts[i] < dts[it]) it = i;
    S.dt = smin(dts[it], P.final_time - S.t);
  }
}
__global__ void __launch_bounds__(kFinalBlock) k_maire_dt_finalize(Batch b, const double *partials) {
  dt_finalize(b, partials);
}

// apply_update's mesh motion x = x0 + dt u (:517-534).  SAVE (the driver
// step's predictor): save_coordinates first, x0 = x, and move from the value
// just saved, so the step needs no separate save launch.
template <bool SAVE, int N>
__global__ void __launch_bounds__(kBlock) k_maire_move(DevMesh m, MaireFields u, Batch b, double factor) {
  for_each_tile(b.vt, [&](const Tile &tile) {
    if (!wave_live(b.s, tile.prob, 0)) return;
    if ((int)threadIdx.x >= tile.count) return;
    const double delta_t = factor * b.s[tile.prob].dt;
    const int vt = tile.first + threadIdx.x;
    double x0[N];
#pragma unroll
    for (int d = 0; d < N; ++d) {
      if (SAVE) {
        x0[d] = m.x[d][vt];
        u.x0[d][vt] = x0[d];
      } else {
        x0[d] = u.x0[d][vt];
      }
    }
    if (!(delta_t > 1e-14)) return;   // do_step, flecsale::config::test_tolerance
    double x[N];
#pragma unroll
    for (int d = 0; d < N; ++d) {
      x[d] = x0[d] + delta_t * u.un[d][vt];
      m.x[d][vt] = x[d];
    }
    if constexpr (N == 3)
      put_xa(u, vt, x[0], x[1], x[2]);
    else
      put_xy(m, vt, x[0], x[1]);
  });
}

// MODE 0: the apply_update task (state from v, e); 1: the driver step's
// predictor, which first saves v0 = v, e0 = e (save_solution); 2: the
// corrector, which starts from v0, e0 (restore_solution) instead of copying
// them back first.
// MOVED (2D driver steps): the cell's vertices are moved here from the
// step's start, x0 + dt u as k_maire_move computes them, instead of read
// back after a move launch (k_maire_move_update moves them in the same
// launch).
// FS (a 2D driver step's predictor, FoldedStep): the step's flag and time
// step as the launch computed them, instead of the scalars it is writing.
// MOVED also leaves unwritten what no kernel of a driver step reads: the
// centroid, volume copy (V), density, pressure, temperature and velocity
// (the velocity is read from its packed copy); k_maire_settle writes them
// at the end of a batch of steps, from the same inputs.
struct FoldedStep {
  int active;
  double dt;
};
// RES (a 2D driver step's corrector, whose time step is known): the cell's
// residual computed here (cell_residual, as k_maire_residual computes and
// stores it) from this half step's nodal solve, instead of in a launch of
// its own.  The gate is the update's: before this launch only an earlier
// launch can have latched an error, which stops both.
template <int MODE, int N, bool MOVED = false, bool FS = false, bool RES = false>
__device__ __forceinline__ void update_tile(const DevMesh &m, const MaireFields &u, const Batch &b, double factor,
                                            int token, const Tile &tile, const FoldedStep &fs = FoldedStep{}) {
  static_assert(!RES || (N == 2 && MOVED), "the residual rides in 2D driver updates");
  {
    if constexpr (FS) {
      const int et = b.s[tile.prob].err_token;
      if (!fs.active || !(et == 0 || et == token)) return;
    } else {
      if (!wave_live(b.s, tile.prob, token)) return;
    }
    if ((int)threadIdx.x >= tile.count) return;
    const ProbParams &P = b.pp[tile.prob];
    const int c = tile.first + threadIdx.x;
    const double delta_t = factor * (FS ? fs.dt : b.s[tile.prob].dt);
    double rr[N + 2], ruv[4][2];
    int rvt[4];
    if constexpr (RES) {
      cell_residual<2, 4>(m, u, c, rr, rvt, ruv);
#pragma unroll
      for (int i = 0; i < N + 2; ++i) u.dudt[i][c] = rr[i];
    }
    double vol;
    if (delta_t > 1e-14) {
      // burton update_geometry() for the moved cell
      if constexpr (N == 2) {
        double px[4], py[4];
        for (int i = 0; i < 4; ++i) {
          const int vt = RES ? rvt[i] : m.cv[(size_t)i * m.cap_c + c];
          if constexpr (RES) {
            const double2 w = u.x0p[vt];
            px[i] = w.x + delta_t * ruv[i][0];
            py[i] = w.y + delta_t * ruv[i][1];
          } else if constexpr (MOVED) {
            const double2 w = u.x0p[vt];
            px[i] = w.x + delta_t * u.un[0][vt];
            py[i] = w.y + delta_t * u.un[1][vt];
          } else {
            const double2 w = m.xy[vt];
            px[i] = w.x;
            py[i] = w.y;
          }
        }
        double cx, cy, ml;
        cell_geom_2d(4, px, py, cx, cy, vol, ml);
        if (!MOVED) {
          m.ccen[0][c] = cx;
          m.ccen[1][c] = cy;
        }
        m.cvol[c] = vol;
        m.cminlen[c] = ml;
      } else {
        double p[8][3], cen[3], ml;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int v = m.cv[(size_t)i * m.cap_c + c];
#pragma unroll
          for (int d = 0; d < 3; ++d) p[i][d] = m.x[d][v];
        }
        hex_geom(p, cen, vol, ml);
#pragma unroll
        for (int d = 0; d < 3; ++d) m.ccen[d][c] = cen[d];
        m.cvol[c] = vol;
        m.cminlen[c] = ml;
      }
    } else {
      vol = m.cvol[c];
    }
    // update_state_from_flux lagrange_eqns.h:310-343
    const double M = u.M[c];
    const double inv_mass = 1 / M;
    double vel[N], e0;
    if constexpr (MOVED && MODE == 1) {
      const double2 w = u.cs2[1][c];
      vel[0] = w.x;
      vel[1] = w.y;
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i) vel[i] = MODE == 2 ? u.v0[i][c] : u.v[i][c];
    }
    e0 = MODE == 2 ? u.e0[c] : u.e[c];
    if (MODE == 1) {
#pragma unroll
      for (int i = 0; i < N; ++i) u.v0[i][c] = vel[i];
      u.e0[c] = e0;
    }
    double rate[N + 1];   // energy, then momentum
#pragma unroll
    for (int i = 0; i <= N; ++i) rate[i] = RES ? rr[i == 0 ? N + 1 : i] : u.dudt[i == 0 ? N + 1 : i][c];
    const double et = e0 + dotN<N>(vel, vel) / 2 + delta_t * inv_mass * rate[0];
#pragma unroll
    for (int i = 0; i < N; ++i) vel[i] += delta_t * inv_mass * rate[1 + i];
    const double ie = et - dotN<N>(vel, vel) / 2;
    // update_volume :352-367
    const double d = M / vol;
    if (!MOVED) {
#pragma unroll
      for (int i = 0; i < N; ++i) u.v[i][c] = vel[i];
      u.V[c] = vol;
      u.d[c] = d;
    }
    u.e[c] = ie;
    // update_state_from_energy lagrange_eqns.h:280-302
    const double g = u.g[c];
    const double p = eos_p_de(g, d, ie);
    const double a = smax(eos_a_de(g, ie), kMinSoundSpeed);
    if (!MOVED) {
      u.p[c] = p;
      u.T[c] = eos_T_de(P.cv, ie);
    }
    u.a[c] = a;
    if constexpr (N == 3)
      put_cs(u, c, d, a, p, vel[0], vel[1], vel[2]);
    else
      put_cs2(u, c, d, a, p, vel[0], vel[1]);
    if (ie < 0 || d < 0) flag_negative(b.s, tile.prob, m.corig[c], token);
  }
}

template <int MODE, int N>
__global__ void __launch_bounds__(kBlock) k_maire_update(DevMesh m, MaireFields u, Batch b, double factor,
                                                         int token) {
  for_each_tile(b.ct, [&](const Tile &tile) { update_tile<MODE, N>(m, u, b, factor, token, tile); });
}

// A 2D driver step's apply_update in one launch: the mesh move x = x0 + dt u
// (:517-534) from the step's start coordinates saved by its first nodal
// solve (x0p), and the cells' update (update_tile MOVED), which takes its
// vertices' moved positions from the same values, so no block waits on
// another.  MODE 1: the predictor, 2: the corrector.  The vertex tiles use
// the launch's token, as the update does: an error this launch raises does
// not stop its moves (a separate move launch would have completed first).
// FOLD (the predictor): the launch also closes the previous step
// (close_step) and finalizes the time step (dt_finalize), which would
// otherwise be a launch of their own before it.  Every block finalizes the time
// step of its tiles' problems itself from the per-tile partials and the
// step recorded by the residual launch (DevScalars::snap_*, k_maire_residual
// TILE), which no block of this launch writes; the block of each problem's
// first cell tile writes the closed step and the new dt back to the
// problem's scalars, which no block of this launch reads.
__device__ inline FoldedStep folded_step(const Batch &b, const double *tp, int prob) {
  const DevScalars &S = b.s[prob];
  const ProbParams &P = b.pp[prob];
  FoldedStep fs{S.snap_active, 0.0};
  // max over the problem's tile partials (block-wide; max is exact, so the
  // order does not matter), four loads in flight per thread
  __shared__ double red[2][kWaves];
  const double2 *pp = reinterpret_cast<const double2 *>(tp) + P.tile_first;
  double2 a4[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  int i = threadIdx.x;
  for (; i + 3 * kBlock < P.tile_count; i += 4 * kBlock) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double2 w = pp[i + k * kBlock];
      a4[k].x = smax(w.x, a4[k].x);
      a4[k].y = smax(w.y, a4[k].y);
    }
  }
  for (; i < P.tile_count; i += kBlock) {
    const double2 w = pp[i];
    a4[0].x = smax(w.x, a4[0].x);
    a4[0].y = smax(w.y, a4[0].y);
  }
  double mx[2] = {smax(smax(a4[0].x, a4[1].x), smax(a4[2].x, a4[3].x)),
                  smax(smax(a4[0].y, a4[1].y), smax(a4[2].y, a4[3].y))};
  __syncthreads();   // red is free (the block's previous call)
  for (int j = 0; j < 2; ++j) {
    mx[j] = wave_max(mx[j]);
    if ((threadIdx.x & 63) == 0) red[j][threadIdx.x >> 6] = mx[j];
  }
  __syncthreads();
  for (int j = 0; j < 2; ++j) {
    mx[j] = red[j][0];
    for (int w = 1; w < kWaves; ++w) mx[j] = smax(red[j][w], mx[j]);
  }
  // evaluate_time_step :179-188 and the driver's cap (driver.h:391), as dt_finalize
  const double dts[3] = {P.cfl_acc / mx[0], P.cfl_vol / mx[1], P.cfl_growth * S.snap_dt};
  int it = 0;
  for (int i = 1; i < 3; ++i)
    if (dts[i] < dts[it]) it = i;
  fs.dt = smin(dts[it], P.final_time - S.snap_t);
  return fs;
}

template <int MODE, bool FOLD = false, bool RES = false>
__global__ void __launch_bounds__(kBlock) k_maire_move_update(DevMesh m, MaireFields u, Batch b, double factor,
                                                              int token) {
  // FOLD: the block's last problem and its folded step (a block's vertex
  // and cell tiles are mostly of one problem: one reduction for both)
  int fprob = -1;
  FoldedStep fcache{0, 0.0};
  auto folded = [&](int prob) {
    if (prob != fprob) {
      fcache = folded_step(b, u.partials, prob);
      fprob = prob;
    }
    return fcache;
  };
  for_each_tile(b.vt, [&](const Tile &tile) {
    FoldedStep fs{0, 0.0};
    if constexpr (FOLD) {
      fs = folded(tile.prob);
      const int et = b.s[tile.prob].err_token;
      if (!fs.active || !(et == 0 || et == token)) return;
    } else {
      if (!wave_live(b.s, tile.prob, token)) return;
    }
    if ((int)threadIdx.x >= tile.count) return;
    const double delta_t = factor * (FOLD ? fs.dt : b.s[tile.prob].dt);
    if (!(delta_t > 1e-14)) return;   // do_step, flecsale::config::test_tolerance
    const int vt = tile.first + threadIdx.x;
    const double2 w = u.x0p[vt];
    const double x = w.x + delta_t * u.un[0][vt], y = w.y + delta_t * u.un[1][vt];
    put_xy(m, vt, x, y);   // m.x follows in k_maire_settle
  });
  for_each_tile(b.ct, [&](const Tile &tile) {
    if constexpr (FOLD) {
      const FoldedStep fs = folded(tile.prob);
      if (threadIdx.x == 0 && tile.first == b.pp[tile.prob].cell_first) {
        DevScalars &S = b.s[tile.prob];
        S.t = S.snap_t;
        S.steps = S.snap_steps;
        S.active = S.snap_active;
        if (fs.active) {
          S.dt = fs.dt;
          S.lazy = 1;
        }
      }
      update_tile<MODE, 2, true, true, RES>(m, u, b, factor, token, tile, fs);
    } else {
      update_tile<MODE, 2, true, false, RES>(m, u, b, factor, token, tile);
    }
  });
}

// The end of a batch of 2D driver steps: for every problem whose steps ran
// (DevScalars::lazy), what the steps' updates left unwritten (update_tile
// MOVED), as they would have written it: the coordinates and the last
// step's saved start coordinates (x0, as save_coordinates) from their pairs,
// the centroids by cell_geom_2d from the coordinates, volume, density,
// pressure and temperature from the cell volume, mass and internal energy,
// and the velocity from its packed copy.  The same functions of the same
// values: the same bits.
__global__ void __launch_bounds__(kBlock) k_maire_settle(DevMesh m, MaireFields u, Batch b) {
  for_each_tile(b.vt, [&](const Tile &tile) {
    if (!b.s[tile.prob].lazy || (int)threadIdx.x >= tile.count) return;
    const int v = tile.first + threadIdx.x;
    const double2 w = m.xy[v], w0 = u.x0p[v];
    m.x[0][v] = w.x;
    m.x[1][v] = w.y;
    u.x0[0][v] = w0.x;
    u.x0[1][v] = w0.y;
  });
  for_each_tile(b.ct, [&](const Tile &tile) {
    if (!b.s[tile.prob].lazy || (int)threadIdx.x >= tile.count) return;
    const ProbParams &P = b.pp[tile.prob];
    const int c = tile.first + threadIdx.x;
    double px[4], py[4];
    for (int i = 0; i < 4; ++i) {
      const double2 w = m.xy[m.cv[(size_t)i * m.cap_c + c]];
      px[i] = w.x;
      py[i] = w.y;
    }
    double cx, cy, vol, ml;
    cell_geom_2d(4, px, py, cx, cy, vol, ml);
    m.ccen[0][c] = cx;
    m.ccen[1][c] = cy;
    const double V = m.cvol[c], ie = u.e[c];
    const double d = u.M[c] / V;
    u.V[c] = V;
    u.d[c] = d;
    u.p[c] = eos_p_de(u.g[c], d, ie);
    u.T[c] = eos_T_de(P.cv, ie);
    const double2 vc = u.cs2[1][c];
    u.v[0][c] = vc.x;
    u.v[1][c] = vc.y;
  });
}

// save (dir 0) / restore (dir 1) of coordinates, velocity and energy
template <int N>
__global__ void __launch_bounds__(kBlock) k_maire_copy(DevMesh m, MaireFields u, Batch b, int dir) {
  for_each_tile(b.vt, [&](const Tile &tile) {
    if (!wave_live(b.s, tile.prob, 0)) return;
    if ((int)threadIdx.x >= tile.count) return;
    const int i = tile.first + threadIdx.x;
    for (int d = 0; d < N; ++d) {
      if (dir == 0)
        u.x0[d][i] = m.x[d][i];
      else
        m.x[d][i] = u.x0[d][i];
    }
    if (dir == 1) {
      if constexpr (N == 3)
        put_xa(u, i, m.x[0][i], m.x[1][i], m.x[2][i]);
      else
        put_xy(m, i, m.x[0][i], m.x[1][i]);
    }
  });
  for_each_tile(b.ct, [&](const Tile &tile) {
    if (!wave_live(b.s, tile.prob, 0)) return;
    if ((int)threadIdx.x >= tile.count) return;
    const int i = tile.first + threadIdx.x;
    if (dir == 0) {
      for (int d = 0; d < N; ++d) u.v0[d][i] = u.v[d][i];
      u.e0[i] = u.e[i];
    } else {
      for (int d = 0; d < N; ++d) u.v[d][i] = u.v0[d][i];
      u.e[i] = u.e0[i];
      if constexpr (N == 3)
        put_cs(u, i, u.d[i], u.a[i], u.p[i], u.v[0][i], u.v[1][i], u.v[2][i]);
      else
        put_cs2(u, i, u.d[i], u.a[i], u.p[i], u.v[0][i], u.v[1][i]);
    }
  });
}

// step bracket of the driver loop (apps/maire_hydro/driver.h:332,514-515)
__global__ void k_maire_begin(Batch b) {
  const int nprob = *b.np;
  for (int ip = blockIdx.x * blockDim.x + threadIdx.x; ip < nprob; ip += gridDim.x * blockDim.x) {
    const int p = b.plist[ip];
    const ProbParams &P = b.pp[p];
    if (P.cell_count == 0) continue;
    DevScalars &S = b.s[p];
    S.active = (S.err_code == 0) && (S.steps < P.max_steps) && (S.t < P.final_time);
  }
}
// ... which also opens the next step as k_maire_begin would, so a step that
// follows another in one batch of steps needs no begin launch
__global__ void k_maire_end(Batch b) {
  const int nprob = *b.np;
  for (int ip = blockIdx.x * blockDim.x + threadIdx.x; ip < nprob; ip += gridDim.x * blockDim.x) {
    const int p = b.plist[ip];
    const ProbParams &P = b.pp[p];
    if (P.cell_count == 0) continue;
    close_step(b.s[p], P);
    b.s[p].lazy = 0;
  }
}

template <class K, class... A>
int launch_persistent(fsale_ctx *ctx, int kid, K kernel, A... args) {
  ProfScope prof(ctx, kid);
  hipLaunchKernelGGL(kernel, dim3(persistent_grid(ctx, kid)), dim3(kBlock), 0, ctx->stream, args...);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int small_grid(int n) {
  const int g = (n + kBlock - 1) / kBlock;
  return g < 1 ? 1 : (g > 1024 ? 1024 : g);
}

bool check_ctx(fsale_ctx *ctx) {
  if (!ctx || ctx->solver != FSALE_SOLVER_MAIRE || !ctx->single || ctx->probs.empty() || !ctx->probs[0].live) {
    set_message("not a Lagrangian context");
    return false;
  }
  return true;
}

// evaluate_nodal_state: in 2D one launch for both kinds of vertex when both
// use one block per tile (timed as the nodal kernel), else interior vertices
// (Cramer) and then the boundary vertices' symmetry solves.  3D keeps two
// launches: its boundary solve (QR up to 6x6) would cap the interior waves'
// registers.
template <int N, bool SAVE = false, bool PEND = false>
int launch_nodal_n(fsale_ctx *ctx) {
  Batch bi = batch_args(ctx, K_MAIRE_NODAL);
  const Batch bb = batch_args(ctx, K_MAIRE_NODAL_BND);
  if constexpr (N == 3) {
    int rc = launch_persistent(ctx, K_MAIRE_NODAL, k_maire_nodal<false, 3>, ctx->m, ctx->ma, bi);
    if (rc) return rc;
    ProfScope prof(ctx, K_MAIRE_NODAL_BND);
    const int gb = persistent_grid(ctx, K_MAIRE_NODAL_BND);
    hipLaunchKernelGGL((k_maire_nodal<true, 3>), dim3(gb), dim3(kBlock), 0, ctx->stream, ctx->m, ctx->ma, bb);
    FSALE_HIP(hipGetLastError());
    // the solves with two or three symmetry rows: only the vertices on two
    // or three tagged faces can have them, a short tail of each problem's
    // list, one block per tile
    Batch bg = bb;
    bg.bt.tiles = ctx->d_bbtiles;
    bg.bt.n = ctx->d_counts + 4;
    bg.bt.direct = nullptr;
    bg.bt.ndirect = 0;
    const int gg = std::max(8, (ctx->n_bbtiles + 7) / 8 * 8);
    hipLaunchKernelGGL(k_maire3_nodal_bnd_big, dim3(gg), dim3(kBlock), 0, ctx->stream, ctx->m, ctx->ma, bg);
    FSALE_HIP(hipGetLastError());
    return FSALE_OK;
  }
  if (bi.vt.direct && bb.bt.direct) {
    bi.bt = bb.bt;
    const int nb = (bb.bt.ndirect + 7) / 8 * 8;
    ProfScope prof(ctx, K_MAIRE_NODAL);
    hipLaunchKernelGGL((k_maire_nodal_both<2, SAVE, PEND>), dim3(nb + bi.vt.ndirect), dim3(kBlock), 0, ctx->stream,
                       ctx->m, ctx->ma, bi, nb);
    FSALE_HIP(hipGetLastError());
    return FSALE_OK;
  }
  int rc = launch_persistent(ctx, K_MAIRE_NODAL, k_maire_nodal<false, N, SAVE, PEND>, ctx->m, ctx->ma,
                             batch_args(ctx, K_MAIRE_NODAL));
  return rc ? rc
            : launch_persistent(ctx, K_MAIRE_NODAL_BND, k_maire_nodal<true, N, SAVE, PEND>, ctx->m, ctx->ma,
                                batch_args(ctx, K_MAIRE_NODAL_BND));
}
int launch_nodal(fsale_ctx *ctx) { return ctx->m.dim == 3 ? launch_nodal_n<3>(ctx) : launch_nodal_n<2>(ctx); }
int launch_residual(fsale_ctx *ctx, bool with_dt = false) {
  const Batch b = batch_args(ctx, K_MAIRE_RESIDUAL);
  if (with_dt)
    return ctx->m.dim == 3
               ? launch_persistent(ctx, K_MAIRE_RESIDUAL, k_maire_residual<3, true>, ctx->m, ctx->ma, b)
               : launch_persistent(ctx, K_MAIRE_RESIDUAL, k_maire_residual<2, true>, ctx->m, ctx->ma, b);
  return ctx->m.dim == 3 ? launch_persistent(ctx, K_MAIRE_RESIDUAL, k_maire_residual<3, false>, ctx->m, ctx->ma, b)
                         : launch_persistent(ctx, K_MAIRE_RESIDUAL, k_maire_residual<2, false>, ctx->m, ctx->ma, b);
}
int launch_dt_finalize(fsale_ctx *ctx) {
  ProfScope prof(ctx, K_MAIRE_DT);
  hipLaunchKernelGGL(k_maire_dt_finalize, dim3(ctx->pgrid), dim3(kFinalBlock), 0, ctx->stream, batch_args(ctx, -1),
                     (const double *)ctx->ma.partials);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}
int launch_dt(fsale_ctx *ctx) {
  int rc = launch_persistent(ctx, K_MAIRE_DT, k_maire_dt_partial, ctx->m, ctx->ma, batch_args(ctx, K_MAIRE_DT));
  return rc ? rc : launch_dt_finalize(ctx);
}
// token: the update launch's id for error attribution (predictor 1,
// corrector 2, task-level apply_update 3)
// stage 0: the apply_update task; 1: the driver step's predictor (saves
// coordinates and solution as it goes); 2: its corrector (from the saved ones)
template <int N>
int launch_apply_n(fsale_ctx *ctx, double factor, int token, int stage) {
  const Batch bm = batch_args(ctx, K_MAIRE_MOVE), bu = batch_args(ctx, K_MAIRE_UPDATE);
  int rc = stage == 1 ? launch_persistent(ctx, K_MAIRE_MOVE, k_maire_move<true, N>, ctx->m, ctx->ma, bm, factor)
                      : launch_persistent(ctx, K_MAIRE_MOVE, k_maire_move<false, N>, ctx->m, ctx->ma, bm, factor);
  if (rc) return rc;
  if (stage == 1)
    return launch_persistent(ctx, K_MAIRE_UPDATE, k_maire_update<1, N>, ctx->m, ctx->ma, bu, factor, token);
  if (stage == 2)
    return launch_persistent(ctx, K_MAIRE_UPDATE, k_maire_update<2, N>, ctx->m, ctx->ma, bu, factor, token);
  return launch_persistent(ctx, K_MAIRE_UPDATE, k_maire_update<0, N>, ctx->m, ctx->ma, bu, factor, token);
}
int launch_apply(fsale_ctx *ctx, double factor, int token, int stage) {
  return ctx->m.dim == 3 ? launch_apply_n<3>(ctx, factor, token, stage) : launch_apply_n<2>(ctx, factor, token, stage);
}
// a 2D driver step's corrector: residual, move and update (token 2)
int launch_corrector_update(fsale_ctx *ctx) {
  return launch_persistent(ctx, K_MAIRE_UPDATE, k_maire_move_update<2, false, true>, ctx->m, ctx->ma,
                           batch_args(ctx, K_MAIRE_UPDATE), 1.0, 2);
}
int launch_copy(fsale_ctx *ctx, int dir) {
  const Batch b = batch_args(ctx, K_MAIRE_COPY);
  return ctx->m.dim == 3 ? launch_persistent(ctx, K_MAIRE_COPY, k_maire_copy<3>, ctx->m, ctx->ma, b, dir)
                         : launch_persistent(ctx, K_MAIRE_COPY, k_maire_copy<2>, ctx->m, ctx->ma, b, dir);
}

}  // namespace

// One driver step (apps/maire_hydro/driver.h:330-538) of every live problem
// (`first` / `last`: the first / last step of a batch of steps):
// predictor half step, then the corrector from the saved state.  The
// driver's save_coordinates / save_solution happen inside the predictor's
// move and update (each saves what it reads before writing), and
// restore_solution / restore_coordinates are not needed: the corrector's
// update reads the saved v0, e0, and its move writes x = x0 + dt u (when it
// does not move, dt <= 1e-14, the predictor did not move either and x is
// still x0).  `first`: the first step of a batch of steps, which opens its
// step with k_maire_begin (later ones are opened by the previous k_maire_end).
int maire_step(fsale_ctx *ctx, bool first, bool last) {
  int rc;
  const int np = (int)ctx->probs.size();
  if (first) {
    hipLaunchKernelGGL(k_maire_begin, dim3(small_grid(np)), dim3(kBlock), 0, ctx->stream, batch_args(ctx, -1));
    FSALE_HIP(hipGetLastError());
  }
  if (ctx->m.dim == 2) {
    // the first nodal solve saves the step's start coordinates, the moves
    // ride in the updates, the corrector's residual in its update, and the
    // close of a step runs with the next step's time step in its
    // predictor's update, or after the last step of the batch of steps
    // (k_maire_settle, k_maire_end): five launches per step
    if ((rc = first ? launch_nodal_n<2, true, false>(ctx) : launch_nodal_n<2, true, true>(ctx))) return rc;
    {
      const Batch b = batch_args(ctx, K_MAIRE_RESIDUAL);
      rc = first ? launch_persistent(ctx, K_MAIRE_RESIDUAL, k_maire_residual<2, true, false, true>, ctx->m, ctx->ma,
                                     b)
                 : launch_persistent(ctx, K_MAIRE_RESIDUAL, k_maire_residual<2, true, true, true>, ctx->m, ctx->ma,
                                     b);
      if (rc) return rc;
    }
    {
      // the time step's finalize (and the previous step's close) ride in
      // the predictor's update
      const Batch b = batch_args(ctx, K_MAIRE_UPDATE);
      if ((rc = launch_persistent(ctx, K_MAIRE_UPDATE, k_maire_move_update<1, true>, ctx->m, ctx->ma, b, 0.5, 1)))
        return rc;
    }
    // the corrector's residual rides in its update: its time step is known
    if ((rc = launch_nodal(ctx))) return rc;
    if ((rc = launch_corrector_update(ctx))) return rc;
    if (last) {
      if ((rc = launch_persistent(ctx, K_MAIRE_UPDATE, k_maire_settle, ctx->m, ctx->ma,
                                  batch_args(ctx, K_MAIRE_UPDATE))))
        return rc;
      hipLaunchKernelGGL(k_maire_end, dim3(small_grid(np)), dim3(kBlock), 0, ctx->stream, batch_args(ctx, -1));
      FSALE_HIP(hipGetLastError());
    }
    return FSALE_OK;
  }
  if ((rc = launch_nodal(ctx))) return rc;
  if ((rc = launch_residual(ctx, true))) return rc;
  if ((rc = launch_dt_finalize(ctx))) return rc;
  if ((rc = launch_apply(ctx, 0.5, 1, 1))) return rc;
  if ((rc = launch_nodal(ctx))) return rc;
  if ((rc = launch_residual(ctx))) return rc;
  if ((rc = launch_apply(ctx, 1.0, 2, 2))) return rc;
  hipLaunchKernelGGL(k_maire_end, dim3(small_grid(np)), dim3(kBlock), 0, ctx->stream, batch_args(ctx, -1));
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int maire_alloc(fsale_ctx *ctx) {
  const DevMesh &m = ctx->m;
  MaireFields &u = ctx->ma;
  int rc;
  double **cell1[] = {&u.V, &u.M, &u.p, &u.d, &u.e, &u.T, &u.a, &u.g, &u.e0};
  for (double **p : cell1)
    if ((rc = dev_alloc_n(ctx, p, m.cap_c))) return rc;
  for (int d = 0; d < m.dim; ++d) {
    if ((rc = dev_alloc_n(ctx, &u.v[d], m.cap_c))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.v0[d], m.cap_c))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.x0[d], m.cap_v))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.un[d], m.cap_v))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.npc[d], (size_t)m.cap_c * m.vpc))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.Fpc[d], (size_t)m.cap_c * m.vpc))) return rc;
  }
  for (int i = 0; i < m.dim + 2; ++i)
    if ((rc = dev_alloc_n(ctx, &u.dudt[i], m.cap_c))) return rc;
  if ((rc = dev_alloc_n(ctx, &u.partials, 2 * (size_t)kWaves * (ctx->ra_ct.capacity + 1)))) return rc;
  if (m.dim == 2) {
    for (int i = 0; i < 2; ++i)
      if ((rc = dev_alloc_n(ctx, &u.cs2[i], m.cap_c))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.x0p, m.cap_v))) return rc;
  }
  if (m.dim == 3) {
    if ((rc = dev_alloc_n(ctx, &u.cv8, (size_t)m.cap_c * 8 + 8))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.xa, (size_t)m.cap_v * 4 + 8))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.cs, (size_t)m.cap_c * 6 + 8))) return rc;
  }
  return FSALE_OK;
}

int maire_pack_problem(fsale_ctx *ctx, int prob, hipStream_t stream) {
  const Problem &P = ctx->probs[prob];
  if (ctx->solver != FSALE_SOLVER_MAIRE) return FSALE_OK;
  const int n = P.nc > P.nv ? P.nc : P.nv;
  if (n == 0) return FSALE_OK;
  const dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
  if (ctx->m.dim == 2)
    hipLaunchKernelGGL(k_maire2_pack, grid, block, 0, stream, ctx->m, ctx->ma, P.c0, P.nc, P.v0, P.nv);
  else
    hipLaunchKernelGGL(k_maire3_pack, grid, block, 0, stream, ctx->m, ctx->ma, P.c0, P.nc, P.v0, P.nv);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int maire_initial_state(fsale_ctx *ctx, int prob, hipStream_t stream) {
  const Problem &P = ctx->probs[prob];
  const int g = (P.nc + kBlock - 1) / kBlock;
  if (g == 0) return FSALE_OK;
  hipLaunchKernelGGL(k_maire_ic, dim3(g), dim3(kBlock), 0, stream, ctx->m, ctx->ma, (const ProbParams *)ctx->d_params,
                     prob);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

// records a boundary condition in the problem's host parameters (the
// caller uploads them)
int set_maire_boundary(fsale_ctx *ctx, int prob, int tag, int kind, const double *value) {
  const int dim = ctx->m.dim;
  if (tag < 0 || tag >= dim || kind < FSALE_BC_SYMMETRY || kind > FSALE_BC_NONE) {
    set_message("boundary condition: tag 0..dim-1, kind FSALE_BC_*");
    return FSALE_ERR_ARG;
  }
  const int nval = kind == FSALE_BC_PRESSURE ? 1 : kind == FSALE_BC_VELOCITY ? dim : 0;
  if (nval && !value) return FSALE_ERR_ARG;
  for (int i = 0; i < nval; ++i)
    if (!std::isfinite(value[i])) {
      set_message("boundary condition value is not finite");
      return FSALE_ERR_ARG;
    }
  ProbParams &P = ctx->probs[prob].params;
  P.bc_kind[tag] = kind;
  for (int d = 0; d < 3; ++d) P.bc_val[tag][d] = d < nval ? value[d] : 0.0;
  return FSALE_OK;
}

int set_problem_field(fsale_ctx *ctx, int prob, int field, const double *host_in);
int set_problem_state(fsale_ctx *ctx, int prob, const double *d, const double *v, const double *p);
int reset_scalars(fsale_ctx *ctx, int prob, double t, double dt);
int batch_steps(fsale_ctx *ctx, int nsteps, int use_graph);

}  // namespace fsale

using namespace fsale;

extern "C" {

int fsale_maire_initial_conditions(fsale_ctx *ctx, double cv, const double *gamma, const double *d,
                                   const double *v, const double *p) {
  if (!check_ctx(ctx) || !gamma || !d || !v || !p || !(cv > 0)) return FSALE_ERR_ARG;
  Problem &P = ctx->probs[0];
  for (int c = 0; c < P.nc; ++c)
    if (!(gamma[c] > 1)) return FSALE_ERR_ARG;
  P.params.cv = cv;
  int rc;
  if ((rc = upload_params(ctx, 0))) return rc;
  if ((rc = set_problem_field(ctx, 0, FSALE_F_GAMMA, gamma))) return rc;
  if ((rc = set_problem_state(ctx, 0, d, v, p))) return rc;
  if ((rc = maire_initial_state(ctx, 0, ctx->stream))) return rc;
  P.have_state = true;
  return reset_scalars(ctx, 0, 0.0, 0.0);
}

int fsale_maire_set_boundary(fsale_ctx *ctx, int tag, int kind, const double *value) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  int rc = set_maire_boundary(ctx, 0, tag, kind, value);
  return rc ? rc : upload_params(ctx, 0);
}

int fsale_maire_evaluate_nodal_state(fsale_ctx *ctx) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  if (!ctx->probs[0].have_state) return FSALE_ERR_STATE;
  int rc = sync_tiles(ctx);
  if (!rc) rc = open_tasks(ctx);
  return rc ? rc : launch_nodal(ctx);
}

int fsale_maire_evaluate_residual(fsale_ctx *ctx) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  if (!ctx->probs[0].have_state) return FSALE_ERR_STATE;
  int rc = sync_tiles(ctx);
  if (!rc) rc = open_tasks(ctx);
  return rc ? rc : launch_residual(ctx);
}

int fsale_maire_evaluate_time_step(fsale_ctx *ctx, double cfl_accoustic, double cfl_volume, double cfl_growth,
                                   double final_time, double *dt_out) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  Problem &P = ctx->probs[0];
  if (!P.have_state) return FSALE_ERR_STATE;
  P.params.cfl_acc = cfl_accoustic;
  P.params.cfl_vol = cfl_volume;
  P.params.cfl_growth = cfl_growth;
  P.params.final_time = final_time;
  int rc;
  if ((rc = upload_params(ctx, 0)) || (rc = sync_tiles(ctx)) || (rc = open_tasks(ctx)) || (rc = launch_dt(ctx)))
    return rc;
  if (!dt_out) return FSALE_OK;
  if ((rc = check_device_error(ctx, 0))) return rc;
  *dt_out = ctx->h_s[0].dt;
  return FSALE_OK;
}

int fsale_maire_apply_update(fsale_ctx *ctx, double factor) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  if (!ctx->probs[0].have_state) return FSALE_ERR_STATE;
  int rc = sync_tiles(ctx);
  if (!rc) rc = open_tasks(ctx);
  return rc ? rc : launch_apply(ctx, factor, 3, 0);
}

int fsale_maire_save(fsale_ctx *ctx) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  int rc = sync_tiles(ctx);
  if (!rc) rc = open_tasks(ctx);
  return rc ? rc : launch_copy(ctx, 0);
}

int fsale_maire_restore(fsale_ctx *ctx) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  int rc = sync_tiles(ctx);
  if (!rc) rc = open_tasks(ctx);
  return rc ? rc : launch_copy(ctx, 1);
}

int fsale_maire_advance(fsale_ctx *ctx, double cfl_accoustic, double cfl_volume, double cfl_growth,
                        double final_time, int nsteps, int use_graph) {
  if (!check_ctx(ctx) || nsteps < 0) return FSALE_ERR_ARG;
  Problem &P = ctx->probs[0];
  if (!P.have_state) return FSALE_ERR_STATE;
  if (nsteps == 0) return FSALE_OK;
  int rc;
  if ((rc = read_scalars(ctx))) return rc;
  P.params.cfl_acc = cfl_accoustic;
  P.params.cfl_vol = cfl_volume;
  P.params.cfl_growth = cfl_growth;
  P.params.final_time = final_time;
  P.params.max_steps = ctx->h_s[0].steps + nsteps;
  if ((rc = upload_params(ctx, 0))) return rc;
  return batch_steps(ctx, nsteps, use_graph);
}

}  // extern "C"
