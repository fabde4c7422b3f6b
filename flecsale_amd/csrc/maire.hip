// Cell-centred Lagrangian (Maire) step of apps/maire_hydro on the device, 2D
// quads.
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
// Corner arrays are k-major [4][cap_c] so the residual gather of a cell's
// four corners is coalesced.
#include "batch_dev.h"

namespace fsale {
namespace {

constexpr double kMinSoundSpeed = 1.e-6;   // flecsale/eqns/lagrange_eqns.h:54

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
  u.a[c] = smax(eos_a_de(g, ie), kMinSoundSpeed);
}

// wedge facets of corner (c, k): wedge 0 on edge v_{k-1} -> v_k, wedge 1 on
// edge v_k -> v_{k+1}; outward unit edge normal, facet_area = |edge
// midpoint - vertex| (burton wedge geometry; pinned by the sedov_2d golden).
__device__ inline void wedge_facets(const DevMesh &m, int c, int k, double n[2][2], double l[2]) {
  const int km = (k + 3) & 3, kp = (k + 1) & 3;
  const int im = m.cv[(size_t)km * m.cap_c + c], ik = m.cv[(size_t)k * m.cap_c + c],
            ip = m.cv[(size_t)kp * m.cap_c + c];
  const double xm = m.x[0][im], ym = m.x[1][im];
  const double xk = m.x[0][ik], yk = m.x[1][ik];
  const double xp = m.x[0][ip], yp = m.x[1][ip];
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
  const double zc = u.d[c] * u.a[c];
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

// Symmetry-constrained nodal velocity of boundary vertex vt
// (apps/maire_hydro/tasks.h:317-420): symmetry normals accumulated per tag
// over the vertex's boundary wedges in corner order, then the augmented system
// [M_p N^T; N 0] solved by flecsale::linalg::qr (:383-418); Cramer's rule
// when no wedge carries a symmetry tag.
constexpr int kNodalKeep = 4;

// keep: wedge normals, facet areas and tags of the first kNodalKeep corners,
// already loaded by the caller
struct KeptWedges {
  double n[kNodalKeep][2][2], l[kNodalKeep][2];
  int tags[kNodalKeep];
};

__device__ inline void boundary_solve(const DevMesh &m, int vt, int cnt, const double Mp[2][2], const double rhs[2],
                                      const KeptWedges &kw, double un[2]) {
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
        if (has[tag]) {
          for (int d = 0; d < 2; ++d) symn[tag][d] += l[w] * n[w][d];
        } else {
          has[tag] = 1;
          for (int d = 0; d < 2; ++d) symn[tag][d] = l[w] * n[w][d];
        }
      }
    }
  };
#pragma unroll
  for (int q = 0; q < kNodalKeep; ++q)
    if (q < cnt && (kw.tags[q] & 0x44)) add_wedges(kw.tags[q], kw.n[q], kw.l[q]);
  for (int q = kNodalKeep; q < cnt; ++q) {
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
    return;
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
}

// evaluate_nodal_state, one work-item per vertex.  The vertex's corners
// (ascending cell id, k-major table) are visited in order; the first four
// corner matrices stay in registers for the final F_pc -= M_pc u_p (a vertex
// with more corners recomputes the rest), and the first four corners' loads
// do not depend on each other.

__device__ inline void corner_force(const MaireFields &u, int c, const double M[2][2], const double np[2],
                                    double F[2]) {
  const double uc[2] = {u.v[0][c], u.v[1][c]};
  const double pc = u.p[c];
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
template <bool BOUNDARY>
__device__ inline void nodal_tile(const DevMesh &m, const MaireFields &u, const Batch &b, const Tile &tile) {
  {
    if (!wave_live(b.s, tile.prob, 0)) return;
    if ((int)threadIdx.x >= tile.count) return;
    int vt = tile.first + threadIdx.x;
    if constexpr (BOUNDARY)
      vt = m.bv[vt];
    else if (m.vbnd[vt])
      return;
    const int cnt = m.vc_cnt[vt];
    double M[kNodalKeep][2][2], np[kNodalKeep][2], F[kNodalKeep][2];
    int cc[kNodalKeep], kk[kNodalKeep];
    KeptWedges kw;
#pragma unroll
    for (int q = 0; q < kNodalKeep; ++q) {
      const int cn = q < cnt ? m.vck[(size_t)q * m.cap_v + vt] : -1;
      cc[q] = cn >> 2;
      kk[q] = cn & 3;
      if (cn >= 0) {
        if constexpr (BOUNDARY) {
          wedge_facets(m, cc[q], kk[q], kw.n[q], kw.l[q]);
          kw.tags[q] = m.cwtag[(size_t)kk[q] * m.cap_c + cc[q]];
          corner_matrix_from(u, cc[q], kw.n[q], kw.l[q], M[q], np[q]);
        } else {
          corner_matrix(m, u, cc[q], kk[q], M[q], np[q]);
        }
        corner_force(u, cc[q], M[q], np[q], F[q]);
      }
    }
    // M_p = sum_c M_pc and rhs = sum_c F_pc in corner order
    double Mp[2][2] = {{0, 0}, {0, 0}}, rhs[2] = {0, 0};
#pragma unroll
    for (int q = 0; q < kNodalKeep; ++q) {
      if (q >= cnt) break;
      for (int r = 0; r < 2; ++r) {
        Mp[r][0] += M[q][r][0];
        Mp[r][1] += M[q][r][1];
        rhs[r] += F[q][r];
      }
    }
    for (int q = kNodalKeep; q < cnt; ++q) {
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
    if constexpr (BOUNDARY) {
      boundary_solve(m, vt, cnt, Mp, rhs, kw, un);
    } else {
      // ristra::math::solve (un-vendored): Cramer's rule
      const double det = Mp[0][0] * Mp[1][1] - Mp[0][1] * Mp[1][0];
      un[0] = (rhs[0] * Mp[1][1] - Mp[0][1] * rhs[1]) / det;
      un[1] = (Mp[0][0] * rhs[1] - rhs[0] * Mp[1][0]) / det;
    }
    u.un[0][vt] = un[0];
    u.un[1][vt] = un[1];
    // F_pc -= M_pc u_p  (matrix_vector(-1, Mpc, un, 1, Fpc), tasks.h:435-448)
    auto finish = [&](int c, int k, const double Mq[2][2], const double npq[2], const double Fq[2]) {
      const size_t ci = (size_t)k * m.cap_c + c;
      u.npc[0][ci] = npq[0];
      u.npc[1][ci] = npq[1];
      for (int r = 0; r < 2; ++r) {
        double t = Mq[r][0] * un[0];
        t += Mq[r][1] * un[1];
        u.Fpc[r][ci] = -1.0 * t + 1.0 * Fq[r];
      }
    };
#pragma unroll
    for (int q = 0; q < kNodalKeep; ++q)
      if (q < cnt) finish(cc[q], kk[q], M[q], np[q], F[q]);
    for (int q = kNodalKeep; q < cnt; ++q) {
      const int cn = m.vck[(size_t)q * m.cap_v + vt];
      double Mq[2][2], npq[2], Fq[2];
      corner_matrix(m, u, cn >> 2, cn & 3, Mq, npq);
      corner_force(u, cn >> 2, Mq, npq, Fq);
      finish(cn >> 2, cn & 3, Mq, npq, Fq);
    }
  }
}

template <bool BOUNDARY>
__global__ void __launch_bounds__(kBlock) k_maire_nodal(DevMesh m, MaireFields u, Batch b) {
  for_each_tile(BOUNDARY ? b.bt : b.vt, [&](const Tile &tile) { nodal_tile<BOUNDARY>(m, u, b, tile); });
}

// Both kinds in one launch over the one-block-per-tile lists: blocks
// [0, nb) take the boundary tiles (nb a multiple of 8, so the interior list
// keeps its XCD arrangement), the rest the interior tiles.  The boundary
// vertices' solves are long serial chains in few waves; issued first, they
// run while the interior waves fill the machine around them instead of
// forming a launch of their own.  Held at 3 waves per SIMD for the interior
// tiles (the boundary path then keeps part of its state in scratch).
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_maire_nodal_both(DevMesh m, MaireFields u, Batch b, int nb) {
  if ((int)blockIdx.x < nb) {
    if ((int)blockIdx.x < b.bt.ndirect) {
      const Tile t = b.bt.direct[blockIdx.x];
      if (t.count > 0) nodal_tile<true>(m, u, b, t);
    }
  } else {
    const int i = (int)blockIdx.x - nb;
    if (i < b.vt.ndirect) {
      const Tile t = b.vt.direct[i];
      if (t.count > 0) nodal_tile<false>(m, u, b, t);
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_maire_residual(DevMesh m, MaireFields u, Batch b) {
  for_each_tile(b.ct, [&](const Tile &tile) {
    if (!wave_live(b.s, tile.prob, 0)) return;
    if ((int)threadIdx.x >= tile.count) return;
    const int c = tile.first + threadIdx.x;
    double r[4] = {0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const size_t ci = (size_t)k * m.cap_c + c;
      const int vt = m.cv[ci];
      const double uv[2] = {u.un[0][vt], u.un[1][vt]};
      const double F[2] = {u.Fpc[0][ci], u.Fpc[1][ci]};
      const double n[2] = {u.npc[0][ci], u.npc[1][ci]};
      // compute_update lagrange_eqns.h:374-389
      for (int d = 0; d < 2; ++d) {
        r[0] += n[d] * uv[d];
        r[3] -= F[d] * uv[d];
        r[1 + d] -= F[d];
      }
    }
    for (int i = 0; i < 4; ++i) u.dudt[i][c] = r[i];
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

// evaluate_time_step :179-188 (first minimum of the three limits) and the
// driver's cap min(dt, t_final - t) (driver.h:391), one block per problem
__global__ void __launch_bounds__(kFinalBlock) k_maire_dt_finalize(Batch b, const double *partials) {
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

// apply_update's mesh motion x = x0 + dt u (:517-534).  SAVE (the driver
// step's predictor): save_coordinates first, x0 = x, and move from the value
// just saved, so the step needs no separate save launch.
template <bool SAVE>
__global__ void __launch_bounds__(kBlock) k_maire_move(DevMesh m, MaireFields u, Batch b, double factor) {
  for_each_tile(b.vt, [&](const Tile &tile) {
    if (!wave_live(b.s, tile.prob, 0)) return;
    if ((int)threadIdx.x >= tile.count) return;
    const double delta_t = factor * b.s[tile.prob].dt;
    const int vt = tile.first + threadIdx.x;
    double x0[2];
    for (int d = 0; d < 2; ++d) {
      if (SAVE) {
        x0[d] = m.x[d][vt];
        u.x0[d][vt] = x0[d];
      } else {
        x0[d] = u.x0[d][vt];
      }
    }
    if (!(delta_t > 1e-14)) return;   // do_step, flecsale::config::test_tolerance
    for (int d = 0; d < 2; ++d) m.x[d][vt] = x0[d] + delta_t * u.un[d][vt];
  });
}

// MODE 0: the apply_update task (state from v, e); 1: the driver step's
// predictor, which first saves v0 = v, e0 = e (save_solution); 2: the
// corrector, which starts from v0, e0 (restore_solution) instead of copying
// them back first.
template <int MODE>
__global__ void __launch_bounds__(kBlock) k_maire_update(DevMesh m, MaireFields u, Batch b, double factor,
                                                         int token) {
  for_each_tile(b.ct, [&](const Tile &tile) {
    if (!wave_live(b.s, tile.prob, token)) return;
    if ((int)threadIdx.x >= tile.count) return;
    const ProbParams &P = b.pp[tile.prob];
    const int c = tile.first + threadIdx.x;
    const double delta_t = factor * b.s[tile.prob].dt;
    double vol;
    if (delta_t > 1e-14) {
      // burton update_geometry() for the moved cell
      double px[4], py[4];
      for (int i = 0; i < 4; ++i) {
        const int v = m.cv[(size_t)i * m.cap_c + c];
        px[i] = m.x[0][v];
        py[i] = m.x[1][v];
      }
      double cx, cy, ml;
      cell_geom_2d(4, px, py, cx, cy, vol, ml);
      m.ccen[0][c] = cx;
      m.ccen[1][c] = cy;
      m.cvol[c] = vol;
      m.cminlen[c] = ml;
    } else {
      vol = m.cvol[c];
    }
    // update_state_from_flux lagrange_eqns.h:310-343
    const double M = u.M[c];
    const double inv_mass = 1 / M;
    double vel[2], e0;
    if (MODE == 2) {
      vel[0] = u.v0[0][c];
      vel[1] = u.v0[1][c];
      e0 = u.e0[c];
    } else {
      vel[0] = u.v[0][c];
      vel[1] = u.v[1][c];
      e0 = u.e[c];
    }
    if (MODE == 1) {
      u.v0[0][c] = vel[0];
      u.v0[1][c] = vel[1];
      u.e0[c] = e0;
    }
    const double et = e0 + dotN<2>(vel, vel) / 2 + delta_t * inv_mass * u.dudt[3][c];
    for (int i = 0; i < 2; ++i) vel[i] += delta_t * inv_mass * u.dudt[1 + i][c];
    const double ie = et - dotN<2>(vel, vel) / 2;
    // update_volume :352-367
    const double d = M / vol;
    u.v[0][c] = vel[0];
    u.v[1][c] = vel[1];
    u.e[c] = ie;
    u.V[c] = vol;
    u.d[c] = d;
    // update_state_from_energy lagrange_eqns.h:280-302
    const double g = u.g[c];
    u.p[c] = eos_p_de(g, d, ie);
    u.T[c] = eos_T_de(P.cv, ie);
    u.a[c] = smax(eos_a_de(g, ie), kMinSoundSpeed);
    if (ie < 0 || d < 0) flag_negative(b.s, tile.prob, m.corig[c], token);
  });
}

// save (dir 0) / restore (dir 1) of coordinates, velocity and energy
__global__ void __launch_bounds__(kBlock) k_maire_copy(DevMesh m, MaireFields u, Batch b, int dir) {
  for_each_tile(b.vt, [&](const Tile &tile) {
    if (!wave_live(b.s, tile.prob, 0)) return;
    if ((int)threadIdx.x >= tile.count) return;
    const int i = tile.first + threadIdx.x;
    for (int d = 0; d < 2; ++d) {
      if (dir == 0)
        u.x0[d][i] = m.x[d][i];
      else
        m.x[d][i] = u.x0[d][i];
    }
  });
  for_each_tile(b.ct, [&](const Tile &tile) {
    if (!wave_live(b.s, tile.prob, 0)) return;
    if ((int)threadIdx.x >= tile.count) return;
    const int i = tile.first + threadIdx.x;
    if (dir == 0) {
      u.v0[0][i] = u.v[0][i];
      u.v0[1][i] = u.v[1][i];
      u.e0[i] = u.e[i];
    } else {
      u.v[0][i] = u.v0[0][i];
      u.v[1][i] = u.v0[1][i];
      u.e[i] = u.e0[i];
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
__global__ void k_maire_end(Batch b) {
  const int nprob = *b.np;
  for (int ip = blockIdx.x * blockDim.x + threadIdx.x; ip < nprob; ip += gridDim.x * blockDim.x) {
    const int p = b.plist[ip];
    if (b.pp[p].cell_count == 0) continue;
    DevScalars &S = b.s[p];
    if (!S.active || S.err_code) continue;
    S.t += S.dt;
    S.steps = S.steps + 1;
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

// evaluate_nodal_state: one launch for both kinds of vertex when both use
// one block per tile (timed as the nodal kernel), else interior vertices
// (Cramer) and then the boundary vertices' symmetry solves
int launch_nodal(fsale_ctx *ctx) {
  Batch bi = batch_args(ctx, K_MAIRE_NODAL);
  const Batch bb = batch_args(ctx, K_MAIRE_NODAL_BND);
  if (bi.vt.direct && bb.bt.direct) {
    bi.bt = bb.bt;
    const int nb = (bb.bt.ndirect + 7) / 8 * 8;
    ProfScope prof(ctx, K_MAIRE_NODAL);
    hipLaunchKernelGGL(k_maire_nodal_both, dim3(nb + bi.vt.ndirect), dim3(kBlock), 0, ctx->stream, ctx->m, ctx->ma,
                       bi, nb);
    FSALE_HIP(hipGetLastError());
    return FSALE_OK;
  }
  int rc = launch_persistent(ctx, K_MAIRE_NODAL, k_maire_nodal<false>, ctx->m, ctx->ma,
                             batch_args(ctx, K_MAIRE_NODAL));
  return rc ? rc
            : launch_persistent(ctx, K_MAIRE_NODAL_BND, k_maire_nodal<true>, ctx->m, ctx->ma,
                                batch_args(ctx, K_MAIRE_NODAL_BND));
}
int launch_residual(fsale_ctx *ctx) {
  return launch_persistent(ctx, K_MAIRE_RESIDUAL, k_maire_residual, ctx->m, ctx->ma,
                           batch_args(ctx, K_MAIRE_RESIDUAL));
}
int launch_dt(fsale_ctx *ctx) {
  int rc = launch_persistent(ctx, K_MAIRE_DT, k_maire_dt_partial, ctx->m, ctx->ma, batch_args(ctx, K_MAIRE_DT));
  if (rc) return rc;
  ProfScope prof(ctx, K_MAIRE_DT);
  hipLaunchKernelGGL(k_maire_dt_finalize, dim3(ctx->pgrid), dim3(kFinalBlock), 0, ctx->stream, batch_args(ctx, -1),
                     (const double *)ctx->ma.partials);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}
// token: the update launch's id for error attribution (predictor 1,
// corrector 2, task-level apply_update 3)
// stage 0: the apply_update task; 1: the driver step's predictor (saves
// coordinates and solution as it goes); 2: its corrector (from the saved ones)
int launch_apply(fsale_ctx *ctx, double factor, int token, int stage) {
  const Batch bm = batch_args(ctx, K_MAIRE_MOVE), bu = batch_args(ctx, K_MAIRE_UPDATE);
  int rc = stage == 1 ? launch_persistent(ctx, K_MAIRE_MOVE, k_maire_move<true>, ctx->m, ctx->ma, bm, factor)
                      : launch_persistent(ctx, K_MAIRE_MOVE, k_maire_move<false>, ctx->m, ctx->ma, bm, factor);
  if (rc) return rc;
  if (stage == 1) return launch_persistent(ctx, K_MAIRE_UPDATE, k_maire_update<1>, ctx->m, ctx->ma, bu, factor, token);
  if (stage == 2) return launch_persistent(ctx, K_MAIRE_UPDATE, k_maire_update<2>, ctx->m, ctx->ma, bu, factor, token);
  return launch_persistent(ctx, K_MAIRE_UPDATE, k_maire_update<0>, ctx->m, ctx->ma, bu, factor, token);
}
int launch_copy(fsale_ctx *ctx, int dir) {
  return launch_persistent(ctx, K_MAIRE_COPY, k_maire_copy, ctx->m, ctx->ma, batch_args(ctx, K_MAIRE_COPY), dir);
}

}  // namespace

// One driver step (apps/maire_hydro/driver.h:330-538) of every live problem:
// predictor half step, then the corrector from the saved state.  The
// driver's save_coordinates / save_solution happen inside the predictor's
// move and update (each saves what it reads before writing), and
// restore_solution / restore_coordinates are not needed: the corrector's
// update reads the saved v0, e0, and its move writes x = x0 + dt u (when it
// does not move, dt <= 1e-14, the predictor did not move either and x is
// still x0).
int maire_step(fsale_ctx *ctx) {
  int rc;
  const int np = (int)ctx->probs.size();
  hipLaunchKernelGGL(k_maire_begin, dim3(small_grid(np)), dim3(kBlock), 0, ctx->stream, batch_args(ctx, -1));
  FSALE_HIP(hipGetLastError());
  if ((rc = launch_nodal(ctx))) return rc;
  if ((rc = launch_residual(ctx))) return rc;
  if ((rc = launch_dt(ctx))) return rc;
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
  for (int d = 0; d < 2; ++d) {
    if ((rc = dev_alloc_n(ctx, &u.v[d], m.cap_c))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.v0[d], m.cap_c))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.x0[d], m.cap_v))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.un[d], m.cap_v))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.npc[d], (size_t)m.cap_c * m.vpc))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.Fpc[d], (size_t)m.cap_c * m.vpc))) return rc;
  }
  for (int i = 0; i < 4; ++i)
    if ((rc = dev_alloc_n(ctx, &u.dudt[i], m.cap_c))) return rc;
  if ((rc = dev_alloc_n(ctx, &u.partials, 2 * (size_t)kWaves * (ctx->ra_ct.capacity + 1)))) return rc;
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

int fsale_maire_evaluate_nodal_state(fsale_ctx *ctx) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  if (!ctx->probs[0].have_state) return FSALE_ERR_STATE;
  int rc = sync_tiles(ctx);
  return rc ? rc : launch_nodal(ctx);
}

int fsale_maire_evaluate_residual(fsale_ctx *ctx) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  if (!ctx->probs[0].have_state) return FSALE_ERR_STATE;
  int rc = sync_tiles(ctx);
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
  if ((rc = upload_params(ctx, 0)) || (rc = sync_tiles(ctx)) || (rc = launch_dt(ctx))) return rc;
  if (!dt_out) return FSALE_OK;
  if ((rc = check_device_error(ctx, 0))) return rc;
  *dt_out = ctx->h_s[0].dt;
  return FSALE_OK;
}

int fsale_maire_apply_update(fsale_ctx *ctx, double factor) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  if (!ctx->probs[0].have_state) return FSALE_ERR_STATE;
  int rc = sync_tiles(ctx);
  return rc ? rc : launch_apply(ctx, factor, 3, 0);
}

int fsale_maire_save(fsale_ctx *ctx) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  int rc = sync_tiles(ctx);
  return rc ? rc : launch_copy(ctx, 0);
}

int fsale_maire_restore(fsale_ctx *ctx) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  int rc = sync_tiles(ctx);
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
