// Eulerian MUSCL finite-volume step of apps/hydro on the device.
//
//   k_eul_dt_partial + k_eul_dt_finalize   evaluate_time_step  apps/hydro/tasks.h:331-376
//   k_eul_grad                             evaluate_gradients  apps/hydro/tasks.h:145-321
//   k_eul_flux                             evaluate_fluxes     apps/hydro/tasks.h:385-536
//   k_eul_update                           apply_update        apps/hydro/tasks.h:589-659
//   k_eul_fluxupd   evaluate_fluxes + apply_update (+ the next step's dt^-1
//                   partials), one work-item per cell: the driver loop's path
//   k_eul_grad_lds  the driver loop's gradients in 2D, tiles staged in LDS
//
// Every kernel but the task-level face kernel walks the batch's tile list
// (batch_dev.h: one block per tile, or a persistent grid), so one launch
// advances every resident problem; problem scalars (EOS, CFL, dt, time,
// flags) come from per-problem records.  SoA FP64 state for the task-level
// kernels; the driver loop works on packed per-cell records (pairs of
// doubles in planes, EulerFields::gs / gg), packed before a batch of steps
// and unpacked after it.  k-major fixed-degree connectivity so neighbouring
// lanes read neighbouring words.  In 2D a tile whose neighbours outside it
// are few (its halo, DevMesh::halo) is staged in LDS with them, and its
// cells read their neighbours there.
//
// k_eul_fluxupd recomputes, for each cell, the numerical flux of each of its
// faces exactly as k_eul_flux does (same left/right states, same operation
// order), so a face shared by two cells gets the bit-identical flux on both
// sides and the update gathers them in the reference's mesh.faces(c) order.
// In 2D the face centroid / normal / length come from the cell's own
// vertices (bit-identical to the stored face geometry: reversing an edge
// negates it exactly).  The flux array never goes through HBM.  Neighbours
// must read the old state, so the state records ping-pong between two
// buffers.
//
// The time step: each wave writes the max of dt^-1 over its cells to a
// partials slot (no atomics, no block barrier); one block per problem folds
// them into dt.  In the driver loop the partials come from the second
// half-step's update, which already holds the new state.
#include "batch_dev.h"

namespace fsale {
namespace {

// ------------------------------------------------------------ state init
__global__ void __launch_bounds__(kBlock) k_eul_ic(EulerFields u, const ProbParams *pp, int prob) {
  const ProbParams &P = pp[prob];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.cell_count) return;
  const int c = P.cell_first + i;
  // update_state_from_pressure euler_eqns.h:333-350
  const double ie = eos_e_dp(P.gamma, u.d[c], u.p[c]);
  u.e[c] = ie;
  u.T[c] = eos_T_de(P.cv, ie);
  u.a[c] = eos_a_de(P.gamma, ie);
}

// ----------------------------------------------------------- time step
template <int N, int FPC>
__global__ void __launch_bounds__(kBlock) k_eul_dt_partial(DevMesh m, EulerFields u, Batch b) {
  for_each_tile(b.ct, [&](const Tile &tile) {
    double dt_inv = 0;
    const int owned = b.pp[tile.prob].owned;
    // mesh.cells(flecsi::owned) (apps/hydro/tasks.h:350): a ghost cell of a
    // partitioned context contributes nothing
    if ((int)threadIdx.x < tile.count &&
        (owned == 0 || m.corig[tile.first + threadIdx.x] < owned)) {
      const int c = tile.first + threadIdx.x;
      double vel[N];
      for (int d = 0; d < N; ++d) vel[d] = u.v[d][c];
      const double a = u.a[c], vol = m.cvol[c];
      double px[FPC], py[FPC];
      if constexpr (N == 2) {
        for (int k = 0; k < FPC; ++k) {
          const double2 w = m.xy[m.cv[(size_t)k * m.cap_c + c]];
          px[k] = w.x;
          py[k] = w.y;
        }
      }
#pragma unroll
      for (int k = 0; k < FPC; ++k) {
        double n[N], area;
        if constexpr (N == 2) {
          // the face's normal up to its sign (|v.n| does not see it) and its
          // length, from the cell's own vertices, bit-identical to the stored
          // face geometry
          double fx, fy;
          const int k1 = (k + 1 == FPC) ? 0 : k + 1;
          face_geom_2d(px[k], py[k], px[k1], py[k1], fx, fy, n[0], n[1], area);
        } else {
          const size_t o = (size_t)k * m.cap_c + c;
          const double2 w1 = m.kfp[1][o], w2 = m.kfp[2][o];
          n[0] = w1.y;
          n[1] = w2.x;
          n[2] = w2.y;
          area = m.kfa[o];
        }
        const double delta_x = vol / area;
        const double dti = (a + fabs(dotN<N>(vel, n))) / delta_x;
        dt_inv = smax(dti, dt_inv);
      }
    }
    dt_inv = wave_max(dt_inv);
    if ((threadIdx.x & 63) == 0) u.partials[tile.slot * kWaves + (threadIdx.x >> 6)] = dt_inv;
  });
}

// Per problem: dt = min(CFL / max dt^-1, t_final - t) and the step flag.
// driver != 0: the step loop (step budget, optional stop at t_final);
// driver == 0: the task-level evaluate_time_step.
// One problem, by a whole block of B threads.
template <int B>
__device__ inline void finalize_problem(const Batch &b, const double *partials, int driver, int p) {
  const ProbParams &P = b.pp[p];
  if (P.cell_count == 0) return;
  double mx[1];
  final_max<1, B>(partials + (size_t)P.tile_first * kWaves, P.tile_count * kWaves, mx);
  if (threadIdx.x != 0) return;
  {
    DevScalars &S = b.s[p];
    const double t = S.t;
    int active = (S.err_code == 0);
    if (driver) {
      if (!(S.steps < P.max_steps)) active = 0;
      if (P.stop_at_final && !(t < P.final_time)) active = 0;
    }
    if (active) {
      if (mx[0] <= 0) {   // apps/hydro/tasks.h:366-367
        S.err_code = FSALE_ERR_DT;
        S.err_token = -1;
        active = 0;
      } else {
        double ts = 1 / mx[0];
        ts *= P.cfl;
        ts = smin(ts, P.final_time - t);
        S.dt = ts;
        if (driver) S.steps = S.steps + 1;
      }
    }
    S.active = active;
  }
}

__global__ void __launch_bounds__(kFinalBlock) k_eul_dt_finalize(Batch b, const double *partials, int driver) {
  const int nprob = *b.np;
  for (int ip = blockIdx.x; ip < nprob; ip += gridDim.x) finalize_problem<kFinalBlock>(b, partials, driver, b.plist[ip]);
}

// The 3D fused kernel streams each cell's face geometry and writes its new
// state with the non-temporal hint (read or written once per launch), so
// that they do not displace the neighbour records it reuses from L2
// (MUSCL3D 128^3: 420 -> 407 us per launch).
#ifndef FSALE_NT3
#define FSALE_NT3 1
#endif

// ----------------------------------------------------------- gradients
// Waves per SIMD the 2D gradient kernel is register-limited to.
#ifndef FSALE_GRAD_WAVES_2D
#define FSALE_GRAD_WAVES_2D 3
#endif

// Packed per-cell records of the driver loop (EulerFields::gs, gg): state
// {d, v[N], e, p, a} then the centroid at C; gradients {d, v, e, p, a} x N.
// A record is stored as pairs of doubles, one plane (array) per pair, so a
// lane loads 16 bytes per instruction and a wave's load covers 1 KiB of
// consecutive memory: half the instructions of one 8-byte load per field,
// the same cache lines.  (A whole record per lane, 64 bytes apart, touches
// four times the cache lines per instruction and measured slower.)
template <int N>
struct Rec {
  static constexpr int S = 2 * N + 4;            // doubles per state record
  static constexpr int C = N + 4;                // centroid offset in it
  static constexpr int G = (N == 2) ? 12 : 22;   // doubles per gradient record (3D: 21 + 1 pad)
  // state planes from here on hold centroid values only: static, written to
  // both state buffers when the records are packed and never by a step
  static constexpr int P_STATIC = (C + 1) / 2;
};

template <int K, int P>
__device__ __forceinline__ void ld_rec(double2 *const (&pl)[P], int c, double (&r)[K]) {
  static_assert(K / 2 <= P, "record larger than its planes");
  const uint32_t off = (uint32_t)c * 16u;
#pragma unroll
  for (int i = 0; i < K / 2; ++i) {
    const double2 w = ld(pl[i], off);
    r[2 * i] = w.x;
    r[2 * i + 1] = w.y;
  }
}
// planes [I0, I1) of a record
template <int K, int P, int I0 = 0, int I1 = K / 2>
__device__ __forceinline__ void st_rec(double2 *const (&pl)[P], int c, const double (&r)[K]) {
  static_assert(K / 2 <= P && I1 <= K / 2, "record larger than its planes");
  const uint32_t off = (uint32_t)c * 16u;
#pragma unroll
  for (int i = I0; i < I1; ++i) st(pl[i], off, make_double2(r[2 * i], r[2 * i + 1]));
}

// LDS staging of a 2D tile: planes of pairs, the tile's cells at slots
// [0, kBlock), its halo from slot kBlock (plane i at sh + i * kLdsW)
constexpr int kLdsW = kBlock + kHalo2;

// One record from the staged planes (plane i at sh + i * W).
template <int K, int W = kLdsW>
__device__ __forceinline__ void lds_rec(const double2 *sh, int slot, double (&r)[K]) {
#pragma unroll
  for (int i = 0; i < K / 2; ++i) {
    const double2 w = sh[i * W + slot];
    r[2 * i] = w.x;
    r[2 * i + 1] = w.y;
  }
}

// Stages one tile with h halo cells: the records of its cells and of its
// halo, planes `plane(i)` for i < NP, into sh (all threads of the block;
// barriers before, so the previous tile's reads are done, and after).  A
// tile of a problem that does not step is staged too (its waves then
// return): reads only, of cells and halo lists in bounds whoever owns them.
// The thread's halo entry (thread j < h stages halo cell j: kHalo2 <=
// kBlock).  The fused kernel loads it ahead of the staging, with its other
// reads, so that the halo's record loads do not wait on it there (the
// gradient kernel, measured slower that way, loads it in the staging).
static_assert(kHalo2 <= kBlock, "one halo cell per thread");
__device__ __forceinline__ int halo_entry(const DevMesh &m, const Tile &tile, int h) {
  const int j = threadIdx.x;
  return j < h ? m.halo[(size_t)tile.slot * kHalo2 + j] : 0;
}
// `again`: the block may have staged another tile before (a persistent
// grid), whose reads must be done first.
template <int NP, bool PRE, class P>
__device__ __forceinline__ void stage_tile(double2 *sh, const DevMesh &m, const Tile &tile, int h, int hcell,
                                           P &&plane, bool again = true) {
  if (again) __syncthreads();
  const int t = threadIdx.x;
  if constexpr (!PRE) hcell = halo_entry(m, tile, h);
  if (t < tile.count) {
    const uint32_t off = (uint32_t)(tile.first + t) * 16u;
#pragma unroll
    for (int i = 0; i < NP; ++i) sh[i * kLdsW + t] = ld(plane(i), off);
  }
  if (t < h) {
    const uint32_t off = (uint32_t)hcell * 16u;
#pragma unroll
    for (int i = 0; i < NP; ++i) sh[i * kLdsW + kBlock + t] = ld(plane(i), off);
  }
  __syncthreads();
}

// A tile of a structured 3D grid (ProbParams::s_nx), for the staging of its
// neighbour planes: the tile's cells, the problem's cell range and strides.
struct S3Tile {
  int first, count, cf, nc, nx, np;
};

// The staging of an 8 x 8 x 4 brick of a structured grid (ProbParams::
// s_brick) for the gradients: the brick and the ring of cells around it,
// 10 x 10 x 6 slots (x fastest; slots off the grid are not written).  For
// one cell: its slot, its place in the grid and the grid's extent, and
// whether the brick lies inside the grid with its whole ring.
constexpr int kB4 = (kBrickX + 2) * (kBrickY + 2) * (kBrickZ + 2);
struct B4Cell {
  int slot, ix, jy, kz, nx, ny, nz, interior;
};

// MODE 0: the per-field arrays (task level); 1: the packed state records
// (gs), as the driver loop does; 2 (2D): the records staged in LDS (sh),
// neighbours by their LDS slots (DevMesh::lnb); 3 (3D structured grids):
// the records of each neighbour plane staged in LDS in turn (sh, a
// contiguous range of cells per plane; every thread of the block takes part
// in the staging, `valid` lanes compute and store); 4 (3D grids in bricks):
// the brick and its ring staged in LDS (sh, kB4 slots per plane, b4).
// Modes 1-4 write the packed gradients (gg).
template <int N, int VPC, int MODE>
__device__ __forceinline__ void grad_cell(const DevMesh &m, const EulerFields &u, int c0, int first_order,
                                 double2 *sh = nullptr, const S3Tile &s3t = S3Tile{}, bool valid = true,
                                 const B4Cell &b4 = B4Cell{}) {
  static_assert(MODE != 2 || N == 2, "LDS staging is 2D");
  static_assert((MODE != 3 && MODE != 4) || N == 3, "plane and brick staging are 3D");
  constexpr bool PK = MODE != 0;
  constexpr int NV = N + 4;
  const double tol = 1.e-9;
  const uint32_t o0 = (uint32_t)c0 * 8u, i0 = (uint32_t)c0 * 4u;
  double u0[NV], umin[NV], umax[NV], dudx[NV][N], dxdx[N][N], x0[N];
  if constexpr (PK) {
    double r[Rec<N>::S];
    if constexpr (MODE == 2)
      lds_rec(sh, (int)threadIdx.x, r);
    else if constexpr (MODE == 4)
      lds_rec<Rec<N>::S, kB4>(sh, b4.slot, r);
    else
      ld_rec(u.gs, c0, r);
    for (int j = 0; j < NV; ++j) u0[j] = r[j];
    for (int d = 0; d < N; ++d) x0[d] = r[Rec<N>::C + d];
  } else {
    for (int d = 0; d < N; ++d) x0[d] = ld(m.ccen[d], o0);
    int i = 0;
    u0[i++] = ld(u.d, o0);
    for (int d = 0; d < N; ++d) u0[i++] = ld(u.v[d], o0);
    u0[i++] = ld(u.e, o0);
    u0[i++] = ld(u.p, o0);
    u0[i++] = ld(u.a, o0);
  }
  for (int j = 0; j < NV; ++j) {
    umin[j] = u0[j];
    umax[j] = u0[j];
    for (int d = 0; d < N; ++d) dudx[j][d] = 0;
  }
  for (int r = 0; r < N; ++r)
    for (int q = 0; q < N; ++q) dxdx[r][q] = 0;
  // Index loads first, all independent: the k-major neighbour head, the
  // neighbour count and (2D) the cell's vertices, whose coordinates the
  // limiter needs; the 3D kernel has no registers to spare for 24
  // coordinates and loads them in the limiter.  A missing neighbour
  // (a -1 slot; the rows are right-aligned, so these come before every real
  // neighbour) reads the cell itself, so no load waits behind a branch, and
  // it changes nothing: std::min(u0, u0) and std::max(u0, u0) are u0 bit for
  // bit, and its products (du = dx = +0) add +0 to sums still at +0.
  constexpr int NBK = (N == 2) ? kNbk2 : kNbk3;
  const size_t row = (size_t)m.cap_c;
  int nb[N == 2 ? NBK : 1];
  if constexpr (MODE == 2) {
    // LDS slots; the tile's plan holds every neighbour in the k-major head
#pragma unroll
    for (int k = 0; k < NBK / 2; ++k) {
      const uint32_t w = (uint32_t)ld(m.lnb + k * row, i0);
      nb[2 * k] = (int)(w & 0xffffu);
      nb[2 * k + 1] = (int)(w >> 16);
    }
  } else if constexpr (N == 2) {
#pragma unroll
    for (int k = 0; k < NBK; ++k) nb[k] = ld(m.nbk + k * row, i0);
  }
  const int cnt = (MODE >= 2) ? 0 : ld(m.nbr_cnt, i0);
  double xv[VPC][N];
  auto load_vertex = [&](int lv) {
    const uint32_t iv = (uint32_t)ld(m.cv + lv * row, i0);
    if constexpr (N == 2) {
      const double2 w = ld(m.xy, iv * 16u);
      xv[lv][0] = w.x;
      xv[lv][1] = w.y;
    } else {
      for (int d = 0; d < N; ++d) xv[lv][d] = ld(m.x[d], iv * 8u);
    }
  };
  // cell_vertex_neighbors in ascending id order: the head, then the CSR tail
  auto neighbour = [&](int c1) {
    const uint32_t o1 = (uint32_t)c1 * 8u;
    double u1[NV], dx[N];
    if constexpr (PK) {
      double r[Rec<N>::S];
      if constexpr (MODE == 2)
        lds_rec(sh, c1, r);
      else if constexpr (MODE == 3)
        lds_rec<Rec<N>::S, kLds3>(sh, c1, r);
      else if constexpr (MODE == 4)
        lds_rec<Rec<N>::S, kB4>(sh, c1, r);
      else
        ld_rec(u.gs, c1, r);
      for (int j = 0; j < NV; ++j) u1[j] = r[j];
      for (int d = 0; d < N; ++d) dx[d] = r[Rec<N>::C + d] - x0[d];
    } else {
      int i = 0;
      u1[i++] = ld(u.d, o1);
      for (int d = 0; d < N; ++d) u1[i++] = ld(u.v[d], o1);
      u1[i++] = ld(u.e, o1);
      u1[i++] = ld(u.p, o1);
      u1[i++] = ld(u.a, o1);
      for (int d = 0; d < N; ++d) dx[d] = ld(m.ccen[d], o1) - x0[d];
    }
    // umin = std::min(u1, umin), umax = std::max(u1, umax) (tasks.h:206-207)
    for (int j = 0; j < NV; ++j) {
      umin[j] = hmin(u1[j], umin[j]);
      umax[j] = hmax(u1[j], umax[j]);
    }
    if constexpr (N == 2) {
      dxdx[0][0] += dx[0] * dx[0];
      dxdx[0][1] += dx[0] * dx[1];
      dxdx[1][1] += dx[1] * dx[1];
    } else {
      dxdx[0][0] += dx[0] * dx[0];
      dxdx[0][1] += dx[0] * dx[1];
      dxdx[0][2] += dx[0] * dx[2];
      dxdx[1][1] += dx[1] * dx[1];
      dxdx[1][2] += dx[1] * dx[2];
      dxdx[2][2] += dx[2] * dx[2];
    }
    for (int j = 0; j < NV; ++j) {
      const double du = u1[j] - u0[j];
      for (int d = 0; d < N; ++d) dudx[j][d] += du * dx[d];
    }
  };
  if constexpr (MODE == 3) {
    // the grid's neighbours in ascending id order, (dz, dy, dx)
    // lexicographic: for each plane dz the tile's neighbours there are the
    // cells [first + dz np - nx - 1, first + count - 1 + dz np + nx + 1] of
    // the problem, staged whole; a neighbour off the grid is skipped, as
    // the neighbour list leaves it out
    constexpr int NPS = Rec<N>::S / 2;
    const int li = c0 - s3t.cf, kz = li / s3t.np, rem = li - kz * s3t.np, jy = rem / s3t.nx, ix = rem - jy * s3t.nx;
    const int ny = s3t.np / s3t.nx, nz = s3t.nc / s3t.np;
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz) {
      const int lo = max(s3t.first + dz * s3t.np - s3t.nx - 1, s3t.cf);
      const int hi = min(s3t.first + s3t.count - 1 + dz * s3t.np + s3t.nx + 1, s3t.cf + s3t.nc - 1);
      __syncthreads();   // the previous plane's (or tile's) reads are done
      for (int q = threadIdx.x; q <= hi - lo; q += kBlock) {
        const uint32_t off = (uint32_t)(lo + q) * 16u;
#pragma unroll
        for (int pl = 0; pl < NPS; ++pl) sh[pl * kLds3 + q] = ld(u.gs[pl], off);
      }
      __syncthreads();
      if (valid && kz + dz >= 0 && kz + dz < nz) {
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
          if (jy + dy < 0 || jy + dy >= ny) continue;
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx) {
            if ((dz == 0 && dy == 0 && dx == 0) || ix + dx < 0 || ix + dx >= s3t.nx) continue;
            neighbour(c0 + dz * s3t.np + dy * s3t.nx + dx - lo);
          }
        }
      }
    }
  } else if constexpr (MODE == 4) {
    // the grid's neighbours in ascending id order, (dz, dy, dx)
    // lexicographic, at slot offsets dz (10 x 10) + dy 10 + dx; a brick with
    // its whole ring inside the grid needs no test (and its loop no branch),
    // otherwise a neighbour off the grid is skipped, as the neighbour list
    // leaves it out
    constexpr int SX = 1, SY = kBrickX + 2, SZ = (kBrickX + 2) * (kBrickY + 2);
    if (b4.interior) {
#pragma unroll
      for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx)
            if (dz != 0 || dy != 0 || dx != 0) neighbour(b4.slot + dz * SZ + dy * SY + dx * SX);
    } else {
#pragma unroll
      for (int dz = -1; dz <= 1; ++dz) {
        if (b4.kz + dz < 0 || b4.kz + dz >= b4.nz) continue;
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy) {
          if (b4.jy + dy < 0 || b4.jy + dy >= b4.ny) continue;
#pragma unroll
          for (int dx = -1; dx <= 1; ++dx) {
            if ((dz == 0 && dy == 0 && dx == 0) || b4.ix + dx < 0 || b4.ix + dx >= b4.nx) continue;
            neighbour(b4.slot + dz * SZ + dy * SY + dx * SX);
          }
        }
      }
    }
  } else if constexpr (MODE == 2) {
#pragma unroll
    for (int k = 0; k < NBK; ++k) neighbour(nb[k]);
  } else if constexpr (N == 2) {
#pragma unroll
    for (int k = 0; k < NBK; ++k) neighbour(nb[k] >= 0 ? nb[k] : c0);
  } else {
    // 28 neighbour slots hoisted at once would spill: groups of four (two
    // for the 12-double packed records), each group's loads issued together
    constexpr int GRP = PK ? 2 : 4;
#pragma unroll 1
    for (int k0 = 0; k0 < NBK; k0 += GRP) {
      int g[GRP];
#pragma unroll
      for (int j = 0; j < GRP; ++j) g[j] = ld(m.nbk + (k0 + j) * row, i0);
#pragma unroll
      for (int j = 0; j < GRP; ++j) neighbour(g[j] >= 0 ? g[j] : c0);
    }
  }
  if (cnt > NBK) {
    const int q0 = ld(m.nbr_start, i0);
    for (int q = q0 + NBK; q < q0 + cnt; ++q) neighbour(m.nbr[q]);
  }
  if constexpr (N == 2) {
#pragma unroll
    for (int lv = 0; lv < VPC; ++lv) load_vertex(lv);
  }
  double gu[NV][N];
  if constexpr (N == 2) {
    // gradient_2d apps/hydro/tasks.h:109-117
    const double denom = 1 / (dxdx[0][0] * dxdx[1][1] - dxdx[0][1] * dxdx[0][1]);
    for (int j = 0; j < NV; ++j) {
      gu[j][0] = (dudx[j][0] * dxdx[1][1] - dudx[j][1] * dxdx[0][1]) * denom;
      gu[j][1] = (dudx[j][1] * dxdx[0][0] - dudx[j][0] * dxdx[0][1]) * denom;
    }
  } else {
    // gradient_3d apps/hydro/tasks.h:119-143
    const double a00 = dxdx[0][0], a01 = dxdx[0][1], a02 = dxdx[0][2];
    const double a11 = dxdx[1][1], a12 = dxdx[1][2], a22 = dxdx[2][2];
    const double min1 = a11 * a22 - a12 * a12;
    const double min2 = a01 * a22 - a12 * a02;
    const double min3 = a01 * a12 - a11 * a02;
    const double denom = 1 / (a00 * min1 - a01 * min2 + a02 * min3);
    for (int j = 0; j < NV; ++j) {
      const double g0 = dudx[j][0], g1 = dudx[j][1], g2 = dudx[j][2];
      gu[j][0] = (g0 * min1 - a01 * (g1 * a22 - a12 * g2) + a02 * (g1 * a12 - a11 * g2)) * denom;
      gu[j][1] = (a00 * (g1 * a22 - a12 * g2) - g0 * min2 + a02 * (g2 * a01 - a02 * g1)) * denom;
      gu[j][2] = (a00 * (g2 * a11 - a12 * g1) - a01 * (g2 * a01 - a02 * g1) + g0 * min3) * denom;
    }
  }
  // limiter apps/hydro/tasks.h:267-299.  The reference takes, over the cell's
  // vertices, phi = min(1, (umax-u0)/du for du > tol, (umin-u0)/du for
  // du < -tol).  Correctly rounded division is monotone in the divisor, so
  // the minimum over the first kind is (umax-u0)/max(du) and over the second
  // (umin-u0)/min(du): two divisions per variable instead of one per vertex,
  // with the same phi (up to the sign of a zero phi, which only flips the
  // sign of a zero gradient and never changes a reconstructed value).  The
  // maximum over every vertex's du is the maximum over those above tol
  // whenever one is, so the extrema need no branch.  A quotient whose
  // numerator is at least its divisor in magnitude rounds to >= 1 and leaves
  // phi alone, so only the limiting cases divide.
  double dmax[NV], dmin[NV];
#pragma unroll
  for (int lv = 0; lv < VPC; ++lv) {
    if constexpr (N != 2) load_vertex(lv);
    double u1[NV];
    for (int j = 0; j < NV; ++j) u1[j] = u0[j];
    for (int d = 0; d < N; ++d) {
      const double dx = xv[lv][d] - x0[d];
      for (int j = 0; j < NV; ++j) u1[j] += dx * gu[j][d];
    }
    for (int j = 0; j < NV; ++j) {
      const double du = u1[j] - u0[j];
      dmax[j] = lv ? hmax(dmax[j], du) : du;
      dmin[j] = lv ? hmin(dmin[j], du) : du;
    }
  }
  for (int j = 0; j < NV; ++j) {
    double phi = 1;
    const double hi = umax[j] - u0[j], lo = umin[j] - u0[j];
    if (dmax[j] > tol && hi < dmax[j]) phi = hi / dmax[j];
    if (dmin[j] < -tol && lo > dmin[j]) phi = smin(phi, lo / dmin[j]);
    phi = smax(phi, 0);
    for (int d = 0; d < N; ++d) gu[j][d] = first_order ? 0.0 : gu[j][d] * phi;
  }
  if (!valid) return;
  if constexpr (PK) {
    // the record is gu row by row: grad v[k][d] (d/dx_d of v_k) at N + k*N + d
    double r[Rec<N>::G];
    for (int j = 0; j < NV; ++j)
      for (int d = 0; d < N; ++d) r[j * N + d] = gu[j][d];
    for (int i = NV * N; i < Rec<N>::G; ++i) r[i] = 0;
    st_rec(u.gg, c0, r);
  } else {
    for (int d = 0; d < N; ++d) {
      st(u.gd[d], o0, gu[0][d]);
      for (int k = 0; k < N; ++k) st(u.gv[k * N + d], o0, gu[1 + k][d]);
      st(u.ge[d], o0, gu[1 + N][d]);
      st(u.gp[d], o0, gu[2 + N][d]);
      st(u.ga[d], o0, gu[3 + N][d]);
    }
  }
}

// FIN (the driver step's first gradient launch): the block of each
// problem's first tile also finalizes the problem's time step (the work of
// k_eul_dt_finalize(driver), one launch less per step).  The step flag is
// then being written during the launch, so no wave waits on it; the tiles of
// a problem that was not packed for this batch of steps (S.packed, set by
// k_eul_pack before the batch and not written during it) are skipped, as
// every other kernel skips them: a finished problem's ranges may already
// hold a new admission.  The gradients of a packed problem that stops
// within the batch are computed over its own ranges and never read.
template <int N, int VPC, int MODE, bool FIN = false>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(N == 2 ? FSALE_GRAD_WAVES_2D : 3)))
k_eul_grad(DevMesh m, EulerFields u, Batch b) {
  for_each_tile(b.ct, [&](const Tile &tile) {
    if constexpr (FIN) {
      if (tile.first == b.pp[tile.prob].cell_first) finalize_problem<kBlock>(b, u.partials, 1, tile.prob);
      if (!b.s[tile.prob].packed) return;
    } else if (!wave_live(b.s, tile.prob, 0)) {
      return;
    }
    if ((int)threadIdx.x < tile.count)
      grad_cell<N, VPC, MODE>(m, u, tile.first + threadIdx.x, b.pp[tile.prob].first_order);
  });
}

// 2D quads with LDS staging plans: a tile with a plan has its state records
// and its halo's staged in LDS and takes the gradients from there; a tile
// without one (halo_cnt -1: a halo over kHalo2 cells or a cell past the
// k-major head) gathers from memory.  The choice is per tile, so uniform
// across the block and its barriers.
// ALL: every tile of the launch has a plan (the memory path is not
// compiled in, which leaves the staged path more registers).
template <int VPC, bool ALL, bool FIN = false>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_eul_grad_lds(DevMesh m, EulerFields u, Batch b) {
  constexpr int NP = Rec<2>::S / 2;
  __shared__ double2 sh[NP * kLdsW];
  for_each_tile(b.ct, [&](const Tile &tile) {
    // the tile's plan and the problem's order are read ahead of the FIN
    // checks, so that the loads overlap (discarded when the tile is skipped)
    const int h = m.halo_cnt[tile.slot];
    const int first_order = b.pp[tile.prob].first_order;
    // FIN: as k_eul_grad
    if (FIN && tile.first == b.pp[tile.prob].cell_first) finalize_problem<kBlock>(b, u.partials, 1, tile.prob);
    if (FIN && !b.s[tile.prob].packed) return;   // uniform across the block
    if (!ALL && h < 0) {
      if (!FIN && !wave_live(b.s, tile.prob, 0)) return;
      if ((int)threadIdx.x < tile.count) grad_cell<2, VPC, 1>(m, u, tile.first + threadIdx.x, first_order);
      return;
    }
    stage_tile<NP, false>(sh, m, tile, h, 0, [&](int i) { return u.gs[i]; }, b.ct.direct == nullptr);
    if (!FIN && !wave_live(b.s, tile.prob, 0)) return;
    if ((int)threadIdx.x < tile.count) grad_cell<2, VPC, 2>(m, u, tile.first + threadIdx.x, first_order, sh);
  });
}

// 3D: the tiles of problems on a structured grid (ProbParams::s_nx) stage
// their neighbour planes in LDS (grad_cell MODE 3); a tile of another
// problem gathers from memory (not compiled in with ALL: every problem is
// on a grid).  The choice is per problem, so uniform across the block.
// The waves of a tile agree on wave_live (a problem's flags do not change
// during the launch outside FIN), so a dead problem's block leaves whole.
template <bool ALL, bool FIN = false>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_eul_grad_s3(DevMesh m, EulerFields u, Batch b) {
  constexpr int NP = Rec<3>::S / 2;
  __shared__ double2 sh[NP * kLds3];
  for_each_tile(b.ct, [&](const Tile &tile) {
    const ProbParams &P = b.pp[tile.prob];
    // FIN: as k_eul_grad
    if (FIN && tile.first == P.cell_first) finalize_problem<kBlock>(b, u.partials, 1, tile.prob);
    if (FIN && !b.s[tile.prob].packed) return;
    if (!FIN && !wave_live(b.s, tile.prob, 0)) return;
    const int first_order = P.first_order;
    if (!ALL && P.s_nx == 0) {
      if ((int)threadIdx.x < tile.count) grad_cell<3, 8, 1>(m, u, tile.first + threadIdx.x, first_order);
      return;
    }
    const bool valid = (int)threadIdx.x < tile.count;
    const S3Tile s3t{tile.first, tile.count, P.cell_first, P.cell_count, P.s_nx, P.s_np};
    grad_cell<3, 8, 3>(m, u, valid ? tile.first + (int)threadIdx.x : tile.first, first_order, sh, s3t, valid);
  });
}

// 3D grids in bricks (every live problem with ProbParams::s_brick; tile
// slot -> brick as the fused step: DevMesh::brick): the block stages its
// brick and the ring around it once and takes every neighbour from LDS.
// FIN and the liveness rules as k_eul_grad_s3.
template <bool FIN = false>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_eul_grad_brick(DevMesh m, EulerFields u, Batch b) {
  constexpr int NP = Rec<3>::S / 2;
  constexpr int RX = kBrickX + 2, RY = kBrickY + 2;
  __shared__ double2 sh[NP * kB4];
  for_each_tile(b.ct, [&](const Tile &tile) {
    const ProbParams &P = b.pp[tile.prob];
    // the brick and the grid's shape are read ahead of the checks below, so
    // that the loads overlap (discarded when the tile is skipped)
    const int base = m.brick[tile.slot];
    const int nx = P.s_nx, np = P.s_np, ncells = P.cell_count;
    if (FIN && tile.first == P.cell_first) finalize_problem<kBlock>(b, u.partials, 1, tile.prob);
    if (FIN && !b.s[tile.prob].packed) return;
    if (!FIN && !wave_live(b.s, tile.prob, 0)) return;
    const int ny = np / nx, nz = ncells / np;
    const int lb = base - P.cell_first;
    const int z0 = lb / np, y0 = (lb - z0 * np) / nx, x0 = lb - z0 * np - y0 * nx;
    const bool interior = x0 > 0 && x0 + kBrickX < nx && y0 > 0 && y0 + kBrickY < ny && z0 > 0 && z0 + kBrickZ < nz;
    __syncthreads();   // a persistent block's previous brick is done with the LDS
    for (int q = threadIdx.x; q < kB4; q += kBlock) {
      const int qi = q % RX, qj = (q / RX) % RY, qk = q / (RX * RY);
      const int gx = x0 + qi - 1, gy = y0 + qj - 1, gz = z0 + qk - 1;
      if (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz >= 0 && gz < nz) {
        const uint32_t off = (uint32_t)(base + (qi - 1) + (qj - 1) * nx + (qk - 1) * np) * 16u;
#pragma unroll
        for (int pl = 0; pl < NP; ++pl) sh[pl * kB4 + q] = ld(u.gs[pl], off);
      }
    }
    __syncthreads();
    const int t = threadIdx.x, ti = t % kBrickX, tj = (t / kBrickX) % kBrickY, tk = t / (kBrickX * kBrickY);
    const B4Cell b4{((tk + 1) * RY + tj + 1) * RX + ti + 1, x0 + ti, y0 + tj, z0 + tk, nx, ny, nz, interior};
    grad_cell<3, 8, 4>(m, u, base + ti + tj * nx + tk * np, P.first_order, sh, S3Tile{}, true, b4);
  });
}

// LDS slots of one problem's cells (DevMesh::lnb, lfn): a neighbour in the
// cell's own tile is its position in the tile, any other its place in the
// tile's sorted halo list (which holds every one for a tile with a plan)
// after the kBlock own slots; padding slots and boundary faces name the
// cell itself.
__global__ void __launch_bounds__(kBlock) k_lds_slots(DevMesh m, int c0, int nc, int ct0) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= nc) return;
  const int j = i / kBlock, slot = ct0 + j;
  const int first = c0 + j * kBlock, count = min(kBlock, nc - j * kBlock);
  const int h = m.halo_cnt[slot];
  if (h < 0) return;
  const int32_t *hl = m.halo + (size_t)slot * kHalo2;
  const int c = c0 + i;
  const size_t row = (size_t)m.cap_c;
  auto local = [&](int n) -> uint32_t {
    if (n < 0) return (uint32_t)(c - first);
    if (n >= first && n < first + count) return (uint32_t)(n - first);
    int lo = 0, hi = h;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (hl[mid] < n)
        lo = mid + 1;
      else
        hi = mid;
    }
    return (uint32_t)(kBlock + lo);
  };
  for (int k = 0; k < kNbk2 / 2; ++k)
    m.lnb[k * row + c] = (int32_t)(local(m.nbk[2 * k * row + c]) | (local(m.nbk[(2 * k + 1) * row + c]) << 16));
  // faces: the slot, bit 14 set when this cell is the face's cells[1], bit
  // 15 on the boundary (what the face-neighbour code holds, so the staged
  // kernel does not read that code)
  auto face = [&](int code) -> uint32_t {
    const int other = (code >> 1) - 1;
    return local(other) | ((uint32_t)(code & 1) << 14) | ((uint32_t)(other < 0) << 15);
  };
  for (int k = 0; k < 2; ++k)
    m.lfn[k * row + c] = (int32_t)(face(m.cfn[2 * k * row + c]) | (face(m.cfn[(2 * k + 1) * row + c]) << 16));
}

// ------------------------------------------------- packed driver records
// The per-field state into the state records (gs) of the problems that step
// in this batch of steps: no error latched, step budget left, and (when the
// loop stops at the final time) time left, the step flag the first step's
// time-step kernel will compute.  The decision is kept in the problem's
// scalars for the unpack.  Problems that do not step are not touched: a
// finished problem's ranges may already hold a new admission's state.
template <int N>
__global__ void __launch_bounds__(kBlock) k_eul_pack(DevMesh m, EulerFields u, Batch b) {
  for_each_tile(b.ct, [&](const Tile &tile) {
    const ProbParams &P = b.pp[tile.prob];
    DevScalars &S = b.s[tile.prob];
    const bool steps = S.err_code == 0 && S.steps < P.max_steps && (!P.stop_at_final || S.t < P.final_time);
    if (tile.first == P.cell_first && threadIdx.x == 0) S.packed = steps;
    if (!steps || (int)threadIdx.x >= tile.count) return;
    const int c = tile.first + threadIdx.x;
    double r[Rec<N>::S];
    r[0] = u.d[c];
    for (int k = 0; k < N; ++k) r[1 + k] = u.v[k][c];
    r[N + 1] = u.e[c];
    r[N + 2] = u.p[c];
    r[N + 3] = u.a[c];
    for (int d = 0; d < N; ++d) r[Rec<N>::C + d] = m.ccen[d][c];
    st_rec(u.gs, c, r);
    st_rec<Rec<N>::S, EulerFields::kPlanesS, Rec<N>::P_STATIC>(u.alt_gs, c, r);
  });
}

// ... and back when the batch of steps ends, for the problems packed; T = e
// / cv as the update computes it (update_state_from_energy,
// flecsale/eqns/euler_eqns.h:359-375).
template <int N>
__global__ void __launch_bounds__(kBlock) k_eul_unpack(EulerFields u, Batch b) {
  for_each_tile(b.ct, [&](const Tile &tile) {
    if (!b.s[tile.prob].packed || (int)threadIdx.x >= tile.count) return;
    const int c = tile.first + threadIdx.x;
    double r[Rec<N>::S];
    ld_rec(u.gs, c, r);
    u.d[c] = r[0];
    for (int k = 0; k < N; ++k) u.v[k][c] = r[1 + k];
    u.e[c] = r[N + 1];
    u.p[c] = r[N + 2];
    u.a[c] = r[N + 3];
    u.T[c] = eos_T_de(b.pp[tile.prob].cv, r[N + 1]);
  });
}

// -------------------------------------------------------------- fluxes
// Cell state + gradients + centroid, as gathered for a face reconstruction.
template <int N>
struct CellData {
  double d, v[N], p, e, a, gd[N], gv[N * N], ge[N], gp[N], ga[N], cx[N];
};

template <int N>
__device__ inline void load_cell(const DevMesh &m, const EulerFields &u, int c, CellData<N> &q) {
  q.d = u.d[c];
  for (int k = 0; k < N; ++k) q.v[k] = u.v[k][c];
  q.p = u.p[c];
  q.e = u.e[c];
  q.a = u.a[c];
  for (int i = 0; i < N; ++i) {
    q.gd[i] = u.gd[i][c];
    q.ge[i] = u.ge[i][c];
    q.gp[i] = u.gp[i][c];
    q.ga[i] = u.ga[i][c];
    q.cx[i] = m.ccen[i][c];
  }
  for (int i = 0; i < N * N; ++i) q.gv[i] = u.gv[i][c];
}

// the same from the driver loop's packed records: from memory, or (2D) from
// a tile staged in LDS, state planes first, then the gradient planes
template <int N, bool LDS = false>
__device__ inline void load_cell_pk(const EulerFields &u, int c, CellData<N> &q, const double2 *sh = nullptr) {
  double r[Rec<N>::S], g[Rec<N>::G];
  if constexpr (LDS) {
    lds_rec(sh, c, r);
    lds_rec(sh + (Rec<N>::S / 2) * kLdsW, c, g);
  } else {
    ld_rec(u.gs, c, r);
    ld_rec(u.gg, c, g);
  }
  q.d = r[0];
  for (int k = 0; k < N; ++k) q.v[k] = r[1 + k];
  q.e = r[N + 1];
  q.p = r[N + 2];
  q.a = r[N + 3];
  for (int i = 0; i < N; ++i) {
    q.cx[i] = r[Rec<N>::C + i];
    q.gd[i] = g[i];
    q.ge[i] = g[(N + 1) * N + i];
    q.gp[i] = g[(N + 2) * N + i];
    q.ga[i] = g[(N + 3) * N + i];
  }
  for (int i = 0; i < N * N; ++i) q.gv[i] = g[N + i];
}

// face state of apps/hydro/tasks.h:418-432: slot 2 (pressure) gets dx.grad_e
// and slot 3 (energy) gets dx.grad_p, as the reference's tuple order does.
template <int N>
__device__ inline void reconstruct(const CellData<N> &q, const double *fx, EState<N> &w) {
  double dx[N];
  for (int d = 0; d < N; ++d) dx[d] = fx[d] - q.cx[d];
  w.d = q.d;
  for (int k = 0; k < N; ++k) w.v[k] = q.v[k];
  w.p = q.p;
  w.e = q.e;
  w.a = q.a;
  for (int i = 0; i < N; ++i) {
    w.d += dx[i] * q.gd[i];
    for (int k = 0; k < N; ++k) w.v[k] += dx[i] * q.gv[k * N + i];
    w.p += dx[i] * q.ge[i];
    w.e += dx[i] * q.gp[i];
    w.a += dx[i] * q.ga[i];
  }
}

// numerical flux of one face times its area from the two face states (wr
// unused on a boundary face)
template <int N>
__device__ __forceinline__ void face_flux_w(int kind, const EState<N> &wl, const EState<N> &wr, bool interior,
                                            const double *n, double area, double *out) {
  if (interior)
    face_flux<N>(kind & 0xf, wl, wr, n, out);
  else if ((kind >> 4) == 1)
    e_flux<N>(wl, n, out);   // transmissive boundary
  else
    wall_flux<N>(wl, n, out);
  for (int i = 0; i < N + 2; ++i) out[i] *= area;
}

// numerical flux of one face times its area (flux(f) of evaluate_fluxes)
template <int N>
__device__ inline void face_flux_area(int kind, const CellData<N> &q0, const CellData<N> &q1, bool interior,
                                      const double *fx, const double *n, double area, double *out) {
  EState<N> wl, wr;
  reconstruct<N>(q0, fx, wl);
  if (interior) reconstruct<N>(q1, fx, wr);
  face_flux_w<N>(kind, wl, wr, interior, n, area, out);
}

// task-level evaluate_fluxes over the faces of one problem
template <int N>
__global__ void __launch_bounds__(kBlock) k_eul_flux(DevMesh m, EulerFields u, Batch b, int prob, int f0,
                                                     int nf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf || !wave_live(b.s, prob, 0)) return;
  const int f = f0 + i;
  const int kind = b.pp[prob].kind;
  double fx[N], n[N], out[N + 2];
  for (int d = 0; d < N; ++d) {
    fx[d] = m.fcen[d][f];
    n[d] = m.fnrm[d][f];
  }
  CellData<N> q0, q1;
  load_cell<N>(m, u, m.fc[f], q0);
  const int c1 = m.fc[(size_t)m.cap_f + f];
  load_cell<N>(m, u, c1 >= 0 ? c1 : m.fc[f], q1);
  face_flux_area<N>(kind, q0, q1, c1 >= 0, fx, n, m.farea[f], out);
  for (int j = 0; j < N + 2; ++j) u.F[j][f] = out[j];
}

// -------------------------------------------------------------- update
// update_state_from_flux euler_eqns.h:383-410 + update_state_from_energy
// :359-375 for one cell; returns false on negative density / energy.
// (den0, vel, ie0) -> (mass, vel, ie)
template <int N>
__device__ inline void update_values(double den0, double *vel, double ie0, double *du, double sc, double &mass,
                                     double &ie) {
  for (int i = 0; i < N + 2; ++i) du[i] *= sc;
  mass = den0 + du[0];
  const double et0 = ie0 + 0.5 * dotN<N>(vel, vel);
  const double ener = den0 * et0 + du[N + 1];
  const double inv_mass = 1 / mass;
  for (int d = 0; d < N; ++d) {
    const double mom = den0 * vel[d] + du[1 + d];
    vel[d] = mom * inv_mass;
  }
  ie = ener * inv_mass - 0.5 * dotN<N>(vel, vel);
}

template <int N>
__device__ inline bool update_cell(const EulerFields &in, const EulerFields &out, int c, double *du, double sc,
                                   double g, double cv, double *vel_out, double &a_out) {
  double vel[N], mass, ie;
  for (int d = 0; d < N; ++d) vel[d] = in.v[d][c];
  update_values<N>(in.d[c], vel, in.e[c], du, sc, mass, ie);
  out.d[c] = mass;
  for (int d = 0; d < N; ++d) {
    out.v[d][c] = vel[d];
    vel_out[d] = vel[d];
  }
  out.e[c] = ie;
  out.p[c] = eos_p_de(g, mass, ie);
  out.T[c] = eos_T_de(cv, ie);
  a_out = eos_a_de(g, ie);
  out.a[c] = a_out;
  return !(ie < 0 || mass < 0);   // apps/hydro/tasks.h:654-655
}

// task-level apply_update gathering the stored face fluxes
template <int N, int FPC>
__global__ void __launch_bounds__(kBlock) k_eul_update(DevMesh m, EulerFields u, Batch b, double factor,
                                                       int token) {
  for_each_tile(b.ct, [&](const Tile &tile) {
    if (!wave_live(b.s, tile.prob, token)) return;
    const ProbParams &P = b.pp[tile.prob];
    double delta_t = b.s[tile.prob].dt;
    delta_t *= factor;
    const int c = tile.first + threadIdx.x;
    if (c == P.cell_first) b.s[tile.prob].t = b.s[tile.prob].t + delta_t;   // update_soln_time :554
    if ((int)threadIdx.x >= tile.count) return;
    if (P.owned != 0 && m.corig[c] >= P.owned) return;   // mesh.cells(flecsi::owned) :614
    double du[N + 2];
    for (int i = 0; i < N + 2; ++i) du[i] = 0;
#pragma unroll
    for (int k = 0; k < FPC; ++k) {
      const int code = m.cf[(size_t)k * m.cap_c + c];
      const int f = code >> 1;
      if (code & 1) {
        for (int i = 0; i < N + 2; ++i) du[i] += u.F[i][f];
      } else {
        for (int i = 0; i < N + 2; ++i) du[i] -= u.F[i][f];
      }
    }
    double vel[N], a;
    if (!update_cell<N>(u, u, c, du, delta_t / m.cvol[c], P.gamma, P.cv, vel, a))
      flag_negative(b.s, tile.prob, m.corig[c], token);
  });
}

// What a 2D staged tile's cell reads from memory besides its tile's records:
// its LDS slot words, its vertices' coordinates and its volume, loaded before
// the tile is staged so that their round trips overlap the staging's.
struct Pre2 {
  uint32_t w[2];
  double px[4], py[4], vol;
};
__device__ __forceinline__ void preload_2d(const DevMesh &m, const Tile &tile, Pre2 &q) {
  const int c = tile.first + ((int)threadIdx.x < tile.count ? (int)threadIdx.x : 0);
#pragma unroll
  for (int k = 0; k < 2; ++k) q.w[k] = (uint32_t)m.lfn[(size_t)k * m.cap_c + c];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double2 w = m.xy[m.cv[(size_t)k * m.cap_c + c]];
    q.px[k] = w.x;
    q.py[k] = w.y;
  }
  q.vol = m.cvol[c];
}

// fused evaluate_fluxes + apply_update (+ dt^-1 partials of the new state)
// of one tile over the packed records: reads u.gs and u.gg, writes out.gs;
// LDS: from the tile's records staged in sh, with the cell's other reads
// preloaded in *pre (2D quads).  FP (3D): a problem whose hex sides pair up
// (ProbParams::s_fp) reads the geometry of a cell's -x, -y and -z sides from
// the +x, +y and +z slots of the neighbours across them (the same faces):
// each face's geometry is then read from memory once, by its lower cell, and
// again from cache by the upper one.
template <int N, int FPC, int VPC, bool LDS, bool FP = false>
__device__ __forceinline__ void fluxupd_tile(const DevMesh &m, const EulerFields &u, const EulerFields &out, const Batch &b,
                                    double factor, int want_dt, int token, const Tile &tile, const double2 *sh,
                                    const Pre2 *pre = nullptr) {
  static_assert(!LDS || (N == 2 && FPC == 4 && VPC == 4), "staged tiles are 2D quads");
  if (!wave_live(b.s, tile.prob, token)) return;
  const ProbParams &P = b.pp[tile.prob];
  const int kind = P.kind;
  double delta_t = b.s[tile.prob].dt;
  delta_t *= factor;
  const int c = tile.first + threadIdx.x;
  if (c == P.cell_first) b.s[tile.prob].t = b.s[tile.prob].t + delta_t;   // update_soln_time :554
  double dt_inv = 0;
  if ((int)threadIdx.x < tile.count) {
    // the face-neighbour codes first: each face's neighbour loads then
    // wait on one round trip, not two
    // LDS: the face-neighbour codes' orientation and boundary bits ride in
    // the LDS slot words (k_lds_slots)
    int cnk[FPC], slot[FPC];
    if constexpr (LDS) {
#pragma unroll
      for (int k = 0; k < FPC / 2; ++k) {
        const uint32_t w = pre->w[k];
        slot[2 * k] = (int)(w & 0x3fffu);
        slot[2 * k + 1] = (int)((w >> 16) & 0x3fffu);
        cnk[2 * k] = (int)((w >> 14) & 1u) | ((w & 0x8000u) ? 0 : 2);
        cnk[2 * k + 1] = (int)((w >> 30) & 1u) | ((w & 0x80000000u) ? 0 : 2);
      }
    } else {
#pragma unroll
      for (int k = 0; k < FPC; ++k) cnk[k] = m.cfn[(size_t)k * m.cap_c + c];
    }
    CellData<N> own;
    if constexpr (LDS)
      load_cell_pk<N, true>(u, (int)threadIdx.x, own, sh);
    else
      load_cell_pk<N>(u, c, own);
    double px[VPC], py[VPC];
    if constexpr (LDS) {
      for (int k = 0; k < VPC; ++k) {
        px[k] = pre->px[k];
        py[k] = pre->py[k];
      }
    } else if constexpr (N == 2) {
      for (int k = 0; k < VPC; ++k) {
        const double2 w = m.xy[m.cv[(size_t)k * m.cap_c + c]];
        px[k] = w.x;
        py[k] = w.y;
      }
    }
    double du[N + 2];
    for (int i = 0; i < N + 2; ++i) du[i] = 0;
    double nk[FPC][N], ak[FPC];
#pragma unroll
    for (int k = 0; k < FPC; ++k) {
      const int cn = cnk[k];
      const bool second = cn & 1;     // this cell is the face's cells[1]
      const int other = (cn >> 1) - 1;
      double fx[N];
      if constexpr (N == 2) {
        const int k1 = (k + 1 == VPC) ? 0 : k + 1;
        // the face runs v_k -> v_k1 out of cells[0], reversed for cells[1]
        if (!second)
          face_geom_2d(px[k], py[k], px[k1], py[k1], fx[0], fx[1], nk[k][0], nk[k][1], ak[k]);
        else
          face_geom_2d(px[k1], py[k1], px[k], py[k], fx[0], fx[1], nk[k][0], nk[k][1], ak[k]);
      } else {
        double2 w0, w1, w2;
        if constexpr (FP) {
          // kHexFaces sides: -x 3 / +x 1, -y 0 / +y 2, -z 4 / +z 5
          const int up = k == 3 ? 1 : k == 0 ? 2 : k == 4 ? 5 : -1;
          const bool pair = up >= 0 && P.s_fp && other >= 0;
          const size_t row = (size_t)(pair ? up : k) * m.cap_c;
          const uint32_t o = (uint32_t)(pair ? other : c) * 8u;
          w0 = ld(m.kfp[0] + row, 2 * o);
          w1 = ld(m.kfp[1] + row, 2 * o);
          w2 = ld(m.kfp[2] + row, 2 * o);
          ak[k] = ld(m.kfa + row, o);
        } else {
          const size_t row = (size_t)k * m.cap_c;
          const uint32_t o = (uint32_t)c * 8u;
#if FSALE_NT3
          w0 = ld_nt(m.kfp[0] + row, 2 * o);
          w1 = ld_nt(m.kfp[1] + row, 2 * o);
          w2 = ld_nt(m.kfp[2] + row, 2 * o);
          ak[k] = ld_nt(m.kfa + row, o);
#else
          w0 = ld(m.kfp[0] + row, 2 * o);
          w1 = ld(m.kfp[1] + row, 2 * o);
          w2 = ld(m.kfp[2] + row, 2 * o);
          ak[k] = ld(m.kfa + row, o);
#endif
        }
        fx[0] = w0.x;
        fx[1] = w0.y;
        fx[2] = w1.x;
        nk[k][0] = w1.y;
        nk[k][1] = w2.x;
        nk[k][2] = w2.y;
      }
      // the neighbour's loads do not wait on a branch (a boundary face
      // loads the cell's own data again and ignores it)
      CellData<N> nb;
      if constexpr (LDS)
        load_cell_pk<N, true>(u, slot[k], nb, sh);
      else
        load_cell_pk<N>(u, other >= 0 ? other : c, nb);
      double fl[N + 2];
      if (!second)
        face_flux_area<N>(kind, own, nb, other >= 0, fx, nk[k], ak[k], fl);
      else
        face_flux_area<N>(kind, nb, own, true, fx, nk[k], ak[k], fl);
      if (second) {
        for (int i = 0; i < N + 2; ++i) du[i] += fl[i];
      } else {
        for (int i = 0; i < N + 2; ++i) du[i] -= fl[i];
      }
    }
    double vol;
    if constexpr (LDS)
      vol = pre->vol;
    else
      vol = m.cvol[c];
    double vel[N], mass, ie;
    for (int d = 0; d < N; ++d) vel[d] = own.v[d];
    update_values<N>(own.d, vel, own.e, du, delta_t / vol, mass, ie);
    const double a = eos_a_de(P.gamma, ie);
    {
      double r[Rec<N>::S];
      r[0] = mass;
      for (int d = 0; d < N; ++d) r[1 + d] = vel[d];
      r[N + 1] = ie;
      r[N + 2] = eos_p_de(P.gamma, mass, ie);
      r[N + 3] = a;
      for (int d = 0; d < N; ++d) r[Rec<N>::C + d] = own.cx[d];
#if FSALE_NT3
      if constexpr (N == 3) {
        const uint32_t off = (uint32_t)c * 16u;
#pragma unroll
        for (int i = 0; i < Rec<N>::P_STATIC; ++i) st_nt(out.gs[i], off, make_double2(r[2 * i], r[2 * i + 1]));
      } else
#endif
      st_rec<Rec<N>::S, EulerFields::kPlanesS, 0, Rec<N>::P_STATIC>(out.gs, c, r);   // the centroid planes stay
    }
    if (ie < 0 || mass < 0)   // apps/hydro/tasks.h:654-655
      flag_negative(b.s, tile.prob, m.corig[c], token);
    if (want_dt) {
#pragma unroll
      for (int k = 0; k < FPC; ++k) {
        const double delta_x = vol / ak[k];
        const double dti = (a + fabs(dotN<N>(vel, nk[k]))) / delta_x;
        dt_inv = smax(dti, dt_inv);
      }
    }
  }
  if (want_dt) {
    dt_inv = wave_max(dt_inv);
    if ((threadIdx.x & 63) == 0) u.partials[tile.slot * kWaves + (threadIdx.x >> 6)] = dt_inv;
  }
}

// LDS (2D quads with staging plans): a tile with a plan is staged in LDS
// and its face neighbours read there (DevMesh::lfn); a tile without one
// gathers from memory.  The choice is per tile, uniform across the block.
// ALL: every tile of the launch has a plan (the memory path not compiled in).
template <int N, int FPC, int VPC, bool LDS, bool ALL = false, bool FP = false>
__global__ void __launch_bounds__(kBlock) k_eul_fluxupd(DevMesh m, EulerFields u, EulerFields out, Batch b,
                                                        double factor, int want_dt, int token) {
  static_assert(!LDS || N == 2, "LDS staging is 2D");
  constexpr int NP = (Rec<N>::S + Rec<N>::G) / 2;
  __shared__ double2 sh[LDS ? NP * kLdsW : 1];
  for_each_tile(b.ct, [&](const Tile &tile) {
    const int h = LDS ? m.halo_cnt[tile.slot] : -1;
    if (ALL || h >= 0) {
      if constexpr (LDS) {
        const int hcell = halo_entry(m, tile, h);
        Pre2 pre;
        preload_2d(m, tile, pre);
        stage_tile<NP, true>(sh, m, tile, h, hcell,
                             [&](int i) { return i < Rec<N>::S / 2 ? u.gs[i] : u.gg[i - Rec<N>::S / 2]; },
                             b.ct.direct == nullptr);
        fluxupd_tile<N, FPC, VPC, true>(m, u, out, b, factor, want_dt, token, tile, sh, &pre);
      }
    } else {
      fluxupd_tile<N, FPC, VPC, false, FP>(m, u, out, b, factor, want_dt, token, tile, sh);
    }
  });
}

// ------------------------------------------- 3D structured grids in bricks
// A problem on a structured hex grid whose sides pair up (ProbParams::s_fp)
// with nx, ny multiples of 8 and nz of 4 runs the fused step over 8 x 8 x 4
// bricks of cells, one per tile slot (DevMesh::brick: the brick's first
// cell).  Each cell reconstructs its own state at its six sides from its own
// record; the bricks' neighbours across the brick's sides (256 cells, one per
// thread) reconstruct theirs at the side they share with the brick.  Those
// face states meet in LDS, one direction at a time: a cell gathers one
// record (its own) and one halo record instead of its six neighbours'.  Each
// face state and each flux is formed by the same operations, on the same
// values, as the gathering kernel forms it, and the fluxes are summed in the
// cell's side order, so the results are the same bits.
constexpr int kBrX = kBrickX, kBrY = kBrickY, kBrZ = kBrickZ;
static_assert(kBrX == 8 && kBrY == 8 && kBrZ == 4, "the thread -> cell and halo maps assume 8 x 8 x 4");
constexpr int kBrW = kBlock;   // LDS plane width: slots of the brick's cells, or of its 256 halo cells

// a face state in LDS: 4 planes of pairs (the 8th double unused)
__device__ __forceinline__ void put_face_state(double2 *sh, int slot, const EState<3> &w) {
  sh[slot] = make_double2(w.d, w.v[0]);
  sh[kBrW + slot] = make_double2(w.v[1], w.v[2]);
  sh[2 * kBrW + slot] = make_double2(w.p, w.e);
  sh[3 * kBrW + slot].x = w.a;
}
__device__ __forceinline__ void get_face_state(const double2 *sh, int slot, EState<3> &w) {
  const double2 a = sh[slot], b = sh[kBrW + slot], c = sh[2 * kBrW + slot];
  w.d = a.x;
  w.v[0] = a.y;
  w.v[1] = b.x;
  w.v[2] = b.y;
  w.p = c.x;
  w.e = c.y;
  w.a = sh[3 * kBrW + slot].x;
}

// Geometry of side k of cell c (the neighbour across it: other, or -1): a
// -x, -y or -z side with a neighbour reads the neighbour's +x, +y, +z slot,
// the same face (kHexFaces sides: -x 3 / +x 1, -y 0 / +y 2, -z 4 / +z 5).
__device__ __forceinline__ void side_geom(const DevMesh &m, int k, int c, int other, double *fx, double *n,
                                          double &area) {
  const int up = k == 3 ? 1 : k == 0 ? 2 : k == 4 ? 5 : -1;
  const bool pair = up >= 0 && other >= 0;
  const size_t row = (size_t)(pair ? up : k) * m.cap_c;
  const uint32_t o = (uint32_t)(pair ? other : c) * 8u;
  const double2 w0 = ld(m.kfp[0] + row, 2 * o), w1 = ld(m.kfp[1] + row, 2 * o), w2 = ld(m.kfp[2] + row, 2 * o);
  area = ld(m.kfa + row, o);
  fx[0] = w0.x;
  fx[1] = w0.y;
  fx[2] = w1.x;
  n[0] = w1.y;
  n[1] = w2.x;
  n[2] = w2.y;
}

// One brick.  LDS: hal (the halo cells' face states, slot = thread), slo /
// shi (the brick's cells' states at their low / high side of the current
// direction).  All threads of the block take part in every barrier; a brick
// of a problem that does not step returns as a whole (the step flag and the
// error token are the same for every wave of the launch).
__device__ __forceinline__ void fluxupd_brick(const DevMesh &m, const EulerFields &u, const EulerFields &out,
                                              const Batch &b, double factor, int want_dt, int token,
                                              const Tile &tile, double2 *hal, double2 *slo, double2 *shi) {
  constexpr int N = 3;
  if (!wave_live(b.s, tile.prob, token)) return;
  const ProbParams &P = b.pp[tile.prob];
  const int kind = P.kind;
  const int nx = P.s_nx, np = P.s_np;
  const int base = m.brick[tile.slot];
  // the brick's place in the grid
  const int lb = base - P.cell_first;
  const int bz = lb / np, by = (lb - bz * np) / nx, bx = lb - bz * np - by * nx;
  const int ny = np / nx, nz = P.cell_count / np;
  const int t = threadIdx.x;
  const int ti = t & 7, tj = (t >> 3) & 7, tk = t >> 6;
  const int c = base + ti + tj * nx + tk * np;
  double delta_t = b.s[tile.prob].dt;
  delta_t *= factor;
  if (c == P.cell_first) b.s[tile.prob].t = b.s[tile.prob].t + delta_t;   // update_soln_time :554

  // the halo: thread t's cell outside the brick and the side it shares with
  // it (t < 64: below in z, < 128: above, < 160: -x, < 192: +x, < 224: -y,
  // else +y), its face state written to hal[t]
  {
    int hc, gk, gc;   // halo cell, the shared face's slot: side gk of cell gc
    bool in;
    if (t < 128) {
      const int i = t & 7, j = (t >> 3) & 7;
      if (t < 64) {
        hc = base + i + j * nx - np;
        in = bz > 0;
        gk = 5;
        gc = hc;
      } else {
        hc = base + i + j * nx + kBrZ * np;
        in = bz + kBrZ < nz;
        gk = 5;
        gc = hc - np;
      }
    } else if (t < 192) {
      const int j = t & 7, k = (t >> 3) & 3;
      if (t < 160) {
        hc = base - 1 + j * nx + k * np;
        in = bx > 0;
        gk = 1;
        gc = hc;
      } else {
        hc = base + kBrX + j * nx + k * np;
        in = bx + kBrX < nx;
        gk = 1;
        gc = hc - 1;
      }
    } else {
      const int i = t & 7, k = (t >> 3) & 3;
      if (t < 224) {
        hc = base + i - nx + k * np;
        in = by > 0;
        gk = 2;
        gc = hc;
      } else {
        hc = base + i + kBrY * nx + k * np;
        in = by + kBrY < ny;
        gk = 2;
        gc = hc - nx;
      }
    }
    if (in) {
      CellData<N> q;
      load_cell_pk<N>(u, hc, q);
      const size_t row = (size_t)gk * m.cap_c;
      const uint32_t o = (uint32_t)gc * 16u;
      const double2 w0 = ld(m.kfp[0] + row, o), w1 = ld(m.kfp[1] + row, o);
      const double fx[3] = {w0.x, w0.y, w1.x};
      EState<N> w;
      reconstruct<N>(q, fx, w);
      put_face_state(hal, t, w);
    }
  }

  int cn[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) cn[k] = m.cfn[(size_t)k * m.cap_c + c];
  CellData<N> own;
  load_cell_pk<N>(u, c, own);
  const double vol = m.cvol[c];
  double du[N + 2];
  for (int i = 0; i < N + 2; ++i) du[i] = 0;
  double nk[6][N], ak[6], f2[N + 2];

  // one direction: sides lo and hi of each cell, neighbours in the brick at
  // thread t -/+ st while the cell's place along it (pos) is inside, else in
  // the halo at slots hlo / hhi.  The fluxes of the two sides, in fl / fh.
  auto pass = [&](int klo, int khi, int pos, int last, int st, int hlo, int hhi, double *fl, double *fh) {
    double fxl[N], fxh[N];
    const int ol = (cn[klo] >> 1) - 1, oh = (cn[khi] >> 1) - 1;
    side_geom(m, klo, c, ol, fxl, nk[klo], ak[klo]);
    side_geom(m, khi, c, oh, fxh, nk[khi], ak[khi]);
    EState<N> wl, wh;
    reconstruct<N>(own, fxl, wl);
    reconstruct<N>(own, fxh, wh);
    __syncthreads();   // the previous direction's reads are done
    put_face_state(slo, t, wl);
    put_face_state(shi, t, wh);
    __syncthreads();
    {
      EState<N> wn;
      get_face_state(pos > 0 ? shi : hal, pos > 0 ? t - st : hlo, wn);
      if (cn[klo] & 1)
        face_flux_w<N>(kind, wn, wl, true, nk[klo], ak[klo], fl);
      else
        face_flux_w<N>(kind, wl, wn, ol >= 0, nk[klo], ak[klo], fl);
    }
    {
      EState<N> wn;
      get_face_state(pos < last ? slo : hal, pos < last ? t + st : hhi, wn);
      if (cn[khi] & 1)
        face_flux_w<N>(kind, wn, wh, true, nk[khi], ak[khi], fh);
      else
        face_flux_w<N>(kind, wh, wn, oh >= 0, nk[khi], ak[khi], fh);
    }
  };
  auto acc = [&](int k, const double *fl) {
    if (cn[k] & 1) {
      for (int i = 0; i < N + 2; ++i) du[i] += fl[i];
    } else {
      for (int i = 0; i < N + 2; ++i) du[i] -= fl[i];
    }
  };
  {
    // y: sides 0 (-y) and 2 (+y); side 0 is the first of the sum
    double f0[N + 2];
    pass(0, 2, tj, kBrY - 1, kBrX, 192 + ti + 8 * tk, 224 + ti + 8 * tk, f0, f2);
    acc(0, f0);
  }
  {
    // x: sides 3 (-x) and 1 (+x)
    double f3[N + 2], f1[N + 2];
    pass(3, 1, ti, kBrX - 1, 1, 128 + tj + 8 * tk, 160 + tj + 8 * tk, f3, f1);
    acc(1, f1);
    acc(2, f2);
    acc(3, f3);
  }
  {
    // z: sides 4 (-z) and 5 (+z)
    double f4[N + 2], f5[N + 2];
    pass(4, 5, tk, kBrZ - 1, kBrX * kBrY, ti + 8 * tj, 64 + ti + 8 * tj, f4, f5);
    acc(4, f4);
    acc(5, f5);
  }

  double vel[N], mass, ie;
  for (int d = 0; d < N; ++d) vel[d] = own.v[d];
  update_values<N>(own.d, vel, own.e, du, delta_t / vol, mass, ie);
  const double a = eos_a_de(P.gamma, ie);
  {
    double r[Rec<N>::S];
    r[0] = mass;
    for (int d = 0; d < N; ++d) r[1 + d] = vel[d];
    r[N + 1] = ie;
    r[N + 2] = eos_p_de(P.gamma, mass, ie);
    r[N + 3] = a;
    for (int d = 0; d < N; ++d) r[Rec<N>::C + d] = own.cx[d];
    const uint32_t off = (uint32_t)c * 16u;
#pragma unroll
    for (int i = 0; i < Rec<N>::P_STATIC; ++i) st_nt(out.gs[i], off, make_double2(r[2 * i], r[2 * i + 1]));
  }
  if (ie < 0 || mass < 0)   // apps/hydro/tasks.h:654-655
    flag_negative(b.s, tile.prob, m.corig[c], token);
  if (want_dt) {
    double dt_inv = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double delta_x = vol / ak[k];
      const double dti = (a + fabs(dotN<N>(vel, nk[k]))) / delta_x;
      dt_inv = smax(dti, dt_inv);
    }
    dt_inv = wave_max(dt_inv);
    if ((t & 63) == 0) u.partials[tile.slot * kWaves + (t >> 6)] = dt_inv;
  }
}

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
k_eul_fluxupd_brick(DevMesh m, EulerFields u, EulerFields out, Batch b, double factor, int want_dt, int token) {
  __shared__ double2 sh[3 * 4 * kBrW];
  for_each_tile(b.ct, [&](const Tile &tile) {
    __syncthreads();   // a persistent block's previous brick is done with the LDS
    fluxupd_brick(m, u, out, b, factor, want_dt, token, tile, sh, sh + 4 * kBrW, sh + 8 * kBrW);
  });
}

// ------------------------------------------------------------- launches
template <class K, class... A>
int launch_persistent(fsale_ctx *ctx, int kid, K kernel, A... args) {
  ProfScope prof(ctx, kid);
  hipLaunchKernelGGL(kernel, dim3(persistent_grid(ctx, kid)), dim3(kBlock), 0, ctx->stream, args...);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

// the same fields with the two state buffers exchanged
EulerFields swapped(const EulerFields &a) {
  EulerFields b = a;
  for (int i = 0; i < EulerFields::kPlanesS; ++i) {
    b.gs[i] = a.alt_gs[i];
    b.alt_gs[i] = a.gs[i];
  }
  return b;
}

int launch_dt_partial(fsale_ctx *ctx) {
  const Batch b = batch_args(ctx, K_EUL_DT);
  if (ctx->m.dim == 2) return launch_persistent(ctx, K_EUL_DT, k_eul_dt_partial<2, 4>, ctx->m, ctx->eu, b);
  return launch_persistent(ctx, K_EUL_DT, k_eul_dt_partial<3, 6>, ctx->m, ctx->eu, b);
}

int launch_finalize(fsale_ctx *ctx, int driver) {
  ProfScope prof(ctx, K_EUL_DT);
  hipLaunchKernelGGL(k_eul_dt_finalize, dim3(ctx->pgrid), dim3(kFinalBlock), 0, ctx->stream, batch_args(ctx, -1),
                     (const double *)ctx->eu.partials, driver);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int launch_grad(fsale_ctx *ctx, const EulerFields &u) {
  const Batch b = batch_args(ctx, K_EUL_GRAD);
  if (ctx->m.dim == 2) return launch_persistent(ctx, K_EUL_GRAD, k_eul_grad<2, 4, 0>, ctx->m, u, b);
  return launch_persistent(ctx, K_EUL_GRAD, k_eul_grad<3, 8, 0>, ctx->m, u, b);
}

// One block per cell tile whatever the kernel's launch shape: the brick
// kernels measured slower on every persistent grid (gradients 338-383 us
// against 259, fused step 380-456 against 319 at 128^3), and the tuner's
// few timed steps have picked one now and then.
template <class K, class... A>
int launch_direct(fsale_ctx *ctx, int kid, K kernel, const DevMesh &m, const EulerFields &u, Batch b, A... args) {
  ProfScope prof(ctx, kid);
  b.ct.direct = ctx->d_cdirect;
  b.ct.ndirect = ctx->n_cdirect;
  hipLaunchKernelGGL(kernel, dim3(std::max(8, ctx->n_cdirect)), dim3(kBlock), 0, ctx->stream, m, u, b, args...);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

// fin: the step's first gradient launch, which also finalizes its time step
template <bool FIN>
int launch_grad_pk(fsale_ctx *ctx, const EulerFields &u) {
  const Batch b = batch_args(ctx, K_EUL_GRAD);
  if (ctx->m.dim == 2 && ctx->lds_all)
    return launch_persistent(ctx, K_EUL_GRAD, k_eul_grad_lds<4, true, FIN>, ctx->m, u, b);
  if (ctx->m.dim == 2 && ctx->lds)
    return launch_persistent(ctx, K_EUL_GRAD, k_eul_grad_lds<4, false, FIN>, ctx->m, u, b);
  if (ctx->m.dim == 2) return launch_persistent(ctx, K_EUL_GRAD, k_eul_grad<2, 4, 1, FIN>, ctx->m, u, b);
  if (ctx->brick_all) return launch_direct(ctx, K_EUL_GRAD, k_eul_grad_brick<FIN>, ctx->m, u, b);
  if (ctx->s3_all) return launch_persistent(ctx, K_EUL_GRAD, k_eul_grad_s3<true, FIN>, ctx->m, u, b);
  if (ctx->s3) return launch_persistent(ctx, K_EUL_GRAD, k_eul_grad_s3<false, FIN>, ctx->m, u, b);
  return launch_persistent(ctx, K_EUL_GRAD, k_eul_grad<3, 8, 1, FIN>, ctx->m, u, b);
}

// token: the launch's id for error attribution (first half step 1, second 2)
int launch_fluxupd(fsale_ctx *ctx, double factor, bool want_dt, int token, const EulerFields &in,
                   const EulerFields &out) {
  const Batch b = batch_args(ctx, K_EUL_FLUXUPD);
  if (ctx->m.dim == 2 && ctx->lds_all)
    return launch_persistent(ctx, K_EUL_FLUXUPD, k_eul_fluxupd<2, 4, 4, true, true>, ctx->m, in, out, b, factor,
                             (int)want_dt, token);
  if (ctx->m.dim == 2 && ctx->lds)
    return launch_persistent(ctx, K_EUL_FLUXUPD, k_eul_fluxupd<2, 4, 4, true>, ctx->m, in, out, b, factor,
                             (int)want_dt, token);
  if (ctx->m.dim == 2)
    return launch_persistent(ctx, K_EUL_FLUXUPD, k_eul_fluxupd<2, 4, 4, false>, ctx->m, in, out, b, factor,
                             (int)want_dt, token);
  if (ctx->brick_all) {
    // one block per brick, as launch_direct
    ProfScope prof(ctx, K_EUL_FLUXUPD);
    Batch bd = b;
    bd.ct.direct = ctx->d_cdirect;
    bd.ct.ndirect = ctx->n_cdirect;
    hipLaunchKernelGGL(k_eul_fluxupd_brick, dim3(std::max(8, ctx->n_cdirect)), dim3(kBlock), 0, ctx->stream, ctx->m,
                       in, out, bd, factor, (int)want_dt, token);
    FSALE_HIP(hipGetLastError());
    return FSALE_OK;
  }
  if (ctx->fp3)
    return launch_persistent(ctx, K_EUL_FLUXUPD, k_eul_fluxupd<3, 6, 8, false, false, true>, ctx->m, in, out, b,
                             factor, (int)want_dt, token);
  return launch_persistent(ctx, K_EUL_FLUXUPD, k_eul_fluxupd<3, 6, 8, false>, ctx->m, in, out, b, factor,
                           (int)want_dt, token);
}

bool check_ctx(fsale_ctx *ctx) {
  if (!ctx || ctx->solver != FSALE_SOLVER_EULER || !ctx->single || ctx->probs.empty() || !ctx->probs[0].live) {
    set_message("not an Eulerian context");
    return false;
  }
  return true;
}

bool valid_kind(int k) { return k >= 0 && (k & 0xf) <= 2 && (k >> 4) <= 1; }

}  // namespace

// One driver step (apps/hydro/driver.h:317-413) of every live problem with
// the fused kernels: the first half step reads the primary state buffer and
// writes the second one, the second half step writes back, so the state is
// in the primary buffer at every step boundary.  The dt^-1 partials of the
// current state must be valid on entry and are valid again on exit.
int euler_step(fsale_ctx *ctx) {
  int rc;
  const EulerFields a = ctx->eu, b = swapped(ctx->eu);
  if ((rc = launch_grad_pk<true>(ctx, a))) return rc;   // + the step's time step
  if ((rc = launch_fluxupd(ctx, 0.5, false, 1, a, b))) return rc;
  if ((rc = launch_grad_pk<false>(ctx, b))) return rc;
  if ((rc = launch_fluxupd(ctx, 0.5, true, 2, b, a))) return rc;
  return FSALE_OK;
}

int euler_dt_partials(fsale_ctx *ctx) { return launch_dt_partial(ctx); }

// Around a batch of driver steps: the per-field state of the problems that
// step into the packed records, and back.
int euler_lds_slots(fsale_ctx *ctx, int prob, hipStream_t stream) {
  const Problem &P = ctx->probs[prob];
  if (!ctx->m.lnb || P.nc == 0) return FSALE_OK;
  hipLaunchKernelGGL(k_lds_slots, dim3((P.nc + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, ctx->m, P.c0, P.nc,
                     P.ct0);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int euler_pack(fsale_ctx *ctx) {
  const Batch b = batch_args(ctx, -1);
  const dim3 grid(persistent_grid(ctx, -1)), block(kBlock);
  if (ctx->m.dim == 2)
    hipLaunchKernelGGL(k_eul_pack<2>, grid, block, 0, ctx->stream, ctx->m, ctx->eu, b);
  else
    hipLaunchKernelGGL(k_eul_pack<3>, grid, block, 0, ctx->stream, ctx->m, ctx->eu, b);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int euler_unpack(fsale_ctx *ctx) {
  const Batch b = batch_args(ctx, -1);
  const dim3 grid(persistent_grid(ctx, -1)), block(kBlock);
  if (ctx->m.dim == 2)
    hipLaunchKernelGGL(k_eul_unpack<2>, grid, block, 0, ctx->stream, ctx->eu, b);
  else
    hipLaunchKernelGGL(k_eul_unpack<3>, grid, block, 0, ctx->stream, ctx->eu, b);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int euler_alloc(fsale_ctx *ctx) {
  const DevMesh &m = ctx->m;
  EulerFields &u = ctx->eu;
  const int N = m.dim;
  int rc;
  double **cell1[] = {&u.d, &u.e, &u.p, &u.T, &u.a};
  for (double **p : cell1)
    if ((rc = dev_alloc_n(ctx, p, m.cap_c))) return rc;
  const int ps = (N == 2 ? Rec<2>::S : Rec<3>::S) / 2, pg = (N == 2 ? Rec<2>::G : Rec<3>::G) / 2;
  for (int i = 0; i < ps; ++i) {
    if ((rc = dev_alloc_n(ctx, &u.gs[i], m.cap_c))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.alt_gs[i], m.cap_c))) return rc;
  }
  for (int i = 0; i < pg; ++i)
    if ((rc = dev_alloc_n(ctx, &u.gg[i], m.cap_c))) return rc;
  for (int d = 0; d < N; ++d) {
    if ((rc = dev_alloc_n(ctx, &u.v[d], m.cap_c))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.gd[d], m.cap_c))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.ge[d], m.cap_c))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.gp[d], m.cap_c))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.ga[d], m.cap_c))) return rc;
  }
  for (int d = 0; d < N * N; ++d)
    if ((rc = dev_alloc_n(ctx, &u.gv[d], m.cap_c))) return rc;
  if (ctx->single)   // the face flux array serves the task-level evaluate_fluxes only
    for (int i = 0; i < N + 2; ++i)
      if ((rc = dev_alloc_n(ctx, &u.F[i], m.cap_f))) return rc;
  if ((rc = dev_alloc_n(ctx, &u.partials, (size_t)kWaves * (ctx->ra_ct.capacity + 1)))) return rc;
  return FSALE_OK;
}

// initial conditions of one problem: d, v, p already in place
int euler_initial_state(fsale_ctx *ctx, int prob, hipStream_t stream) {
  const Problem &P = ctx->probs[prob];
  const int g = (P.nc + kBlock - 1) / kBlock;
  if (g == 0) return FSALE_OK;
  hipLaunchKernelGGL(k_eul_ic, dim3(g), dim3(kBlock), 0, stream, ctx->eu, (const ProbParams *)ctx->d_params, prob);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

}  // namespace fsale

using namespace fsale;

namespace fsale {
int set_problem_state(fsale_ctx *ctx, int prob, const double *d, const double *v, const double *p);
int reset_scalars(fsale_ctx *ctx, int prob, double t, double dt);
int get_problem_field(fsale_ctx *ctx, int prob, int field, double *out);
}  // namespace fsale

extern "C" {

int fsale_eul_initial_conditions(fsale_ctx *ctx, fsale_eos eos, const double *d, const double *v,
                                 const double *p) {
  if (!check_ctx(ctx) || !d || !v || !p) return FSALE_ERR_ARG;
  if (!(eos.gamma > 1) || !(eos.cv > 0)) return FSALE_ERR_ARG;
  Problem &P = ctx->probs[0];
  P.params.gamma = eos.gamma;
  P.params.cv = eos.cv;
  int rc;
  if ((rc = upload_params(ctx, 0))) return rc;
  if ((rc = set_problem_state(ctx, 0, d, v, p))) return rc;
  if ((rc = euler_initial_state(ctx, 0, ctx->stream))) return rc;
  P.have_state = true;
  P.partials_valid = false;
  return reset_scalars(ctx, 0, 0.0, 0.0);
}

int fsale_eul_evaluate_time_step(fsale_ctx *ctx, double cfl, double final_time, double *dt_out) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  Problem &P = ctx->probs[0];
  if (!P.have_state) return FSALE_ERR_STATE;
  P.params.cfl = cfl;
  P.params.final_time = final_time;
  int rc;
  if ((rc = upload_params(ctx, 0))) return rc;
  if ((rc = sync_tiles(ctx)) || (rc = open_tasks(ctx))) return rc;
  if ((rc = launch_dt_partial(ctx))) return rc;
  if ((rc = launch_finalize(ctx, 0))) return rc;
  P.partials_valid = true;
  if (!dt_out) return FSALE_OK;
  if ((rc = check_device_error(ctx, 0))) return rc;
  *dt_out = ctx->h_s[0].dt;
  return FSALE_OK;
}

int fsale_eul_set_owned_cells(fsale_ctx *ctx, int32_t num_owned) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  Problem &P = ctx->probs[0];
  if (num_owned <= 0 || num_owned > P.nc) {
    set_message("the owned cells are the first 1..num_cells cells");
    return FSALE_ERR_ARG;
  }
  P.params.owned = num_owned == P.nc ? 0 : num_owned;
  P.partials_valid = false;
  return upload_params(ctx, 0);
}

int fsale_eul_evaluate_gradients(fsale_ctx *ctx, int first_order) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  Problem &P = ctx->probs[0];
  if (!P.have_state) return FSALE_ERR_STATE;
  int rc;
  P.params.first_order = first_order != 0;
  if ((rc = upload_params(ctx, 0))) return rc;
  if ((rc = sync_tiles(ctx)) || (rc = open_tasks(ctx))) return rc;
  return launch_grad(ctx, ctx->eu);
}

int fsale_eul_evaluate_fluxes(fsale_ctx *ctx, int flux_kind) {
  if (!check_ctx(ctx) || !valid_kind(flux_kind)) return FSALE_ERR_ARG;
  Problem &P = ctx->probs[0];
  if (!P.have_state) return FSALE_ERR_STATE;
  P.params.kind = flux_kind;
  int rc;
  if ((rc = upload_params(ctx, 0)) || (rc = open_tasks(ctx))) return rc;
  ProfScope prof(ctx, K_EUL_FLUX);
  const int g = (P.nf + kBlock - 1) / kBlock;
  if (ctx->m.dim == 2)
    hipLaunchKernelGGL(k_eul_flux<2>, dim3(g), dim3(kBlock), 0, ctx->stream, ctx->m, ctx->eu, batch_args(ctx, -1), 0,
                       P.f0, P.nf);
  else
    hipLaunchKernelGGL(k_eul_flux<3>, dim3(g), dim3(kBlock), 0, ctx->stream, ctx->m, ctx->eu, batch_args(ctx, -1), 0,
                       P.f0, P.nf);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int fsale_eul_apply_update(fsale_ctx *ctx, fsale_eos eos, double factor) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  Problem &P = ctx->probs[0];
  if (!P.have_state) return FSALE_ERR_STATE;
  P.params.gamma = eos.gamma;
  P.params.cv = eos.cv;
  P.partials_valid = false;
  int rc;
  if ((rc = upload_params(ctx, 0))) return rc;
  if ((rc = sync_tiles(ctx)) || (rc = open_tasks(ctx))) return rc;
  const Batch b = batch_args(ctx, K_EUL_UPDATE);
  if (ctx->m.dim == 2)
    return launch_persistent(ctx, K_EUL_UPDATE, k_eul_update<2, 4>, ctx->m, ctx->eu, b, factor, 3);
  return launch_persistent(ctx, K_EUL_UPDATE, k_eul_update<3, 6>, ctx->m, ctx->eu, b, factor, 3);
}

int fsale_eul_advance(fsale_ctx *ctx, fsale_eos eos, double cfl, double final_time, int nsteps, int flux_kind,
                      int first_order, int stop_at_final, int use_graph) {
  if (!check_ctx(ctx) || nsteps < 0 || !valid_kind(flux_kind)) return FSALE_ERR_ARG;
  Problem &P = ctx->probs[0];
  if (!P.have_state) return FSALE_ERR_STATE;
  if (P.params.owned != 0) {
    set_message("a context with ghost cells steps through the task-level entry points and halo exchanges");
    return FSALE_ERR_STATE;
  }
  if (nsteps == 0) return FSALE_OK;
  int rc;
  if ((rc = read_scalars(ctx))) return rc;
  P.params.gamma = eos.gamma;
  P.params.cv = eos.cv;
  P.params.cfl = cfl;
  P.params.final_time = final_time;
  P.params.kind = flux_kind;
  P.params.first_order = first_order != 0;
  P.params.stop_at_final = stop_at_final != 0;
  P.params.max_steps = ctx->h_s[0].steps + nsteps;
  if ((rc = upload_params(ctx, 0))) return rc;
  return fsale_batch_step(ctx, nsteps, use_graph);
}

}  // extern "C"
