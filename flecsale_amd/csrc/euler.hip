// Eulerian MUSCL finite-volume step of apps/hydro on the device.
//
//   k_eul_dt      evaluate_time_step  apps/hydro/tasks.h:331-376
//   k_eul_grad    evaluate_gradients  apps/hydro/tasks.h:145-321
//   k_eul_flux    evaluate_fluxes     apps/hydro/tasks.h:385-536
//   k_eul_update  apply_update        apps/hydro/tasks.h:589-659
//
// One work-item per cell (dt, gradients, update) or per face (fluxes), SoA
// FP64 state, k-major fixed-degree connectivity so neighbouring lanes read
// neighbouring words.  The update gathers the face fluxes of its cell in the
// reference's mesh.faces(c) order (no atomics).  All four kernels are
// HBM-bound (no dense contraction): see the byte model in bench.py.
#include "internal.h"
#include "kernels_common.h"

namespace fsale {
namespace {

// ------------------------------------------------------------ state init
__global__ void __launch_bounds__(kBlock) k_eul_ic(int nc, EulerFields u, double g, double cv) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  // update_state_from_pressure euler_eqns.h:333-350
  const double ie = eos_e_dp(g, u.d[c], u.p[c]);
  u.e[c] = ie;
  u.T[c] = eos_T_de(cv, ie);
  u.a[c] = eos_a_de(g, ie);
}

// ----------------------------------------------------------- time step
// Block-level max of dt^-1, combined across blocks with an atomic max on the
// IEEE bit pattern (all values are >= 0); the last block to finish turns it
// into dt = min(CFL/dt_inv, t_final - t) and re-arms the scratch.
template <int N, int FPC>
__global__ void __launch_bounds__(kBlock) k_eul_dt(DevMesh m, EulerFields u, DevScalars *s, double cfl,
                                                   double final_time, int stop_at_final) {
  __shared__ double red[kBlock / 64];
  __shared__ int last;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double dt_inv = 0;
  if (c < m.nc) {
    double vel[N];
    for (int d = 0; d < N; ++d) vel[d] = u.v[d][c];
    const double a = u.a[c], vol = m.cvol[c];
#pragma unroll
    for (int k = 0; k < FPC; ++k) {
      const int f = m.cf[(size_t)k * m.nc + c] >> 1;
      double n[N];
      for (int d = 0; d < N; ++d) n[d] = m.fnrm[d][f];
      const double delta_x = vol / m.farea[f];
      const double vn = dotN<N>(vel, n);
      const double dti = (a + fabs(vn)) / delta_x;
      dt_inv = smax(dti, dt_inv);
    }
  }
  dt_inv = wave_max(dt_inv);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dt_inv;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = red[0];
    for (int w = 1; w < kBlock / 64; ++w) b = smax(red[w], b);
    atomic_max_nonneg(&s->red[0], b);
    __threadfence();
    const unsigned prev = atomicAdd(&s->done, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    const double mx = __longlong_as_double(
        (long long)atomicAdd(reinterpret_cast<unsigned long long *>(&s->red[0]), 0ull));
    s->red[0] = 0;
    s->done = 0;
    const double t = s->t;
    int active = (s->err_code == 0);
    if (stop_at_final && !(t < final_time)) active = 0;
    if (active) {
      if (mx <= 0) {
        atomicCAS(&s->err_code, 0, (int)FSALE_ERR_DT);
        active = 0;
      } else {
        double ts = 1 / mx;
        ts *= cfl;
        ts = smin(ts, final_time - t);
        s->dt = ts;
        s->steps = s->steps + 1;
      }
    }
    s->active = active;
  }
}

// ----------------------------------------------------------- gradients
template <int N, int VPC>
__global__ void __launch_bounds__(kBlock) k_eul_grad(DevMesh m, EulerFields u, const DevScalars *s,
                                                     int first_order) {
  constexpr int NV = N + 4;
  const int c0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (c0 >= m.nc || !s->active) return;
  const double tol = 1.e-9;
  double u0[NV], umin[NV], umax[NV], dudx[NV][N], dxdx[N][N], x0[N];
  for (int d = 0; d < N; ++d) x0[d] = m.ccen[d][c0];
  {
    int i = 0;
    u0[i++] = u.d[c0];
    for (int d = 0; d < N; ++d) u0[i++] = u.v[d][c0];
    u0[i++] = u.e[c0];
    u0[i++] = u.p[c0];
    u0[i++] = u.a[c0];
  }
  for (int j = 0; j < NV; ++j) {
    umin[j] = u0[j];
    umax[j] = u0[j];
    for (int d = 0; d < N; ++d) dudx[j][d] = 0;
  }
  for (int r = 0; r < N; ++r)
    for (int q = 0; q < N; ++q) dxdx[r][q] = 0;
  const int q0 = m.nbr_off[c0], q1 = m.nbr_off[c0 + 1];
  for (int q = q0; q < q1; ++q) {
    const int c1 = m.nbr[q];
    double u1[NV], dx[N];
    {
      int i = 0;
      u1[i++] = u.d[c1];
      for (int d = 0; d < N; ++d) u1[i++] = u.v[d][c1];
      u1[i++] = u.e[c1];
      u1[i++] = u.p[c1];
      u1[i++] = u.a[c1];
    }
    for (int j = 0; j < NV; ++j) {
      umin[j] = smin(u1[j], umin[j]);
      umax[j] = smax(u1[j], umax[j]);
    }
    for (int d = 0; d < N; ++d) dx[d] = m.ccen[d][c1] - x0[d];
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
  }
  double gu[NV][N];
  if constexpr (N == 2) {
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
  // limiter apps/hydro/tasks.h:267-299
  double phi[NV];
  for (int j = 0; j < NV; ++j) phi[j] = 1;
#pragma unroll
  for (int lv = 0; lv < VPC; ++lv) {
    const int vid = m.cv[(size_t)lv * m.nc + c0];
    double xv[N];
    for (int d = 0; d < N; ++d) xv[d] = m.x[d][vid];
    double u1[NV];
    for (int j = 0; j < NV; ++j) u1[j] = u0[j];
    for (int d = 0; d < N; ++d) {
      const double dx = xv[d] - x0[d];
      for (int j = 0; j < NV; ++j) u1[j] += dx * gu[j][d];
    }
    for (int j = 0; j < NV; ++j) {
      const double du = u1[j] - u0[j];
      if (du > tol)
        phi[j] = smin(phi[j], (umax[j] - u0[j]) / du);
      else if (du < -tol)
        phi[j] = smin(phi[j], (umin[j] - u0[j]) / du);
    }
  }
  for (int j = 0; j < NV; ++j) {
    phi[j] = smax(phi[j], 0);
    for (int d = 0; d < N; ++d) gu[j][d] = first_order ? 0.0 : gu[j][d] * phi[j];
  }
  for (int d = 0; d < N; ++d) {
    u.gd[d][c0] = gu[0][d];
    for (int k = 0; k < N; ++k) u.gv[k * N + d][c0] = gu[1 + k][d];
    u.ge[d][c0] = gu[1 + N][d];
    u.gp[d][c0] = gu[2 + N][d];
    u.ga[d][c0] = gu[3 + N][d];
  }
}

// -------------------------------------------------------------- fluxes
// face state of apps/hydro/tasks.h:418-432: slot 2 (pressure) gets dx.grad_e
// and slot 3 (energy) gets dx.grad_p, as the reference's tuple order does.
template <int N>
__device__ inline void reconstruct(const DevMesh &m, const EulerFields &u, int c, const double *fx,
                                   EState<N> &w) {
  double dx[N];
  for (int d = 0; d < N; ++d) dx[d] = fx[d] - m.ccen[d][c];
  w.d = u.d[c];
  for (int k = 0; k < N; ++k) w.v[k] = u.v[k][c];
  w.p = u.p[c];
  w.e = u.e[c];
  w.a = u.a[c];
  for (int i = 0; i < N; ++i) {
    w.d += dx[i] * u.gd[i][c];
    for (int k = 0; k < N; ++k) w.v[k] += dx[i] * u.gv[k * N + i][c];
    w.p += dx[i] * u.ge[i][c];
    w.e += dx[i] * u.gp[i][c];
    w.a += dx[i] * u.ga[i][c];
  }
}

template <int N>
__global__ void __launch_bounds__(kBlock) k_eul_flux(DevMesh m, EulerFields u, const DevScalars *s,
                                                     int kind) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nf || !s->active) return;
  double fx[N], n[N], out[N + 2];
  for (int d = 0; d < N; ++d) {
    fx[d] = m.fcen[d][f];
    n[d] = m.fnrm[d][f];
  }
  EState<N> wl, wr;
  reconstruct<N>(m, u, m.fc[f], fx, wl);
  const int c1 = m.fc[(size_t)m.nf + f];
  if (c1 >= 0) {
    reconstruct<N>(m, u, c1, fx, wr);
    face_flux<N>(kind & 0xf, wl, wr, n, out);
  } else if ((kind >> 4) == 1) {
    e_flux<N>(wl, n, out);   // transmissive boundary
  } else {
    wall_flux<N>(wl, n, out);
  }
  const double area = m.farea[f];
  for (int i = 0; i < N + 2; ++i) u.F[i][f] = out[i] * area;
}

// -------------------------------------------------------------- update
template <int N, int FPC>
__global__ void __launch_bounds__(kBlock) k_eul_update(DevMesh m, EulerFields u, DevScalars *s, double g,
                                                       double cv, double factor) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (!s->active) return;
  double delta_t = s->dt;
  delta_t *= factor;
  if (c == 0) s->t = s->t + delta_t;   // update_soln_time :554 (initial time 0)
  if (c >= m.nc) return;
  double du[N + 2];
  for (int i = 0; i < N + 2; ++i) du[i] = 0;
#pragma unroll
  for (int k = 0; k < FPC; ++k) {
    const int code = m.cf[(size_t)k * m.nc + c];
    const int f = code >> 1;
    if (code & 1) {
      for (int i = 0; i < N + 2; ++i) du[i] += u.F[i][f];
    } else {
      for (int i = 0; i < N + 2; ++i) du[i] -= u.F[i][f];
    }
  }
  const double sc = delta_t / m.cvol[c];
  for (int i = 0; i < N + 2; ++i) du[i] *= sc;
  // update_state_from_flux euler_eqns.h:383-410
  double vel[N];
  for (int d = 0; d < N; ++d) vel[d] = u.v[d][c];
  const double den0 = u.d[c];
  const double mass = den0 + du[0];
  const double et0 = u.e[c] + 0.5 * dotN<N>(vel, vel);
  const double ener = den0 * et0 + du[N + 1];
  const double inv_mass = 1 / mass;
  for (int d = 0; d < N; ++d) {
    const double mom = den0 * vel[d] + du[1 + d];
    vel[d] = mom * inv_mass;
  }
  const double ie = ener * inv_mass - 0.5 * dotN<N>(vel, vel);
  u.d[c] = mass;
  for (int d = 0; d < N; ++d) u.v[d][c] = vel[d];
  u.e[c] = ie;
  // update_state_from_energy euler_eqns.h:359-375
  u.p[c] = eos_p_de(g, mass, ie);
  u.T[c] = eos_T_de(cv, ie);
  u.a[c] = eos_a_de(g, ie);
  if (ie < 0 || mass < 0) {   // apps/hydro/tasks.h:654-655
    atomicCAS(&s->err_code, 0, (int)FSALE_ERR_NEG_STATE);
    atomicMin(reinterpret_cast<unsigned long long *>(&s->err_cell), (unsigned long long)c);
    s->active = 0;
  }
}

inline int grid(int n) { return (n + kBlock - 1) / kBlock; }

int launch_dt(fsale_ctx *ctx, double cfl, double final_time, int stop_at_final) {
  const DevMesh &m = ctx->m;
  ProfScope prof(ctx, K_EUL_DT);
  if (m.dim == 2)
    hipLaunchKernelGGL((k_eul_dt<2, 4>), dim3(grid(m.nc)), dim3(kBlock), 0, ctx->stream, m, ctx->eu, ctx->s,
                       cfl, final_time, stop_at_final);
  else
    hipLaunchKernelGGL((k_eul_dt<3, 6>), dim3(grid(m.nc)), dim3(kBlock), 0, ctx->stream, m, ctx->eu, ctx->s,
                       cfl, final_time, stop_at_final);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int launch_grad(fsale_ctx *ctx, int first_order) {
  const DevMesh &m = ctx->m;
  ProfScope prof(ctx, K_EUL_GRAD);
  if (m.dim == 2)
    hipLaunchKernelGGL((k_eul_grad<2, 4>), dim3(grid(m.nc)), dim3(kBlock), 0, ctx->stream, m, ctx->eu, ctx->s,
                       first_order);
  else
    hipLaunchKernelGGL((k_eul_grad<3, 8>), dim3(grid(m.nc)), dim3(kBlock), 0, ctx->stream, m, ctx->eu, ctx->s,
                       first_order);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int launch_flux(fsale_ctx *ctx, int kind) {
  const DevMesh &m = ctx->m;
  ProfScope prof(ctx, K_EUL_FLUX);
  if (m.dim == 2)
    hipLaunchKernelGGL((k_eul_flux<2>), dim3(grid(m.nf)), dim3(kBlock), 0, ctx->stream, m, ctx->eu, ctx->s, kind);
  else
    hipLaunchKernelGGL((k_eul_flux<3>), dim3(grid(m.nf)), dim3(kBlock), 0, ctx->stream, m, ctx->eu, ctx->s, kind);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int launch_update(fsale_ctx *ctx, fsale_eos eos, double factor) {
  const DevMesh &m = ctx->m;
  ProfScope prof(ctx, K_EUL_UPDATE);
  if (m.dim == 2)
    hipLaunchKernelGGL((k_eul_update<2, 4>), dim3(grid(m.nc)), dim3(kBlock), 0, ctx->stream, m, ctx->eu, ctx->s,
                       eos.gamma, eos.cv, factor);
  else
    hipLaunchKernelGGL((k_eul_update<3, 6>), dim3(grid(m.nc)), dim3(kBlock), 0, ctx->stream, m, ctx->eu, ctx->s,
                       eos.gamma, eos.cv, factor);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int launch_step(fsale_ctx *ctx, fsale_eos eos, double cfl, double final_time, int kind, int first_order,
                int stop_at_final) {
  int rc;
  if ((rc = launch_dt(ctx, cfl, final_time, stop_at_final))) return rc;
  for (int half = 0; half < 2; ++half) {
    if ((rc = launch_grad(ctx, first_order))) return rc;
    if ((rc = launch_flux(ctx, kind))) return rc;
    if ((rc = launch_update(ctx, eos, 0.5))) return rc;
  }
  return FSALE_OK;
}

bool check_ctx(fsale_ctx *ctx) {
  if (!ctx || ctx->solver != FSALE_SOLVER_EULER) {
    set_message("not an Eulerian context");
    return false;
  }
  return true;
}

}  // namespace

int euler_alloc(fsale_ctx *ctx) {
  const DevMesh &m = ctx->m;
  EulerFields &u = ctx->eu;
  const int N = m.dim;
  int rc;
  double **cell1[] = {&u.d, &u.e, &u.p, &u.T, &u.a};
  for (double **p : cell1)
    if ((rc = dev_alloc_n(ctx, p, m.nc))) return rc;
  for (int d = 0; d < N; ++d) {
    if ((rc = dev_alloc_n(ctx, &u.v[d], m.nc))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.gd[d], m.nc))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.ge[d], m.nc))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.gp[d], m.nc))) return rc;
    if ((rc = dev_alloc_n(ctx, &u.ga[d], m.nc))) return rc;
  }
  for (int d = 0; d < N * N; ++d)
    if ((rc = dev_alloc_n(ctx, &u.gv[d], m.nc))) return rc;
  for (int i = 0; i < N + 2; ++i)
    if ((rc = dev_alloc_n(ctx, &u.F[i], m.nf))) return rc;
  return FSALE_OK;
}

}  // namespace fsale

using namespace fsale;

extern "C" {

int fsale_eul_initial_conditions(fsale_ctx *ctx, fsale_eos eos, const double *d, const double *v,
                                 const double *p) {
  if (!check_ctx(ctx) || !d || !v || !p) return FSALE_ERR_ARG;
  if (!(eos.gamma > 1) || !(eos.cv > 0)) return FSALE_ERR_ARG;
  int rc;
  if ((rc = fsale_set_field(ctx, FSALE_F_DENSITY, d))) return rc;
  if ((rc = fsale_set_field(ctx, FSALE_F_VELOCITY, v))) return rc;
  if ((rc = fsale_set_field(ctx, FSALE_F_PRESSURE, p))) return rc;
  hipLaunchKernelGGL(k_eul_ic, dim3(grid(ctx->m.nc)), dim3(kBlock), 0, ctx->stream, ctx->m.nc, ctx->eu,
                     eos.gamma, eos.cv);
  FSALE_HIP(hipGetLastError());
  ctx->eu.gamma = eos.gamma;
  ctx->eu.cv = eos.cv;
  ctx->have_state = true;
  // fresh run: time 0, no error, active
  if ((rc = read_scalars(ctx))) return rc;
  DevScalars h = *ctx->h_s;
  h.t = 0;
  h.dt = 0;
  h.steps = 0;
  h.err_code = 0;
  h.err_cell = INT64_MAX;
  h.active = 1;
  h.red[0] = h.red[1] = 0;
  h.done = 0;
  *ctx->h_s = h;
  FSALE_HIP(hipMemcpyAsync(ctx->s, ctx->h_s, sizeof(DevScalars), hipMemcpyHostToDevice, ctx->stream));
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  return FSALE_OK;
}

int fsale_eul_evaluate_time_step(fsale_ctx *ctx, double cfl, double final_time, double *dt_out) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  if (!ctx->have_state) return FSALE_ERR_STATE;
  int rc = launch_dt(ctx, cfl, final_time, 0);
  if (rc || !dt_out) return rc;
  if ((rc = check_device_error(ctx))) return rc;
  *dt_out = ctx->h_s->dt;
  return FSALE_OK;
}

int fsale_eul_evaluate_gradients(fsale_ctx *ctx, int first_order) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  if (!ctx->have_state) return FSALE_ERR_STATE;
  return launch_grad(ctx, first_order);
}

bool valid_kind(int k) { return k >= 0 && (k & 0xf) <= 2 && (k >> 4) <= 1; }

int fsale_eul_evaluate_fluxes(fsale_ctx *ctx, int flux_kind) {
  if (!check_ctx(ctx) || !valid_kind(flux_kind)) return FSALE_ERR_ARG;
  if (!ctx->have_state) return FSALE_ERR_STATE;
  return launch_flux(ctx, flux_kind);
}

int fsale_eul_apply_update(fsale_ctx *ctx, fsale_eos eos, double factor) {
  if (!check_ctx(ctx)) return FSALE_ERR_ARG;
  if (!ctx->have_state) return FSALE_ERR_STATE;
  return launch_update(ctx, eos, factor);
}

int fsale_eul_advance(fsale_ctx *ctx, fsale_eos eos, double cfl, double final_time, int nsteps, int flux_kind,
                      int first_order, int stop_at_final, int use_graph) {
  if (!check_ctx(ctx) || nsteps < 0 || !valid_kind(flux_kind)) return FSALE_ERR_ARG;
  if (!ctx->have_state) return FSALE_ERR_STATE;
  int rc;
  if (!use_graph || nsteps < 2 || ctx->profiling) {
    for (int n = 0; n < nsteps; ++n)
      if ((rc = launch_step(ctx, eos, cfl, final_time, flux_kind, first_order, stop_at_final))) return rc;
    return FSALE_OK;
  }
  // capture a chunk of steps once, replay it; the key holds every baked-in argument
  const int chunk = nsteps < 10 ? nsteps : 10;
  std::vector<double> key = {0.0, eos.gamma, eos.cv, cfl, final_time, (double)flux_kind, (double)first_order,
                             (double)stop_at_final};
  if (!ctx->graph.exec || ctx->graph.key != key || ctx->graph.steps != chunk) {
    if (ctx->graph.exec) {
      (void)hipGraphExecDestroy(ctx->graph.exec);
      ctx->graph.exec = nullptr;
    }
    hipGraph_t graph;
    FSALE_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int n = 0; n < chunk; ++n) {
      rc = launch_step(ctx, eos, cfl, final_time, flux_kind, first_order, stop_at_final);
      if (rc) {
        (void)hipStreamEndCapture(ctx->stream, &graph);
        return rc;
      }
    }
    FSALE_HIP(hipStreamEndCapture(ctx->stream, &graph));
    FSALE_HIP(hipGraphInstantiate(&ctx->graph.exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    ctx->graph.key = key;
    ctx->graph.steps = chunk;
  }
  int n = 0;
  for (; n + chunk <= nsteps; n += chunk) FSALE_HIP(hipGraphLaunch(ctx->graph.exec, ctx->stream));
  for (; n < nsteps; ++n)
    if ((rc = launch_step(ctx, eos, cfl, final_time, flux_kind, first_order, stop_at_final))) return rc;
  return FSALE_OK;
}

}  // extern "C"
