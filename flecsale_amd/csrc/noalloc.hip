// evaluate_time_step (apps/hydro/tasks.h:331-376) over caller-provided
// device arrays, for callers that cannot allocate device memory.
//
// fsale_eul_evaluate_time_step works on a context, whose partials buffer the
// library allocated at context creation.  This entry point takes the mesh
// and state as plain device pointers and keeps its per-wave partial maxima
// in `scratch`, a file-scope device array: it calls no hipMalloc, hipFree or
// allocating library routine.  Each device holds its own copy of the array
// until the process ends; its address is looked up once per device on the
// first call there (hipGetSymbolAddress), and that address, never the
// array's host-side name, is what the kernels receive.  Calls on one device
// share the array: calls that overlap on one device (different streams, or
// host threads) race on it and must be serialised by the caller.
//
// Arithmetic: per cell, over its faces in cell_faces order, dt^-1 =
// max((a + |v.n|) / (V / A)); the maximum over cells through per-wave
// partials (maximum is exact, so the order of the reduction does not change
// a bit); dt = min(CFL * (1 / max), t_final - t).  A non-positive maximum is
// the reference's "infinite delta t" (tasks.h:366-367): status
// FSALE_ERR_DT in *status_out, dt_out untouched.
#include <mutex>

#include "internal.h"
#include "kernels_common.h"

namespace fsale {
namespace {

// one partial per wave of 64 cells: up to 2^24 cells
constexpr int kScratchPartials = 1 << 18;
constexpr int64_t kNoAllocMaxCells = (int64_t)kScratchPartials * 64;
constexpr int kNaBlock = 256;
constexpr int kNaFinal = 1024;
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kScratchPartials];

std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

template <int N>
__global__ void __launch_bounds__(kNaBlock)
k_na_dt_partial(int nc, int fpc, const int32_t *__restrict__ cell_faces, const double *__restrict__ fnrm,
                const double *__restrict__ farea, const double *__restrict__ cvol,
                const double *__restrict__ vel, const double *__restrict__ snd, double *__restrict__ partials) {
  const int c = blockIdx.x * kNaBlock + threadIdx.x;
  double dt_inv = 0;
  if (c < nc) {
    double v[N];
    for (int d = 0; d < N; ++d) v[d] = vel[(size_t)c * N + d];
    const double a = snd[c], vol = cvol[c];
    for (int k = 0; k < fpc; ++k) {
      const int f = cell_faces[(size_t)c * fpc + k];
      double n[N];
      for (int d = 0; d < N; ++d) n[d] = fnrm[(size_t)f * N + d];
      const double delta_x = vol / farea[f];
      const double dti = (a + fabs(dotN<N>(v, n))) / delta_x;
      dt_inv = smax(dti, dt_inv);
    }
  }
  dt_inv = wave_max(dt_inv);
  if ((threadIdx.x & 63) == 0) partials[(size_t)blockIdx.x * (kNaBlock / 64) + (threadIdx.x >> 6)] = dt_inv;
}

__global__ void __launch_bounds__(kNaFinal)
k_na_dt_finalize(const double *__restrict__ partials, int n, double cfl, double final_time, double soln_time,
                 double *dt_out, int32_t *status_out) {
  __shared__ double red[kNaFinal / 64];
  double m = 0;
  for (int i = threadIdx.x; i < n; i += kNaFinal) m = smax(partials[i], m);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double mx = red[0];
  for (int w = 1; w < kNaFinal / 64; ++w) mx = smax(red[w], mx);
  if (mx <= 0) {
    *status_out = FSALE_ERR_DT;
    return;
  }
  double ts = 1 / mx;
  ts *= cfl;
  ts = smin(ts, final_time - soln_time);
  *dt_out = ts;
  *status_out = FSALE_OK;
}

// the device address of this device's scratch array (looked up once)
int scratch_for_current_device(double **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_message("fsale_eul_time_step_noalloc: no current HIP device");
    return FSALE_ERR_NO_DEVICE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_message(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return FSALE_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return FSALE_OK;
}

}  // namespace
}  // namespace fsale

using namespace fsale;

extern "C" int fsale_eul_time_step_noalloc(int32_t dim, int32_t num_cells, int32_t faces_per_cell,
                                           const int32_t *cell_faces, const double *face_normal,
                                           const double *face_area, const double *cell_volume,
                                           const double *velocity, const double *sound_speed, double cfl,
                                           double final_time, double soln_time, double *dt_out,
                                           int32_t *status_out, void *stream) {
  if ((dim != 2 && dim != 3) || num_cells <= 0 || faces_per_cell <= 0 || faces_per_cell > 8 || !cell_faces ||
      !face_normal || !face_area || !cell_volume || !velocity || !sound_speed || !dt_out || !status_out) {
    set_message("fsale_eul_time_step_noalloc: bad argument");
    return FSALE_ERR_ARG;
  }
  if ((int64_t)num_cells > kNoAllocMaxCells) {
    set_message("fsale_eul_time_step_noalloc: " + std::to_string(num_cells) + " cells exceed the scratch limit of " +
                std::to_string(kNoAllocMaxCells));
    return FSALE_ERR_ARG;
  }
  double *partials = nullptr;
  int rc = scratch_for_current_device(&partials);
  if (rc != FSALE_OK) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int blocks = (num_cells + kNaBlock - 1) / kNaBlock;
  const int nparts = blocks * (kNaBlock / 64);
  if (dim == 2)
    hipLaunchKernelGGL(k_na_dt_partial<2>, dim3(blocks), dim3(kNaBlock), 0, s, num_cells, faces_per_cell,
                       cell_faces, face_normal, face_area, cell_volume, velocity, sound_speed, partials);
  else
    hipLaunchKernelGGL(k_na_dt_partial<3>, dim3(blocks), dim3(kNaBlock), 0, s, num_cells, faces_per_cell,
                       cell_faces, face_normal, face_area, cell_volume, velocity, sound_speed, partials);
  hipLaunchKernelGGL(k_na_dt_finalize, dim3(1), dim3(kNaFinal), 0, s, partials, nparts, cfl, final_time,
                     soln_time, dt_out, status_out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_message(std::string("fsale_eul_time_step_noalloc launch: ") + hipGetErrorString(e));
    return FSALE_ERR_HIP;
  }
  return FSALE_OK;
}
