// Mesh geometry on the device: cell centroid / volume / min_length and face
// centroid / unit normal / area, the quantities burton's update_geometry()
// provides to the hydro tasks (apps/hydro/tasks.h:32-36; c->centroid(),
// c->volume(), c->min_length(), f->centroid(), f->normal(), f->area()).
// The burton sources are not vendored; the 2D formulas below reproduce the
// reference goldens (shoelace centroid about the origin), the 3D hex formulas
// are build-defined (triangles / tetrahedra about vertex averages).
#include "internal.h"
#include "kernels_common.h"

namespace fsale {
namespace {

__global__ void __launch_bounds__(kBlock) k_cell_geom_2d(DevMesh m) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nc) return;
  const int n = m.vpc;
  double px[8], py[8];
  for (int i = 0; i < n; ++i) {
    const int v = m.cv[(size_t)i * m.nc + c];
    px[i] = m.x[0][v];
    py[i] = m.x[1][v];
  }
  double cx, cy, vol, ml;
  cell_geom_2d(n, px, py, cx, cy, vol, ml);
  m.ccen[0][c] = cx;
  m.ccen[1][c] = cy;
  m.cvol[c] = vol;
  m.cminlen[c] = ml;
}

__global__ void __launch_bounds__(kBlock) k_face_geom_2d(DevMesh m) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nf) return;
  const int a = m.fv[f], b = m.fv[(size_t)m.nf + f];
  double cx, cy, nx, ny, area;
  face_geom_2d(m.x[0][a], m.x[1][a], m.x[0][b], m.x[1][b], cx, cy, nx, ny, area);
  m.fcen[0][f] = cx;
  m.fcen[1][f] = cy;
  m.fnrm[0][f] = nx;
  m.fnrm[1][f] = ny;
  m.farea[f] = area;
}

__global__ void __launch_bounds__(kBlock) k_cell_geom_3d(DevMesh m) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nc) return;
  double p[8][3];
  for (int i = 0; i < 8; ++i) {
    const int v = m.cv[(size_t)i * m.nc + c];
    p[i][0] = m.x[0][v];
    p[i][1] = m.x[1][v];
    p[i][2] = m.x[2][v];
  }
  double cen[3], vol, ml;
  hex_geom(p, cen, vol, ml);
  for (int d = 0; d < 3; ++d) m.ccen[d][c] = cen[d];
  m.cvol[c] = vol;
  m.cminlen[c] = ml;
}

__global__ void __launch_bounds__(kBlock) k_face_geom_3d(DevMesh m) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nf) return;
  double q[4][3];
  for (int i = 0; i < 4; ++i) {
    const int v = m.fv[(size_t)i * m.nf + f];
    q[i][0] = m.x[0][v];
    q[i][1] = m.x[1][v];
    q[i][2] = m.x[2][v];
  }
  double cen[3], nrm[3], area;
  quad3_geom(q, cen, nrm, area);
  for (int d = 0; d < 3; ++d) {
    m.fcen[d][f] = cen[d];
    m.fnrm[d][f] = nrm[d];
  }
  m.farea[f] = area;
}

}  // namespace

int launch_geometry(fsale_ctx *ctx, bool cells, bool faces) {
  const DevMesh &m = ctx->m;
  if (cells) {
    const int g = (m.nc + kBlock - 1) / kBlock;
    if (m.dim == 2)
      hipLaunchKernelGGL(k_cell_geom_2d, dim3(g), dim3(kBlock), 0, ctx->stream, m);
    else
      hipLaunchKernelGGL(k_cell_geom_3d, dim3(g), dim3(kBlock), 0, ctx->stream, m);
    FSALE_HIP(hipGetLastError());
  }
  if (faces) {
    const int g = (m.nf + kBlock - 1) / kBlock;
    if (m.dim == 2)
      hipLaunchKernelGGL(k_face_geom_2d, dim3(g), dim3(kBlock), 0, ctx->stream, m);
    else
      hipLaunchKernelGGL(k_face_geom_3d, dim3(g), dim3(kBlock), 0, ctx->stream, m);
    FSALE_HIP(hipGetLastError());
  }
  return FSALE_OK;
}

}  // namespace fsale
