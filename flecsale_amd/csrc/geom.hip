// Mesh geometry on the device: cell centroid / volume / min_length and face
// centroid / unit normal / area, the quantities burton's update_geometry()
// provides to the hydro tasks (apps/hydro/tasks.h:32-36; c->centroid(),
// c->volume(), c->min_length(), f->centroid(), f->normal(), f->area()).
// The burton sources are not vendored; the 2D formulas below reproduce the
// reference goldens (shoelace centroid about the origin), the 3D hex formulas
// are build-defined (triangles / tetrahedra about vertex averages).
// Each launch covers one problem's cell range [c0, c0+n) or face range.
#include "internal.h"
#include "kernels_common.h"

namespace fsale {
namespace {

__global__ void __launch_bounds__(kBlock) k_cell_geom_2d(DevMesh m, int c0, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = c0 + i;
  const int nv = m.vpc;
  double px[8], py[8];
  for (int k = 0; k < nv; ++k) {
    const int v = m.cv[(size_t)k * m.cap_c + c];
    px[k] = m.x[0][v];
    py[k] = m.x[1][v];
  }
  double cx, cy, vol, ml;
  cell_geom_2d(nv, px, py, cx, cy, vol, ml);
  m.ccen[0][c] = cx;
  m.ccen[1][c] = cy;
  m.cvol[c] = vol;
  m.cminlen[c] = ml;
}

__global__ void __launch_bounds__(kBlock) k_face_geom_2d(DevMesh m, int f0, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int f = f0 + i;
  const int a = m.fv[f], b = m.fv[(size_t)m.cap_f + f];
  double cx, cy, nx, ny, area;
  face_geom_2d(m.x[0][a], m.x[1][a], m.x[0][b], m.x[1][b], cx, cy, nx, ny, area);
  m.fcen[0][f] = cx;
  m.fcen[1][f] = cy;
  m.fnrm[0][f] = nx;
  m.fnrm[1][f] = ny;
  m.farea[f] = area;
}

__global__ void __launch_bounds__(kBlock) k_cell_geom_3d(DevMesh m, int c0, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = c0 + i;
  double p[8][3];
  for (int k = 0; k < 8; ++k) {
    const int v = m.cv[(size_t)k * m.cap_c + c];
    p[k][0] = m.x[0][v];
    p[k][1] = m.x[1][v];
    p[k][2] = m.x[2][v];
  }
  double cen[3], vol, ml;
  hex_geom(p, cen, vol, ml);
  for (int d = 0; d < 3; ++d) m.ccen[d][c] = cen[d];
  m.cvol[c] = vol;
  m.cminlen[c] = ml;
}

__global__ void __launch_bounds__(kBlock) k_face_geom_3d(DevMesh m, int f0, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int f = f0 + i;
  double q[4][3];
  for (int k = 0; k < 4; ++k) {
    const int v = m.fv[(size_t)k * m.cap_f + f];
    q[k][0] = m.x[0][v];
    q[k][1] = m.x[1][v];
    q[k][2] = m.x[2][v];
  }
  double cen[3], nrm[3], area;
  quad3_geom(q, cen, nrm, area);
  for (int d = 0; d < 3; ++d) {
    m.fcen[d][f] = cen[d];
    m.fnrm[d][f] = nrm[d];
  }
  m.farea[f] = area;
}

// copies of the face geometry into each cell's local-face slots
__global__ void __launch_bounds__(kBlock) k_cellface_geom_3d(DevMesh m, int c0, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = c0 + i;
  for (int k = 0; k < m.fpc; ++k) {
    const int f = m.cf[(size_t)k * m.cap_c + c] >> 1;
    const size_t o = (size_t)k * m.cap_c + c;
    m.kfp[0][o] = make_double2(m.fcen[0][f], m.fcen[1][f]);
    m.kfp[1][o] = make_double2(m.fcen[2][f], m.fnrm[0][f]);
    m.kfp[2][o] = make_double2(m.fnrm[1][f], m.fnrm[2][f]);
    m.kfa[o] = m.farea[f];
  }
}

__global__ void __launch_bounds__(kBlock) k_vertex_pairs_2d(DevMesh m, int v0, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = v0 + i;
  m.xy[v] = make_double2(m.x[0][v], m.x[1][v]);
}

}  // namespace

int launch_geometry(fsale_ctx *ctx, int prob, bool cells, bool faces, hipStream_t stream) {
  const DevMesh &m = ctx->m;
  const Problem &P = ctx->probs[prob];
  if (cells && P.nc > 0) {
    const int g = (P.nc + kBlock - 1) / kBlock;
    if (m.dim == 2)
      hipLaunchKernelGGL(k_cell_geom_2d, dim3(g), dim3(kBlock), 0, stream, m, P.c0, P.nc);
    else
      hipLaunchKernelGGL(k_cell_geom_3d, dim3(g), dim3(kBlock), 0, stream, m, P.c0, P.nc);
    FSALE_HIP(hipGetLastError());
    if (m.xy && P.nv > 0) {
      hipLaunchKernelGGL(k_vertex_pairs_2d, dim3((P.nv + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, m, P.v0,
                         P.nv);
      FSALE_HIP(hipGetLastError());
    }
  }
  if (faces && P.nf > 0) {
    const int g = (P.nf + kBlock - 1) / kBlock;
    if (m.dim == 2)
      hipLaunchKernelGGL(k_face_geom_2d, dim3(g), dim3(kBlock), 0, stream, m, P.f0, P.nf);
    else
      hipLaunchKernelGGL(k_face_geom_3d, dim3(g), dim3(kBlock), 0, stream, m, P.f0, P.nf);
    FSALE_HIP(hipGetLastError());
    if (m.kfa && P.nc > 0) {
      hipLaunchKernelGGL(k_cellface_geom_3d, dim3((P.nc + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, m, P.c0,
                         P.nc);
      FSALE_HIP(hipGetLastError());
    }
  }
  return FSALE_OK;
}

}  // namespace fsale
