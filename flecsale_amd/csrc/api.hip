// Context life cycle, field I/O and error reporting of the C ABI
// (include/flecsale_amd.h).  The context replaces the FleCSI runtime's field
// storage and client handles (apps/hydro/driver.h:35-158,215-232;
// apps/maire_hydro/driver.h:36-166,237-258): all state lives in device
// memory as SoA FP64 arrays.
#include <cstring>
#include <mutex>

#include "internal.h"
#include "kernels_common.h"

namespace fsale {

static thread_local std::string g_message;
void set_message(const std::string &msg) { g_message = msg; }

int dev_alloc(fsale_ctx *ctx, void **p, size_t bytes) {
  *p = nullptr;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    set_message(std::string("hipMalloc: ") + hipGetErrorString(e));
    *p = nullptr;
    return FSALE_ERR_OOM;
  }
  ctx->allocations.push_back(*p);
  return FSALE_OK;
}

ProfScope::ProfScope(fsale_ctx *c, int kid) : ctx(c), id(kid) {
  if (!ctx->profiling) return;
  hipEvent_t start;
  if (hipEventCreate(&start) != hipSuccess) return;
  if (hipEventCreate(&stop) != hipSuccess) {
    (void)hipEventDestroy(start);
    stop = nullptr;
    return;
  }
  (void)hipEventRecord(start, ctx->stream);
  ctx->prof[id].start.push_back(start);
}

ProfScope::~ProfScope() {
  if (!ctx->profiling || !stop) return;
  (void)hipEventRecord(stop, ctx->stream);
  ctx->prof[id].stop.push_back(stop);
}

int read_scalars(fsale_ctx *ctx) {
  FSALE_HIP(hipMemcpyAsync(ctx->h_s, ctx->s, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  return FSALE_OK;
}

int check_device_error(fsale_ctx *ctx) {
  int rc = read_scalars(ctx);
  if (rc) return rc;
  if (ctx->h_s->err_code) {
    if (ctx->h_s->err_code == FSALE_ERR_NEG_STATE)
      set_message("Negative density or internal energy encountered! (cell " +
                  std::to_string(ctx->h_s->err_cell) + ")");
    else if (ctx->h_s->err_code == FSALE_ERR_DT)
      set_message("infinite delta t");
    return ctx->h_s->err_code;
  }
  return FSALE_OK;
}

namespace {

template <class T>
int upload_kmajor(fsale_ctx *ctx, T **dst, const std::vector<T> &src, int n, int k) {
  // src is entity-major [n][k]; device layout is k-major [k][n]
  std::vector<T> tmp((size_t)n * k);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < k; ++j) tmp[(size_t)j * n + i] = src[(size_t)i * k + j];
  int rc = dev_alloc_n(ctx, dst, tmp.size());
  if (rc) return rc;
  FSALE_HIP(hipMemcpy(*dst, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
  return FSALE_OK;
}

template <class T>
int upload(fsale_ctx *ctx, T **dst, const std::vector<T> &src) {
  int rc = dev_alloc_n(ctx, dst, src.size() + 1);
  if (rc) return rc;
  if (!src.empty()) FSALE_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return FSALE_OK;
}

int create_impl(const fsale_mesh_desc *mesh, int solver, fsale_ctx *ctx) {
  std::string err;
  int rc = build_topology(*mesh, ctx->topo, err);
  if (rc) {
    set_message(err);
    return rc;
  }
  const HostTopology &t = ctx->topo;
  if (solver == FSALE_SOLVER_MAIRE && (t.dim != 2 || t.vpc != 4)) {
    set_message("the Lagrangian solver supports 2D quad meshes");
    return FSALE_ERR_MESH;
  }
  for (int v = 0; v < t.nv; ++v)
    if (t.vc_off[v + 1] - t.vc_off[v] > kMaxCornersPerVertex) {
      set_message("more than 8 cells around a vertex");
      return FSALE_ERR_MESH;
    }
  DevMesh &m = ctx->m;
  m.dim = t.dim;
  m.nc = t.nc;
  m.nv = t.nv;
  m.nf = t.nf;
  m.vpc = t.vpc;
  m.fpc = t.fpc;
  m.vpf = t.vpf;
  m.nnbr = (int)t.nbr.size();
  for (int d = 0; d < t.dim; ++d) {
    std::vector<double> xd(t.nv);
    for (int v = 0; v < t.nv; ++v) xd[v] = mesh->coords[(size_t)v * t.dim + d];
    if ((rc = upload(ctx, &m.x[d], xd))) return rc;
  }
  if ((rc = upload_kmajor(ctx, &m.cv, t.cv, t.nc, t.vpc))) return rc;
  {
    std::vector<int32_t> cf(t.cf.size());
    for (int c = 0; c < t.nc; ++c)
      for (int k = 0; k < t.fpc; ++k) {
        const int f = t.cf[(size_t)c * t.fpc + k];
        cf[(size_t)c * t.fpc + k] = 2 * f + (t.fc[2 * (size_t)f] == c ? 0 : 1);
      }
    if ((rc = upload_kmajor(ctx, &m.cf, cf, t.nc, t.fpc))) return rc;
  }
  if ((rc = upload_kmajor(ctx, &m.fc, t.fc, t.nf, 2))) return rc;
  if ((rc = upload_kmajor(ctx, &m.fv, t.fv, t.nf, t.vpf))) return rc;
  if ((rc = upload(ctx, &m.nbr_off, t.nbr_off))) return rc;
  if ((rc = upload(ctx, &m.nbr, t.nbr))) return rc;
  if ((rc = upload(ctx, &m.vc_off, t.vc_off))) return rc;
  if ((rc = upload(ctx, &m.vc, t.vc))) return rc;
  for (int d = 0; d < t.dim; ++d) {
    if ((rc = dev_alloc_n(ctx, &m.ccen[d], t.nc))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.fcen[d], t.nf))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.fnrm[d], t.nf))) return rc;
  }
  if ((rc = dev_alloc_n(ctx, &m.cvol, t.nc))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.cminlen, t.nc))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.farea, t.nf))) return rc;
  if ((rc = dev_alloc_n(ctx, &ctx->s, 1))) return rc;
  FSALE_HIP(hipHostMalloc((void **)&ctx->h_s, sizeof(DevScalars), hipHostMallocDefault));
  std::memset(ctx->h_s, 0, sizeof(DevScalars));
  FSALE_HIP(hipMemcpy(ctx->s, ctx->h_s, sizeof(DevScalars), hipMemcpyHostToDevice));
  if ((rc = launch_geometry(ctx, true, true))) return rc;

  if (solver == FSALE_SOLVER_MAIRE) {
    // boundary tags from the initial face normals: tag d when |n_d| is 1
    // within 1e-14 (apps/maire_hydro/2d/inputs_regression.cc:134-162)
    std::vector<double> fn[2];
    for (int d = 0; d < 2; ++d) {
      fn[d].resize(t.nf);
      FSALE_HIP(hipMemcpyAsync(fn[d].data(), m.fnrm[d], t.nf * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
    }
    FSALE_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<uint8_t> ftag(t.nf, 0), vbnd(t.nv, 0);
    for (int f = 0; f < t.nf; ++f) {
      if (t.fc[2 * (size_t)f + 1] >= 0) continue;
      uint8_t tag = 0x4;  // bit 2: boundary face
      for (int d = 0; d < 2; ++d)
        if (std::fabs(std::fabs(fn[d][f]) - 1.0) <= 1e-14) tag |= (uint8_t)(1u << d);
      ftag[f] = tag;
      for (int i = 0; i < t.vpf; ++i) vbnd[t.fv[(size_t)f * t.vpf + i]] = 1;
    }
    // corner (c,k): low nibble = face of edge (k-1 -> k), high nibble = edge (k -> k+1)
    std::vector<uint8_t> cw((size_t)t.nc * t.vpc);
    for (int c = 0; c < t.nc; ++c)
      for (int k = 0; k < t.vpc; ++k) {
        const int km = (k + t.vpc - 1) % t.vpc;
        const int fm = t.cf[(size_t)c * t.fpc + km], fp = t.cf[(size_t)c * t.fpc + k];
        cw[(size_t)c * t.vpc + k] = (uint8_t)(ftag[fm] | (ftag[fp] << 4));
      }
    if ((rc = upload_kmajor(ctx, &m.cwtag, cw, t.nc, t.vpc))) return rc;
    if ((rc = upload(ctx, &m.vbnd, vbnd))) return rc;
    if ((rc = maire_alloc(ctx))) return rc;
  } else {
    if ((rc = euler_alloc(ctx))) return rc;
  }
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  return FSALE_OK;
}

// host <-> device field conversions
struct FieldView {
  double *const *ptrs = nullptr;   // component arrays on the device
  int ncomp = 0;                   // interleaved components on the host
  int n = 0;                       // entities
  bool corner = false;             // host corner id c*4+k, device [k][nc]
  bool writable = true;
};

int field_view(fsale_ctx *ctx, int field, FieldView &fv, double *tmp[9]) {
  DevMesh &m = ctx->m;
  const int N = m.dim;
  auto one = [&](double *p, int n) {
    tmp[0] = p;
    fv.ptrs = tmp;
    fv.ncomp = 1;
    fv.n = n;
  };
  auto vec = [&](double *const *p, int ncomp, int n) {
    for (int i = 0; i < ncomp; ++i) tmp[i] = p[i];
    fv.ptrs = tmp;
    fv.ncomp = ncomp;
    fv.n = n;
  };
  if (ctx->solver == FSALE_SOLVER_EULER) {
    EulerFields &u = ctx->eu;
    switch (field) {
      case FSALE_F_DENSITY: one(u.d, m.nc); return FSALE_OK;
      case FSALE_F_VELOCITY: vec(u.v, N, m.nc); return FSALE_OK;
      case FSALE_F_INTERNAL_ENERGY: one(u.e, m.nc); return FSALE_OK;
      case FSALE_F_PRESSURE: one(u.p, m.nc); return FSALE_OK;
      case FSALE_F_TEMPERATURE: one(u.T, m.nc); return FSALE_OK;
      case FSALE_F_SOUND_SPEED: one(u.a, m.nc); return FSALE_OK;
      case FSALE_F_GRAD_DENSITY: vec(u.gd, N, m.nc); return FSALE_OK;
      case FSALE_F_GRAD_VELOCITY: vec(u.gv, N * N, m.nc); return FSALE_OK;
      case FSALE_F_GRAD_ENERGY: vec(u.ge, N, m.nc); return FSALE_OK;
      case FSALE_F_GRAD_PRESSURE: vec(u.gp, N, m.nc); return FSALE_OK;
      case FSALE_F_GRAD_SOUND_SPEED: vec(u.ga, N, m.nc); return FSALE_OK;
      case FSALE_F_FLUX: vec(u.F, N + 2, m.nf); return FSALE_OK;
      default: break;
    }
  } else {
    MaireFields &u = ctx->ma;
    switch (field) {
      case FSALE_F_DENSITY: one(u.d, m.nc); return FSALE_OK;
      case FSALE_F_VELOCITY: vec(u.v, 2, m.nc); return FSALE_OK;
      case FSALE_F_INTERNAL_ENERGY: one(u.e, m.nc); return FSALE_OK;
      case FSALE_F_PRESSURE: one(u.p, m.nc); return FSALE_OK;
      case FSALE_F_TEMPERATURE: one(u.T, m.nc); return FSALE_OK;
      case FSALE_F_SOUND_SPEED: one(u.a, m.nc); return FSALE_OK;
      case FSALE_F_MASS: one(u.M, m.nc); return FSALE_OK;
      case FSALE_F_VOLUME: one(u.V, m.nc); return FSALE_OK;
      case FSALE_F_GAMMA: one(u.g, m.nc); return FSALE_OK;
      case FSALE_F_NODE_VELOCITY: vec(u.un, 2, m.nv); return FSALE_OK;
      case FSALE_F_CORNER_NORMAL: vec(u.npc, 2, m.nc * m.vpc); fv.corner = true; return FSALE_OK;
      case FSALE_F_CORNER_FORCE: vec(u.Fpc, 2, m.nc * m.vpc); fv.corner = true; return FSALE_OK;
      case FSALE_F_RESIDUAL: vec(u.dudt, 4, m.nc); return FSALE_OK;
      default: break;
    }
  }
  switch (field) {
    case FSALE_F_CELL_CENTROID: vec(m.ccen, N, m.nc); fv.writable = false; return FSALE_OK;
    case FSALE_F_CELL_VOLUME: one(m.cvol, m.nc); fv.writable = false; return FSALE_OK;
    case FSALE_F_CELL_MIN_LENGTH: one(m.cminlen, m.nc); fv.writable = false; return FSALE_OK;
    case FSALE_F_FACE_CENTROID: vec(m.fcen, N, m.nf); fv.writable = false; return FSALE_OK;
    case FSALE_F_FACE_NORMAL: vec(m.fnrm, N, m.nf); fv.writable = false; return FSALE_OK;
    case FSALE_F_FACE_AREA: one(m.farea, m.nf); fv.writable = false; return FSALE_OK;
    case FSALE_F_COORDS: vec(m.x, N, m.nv); return FSALE_OK;
    default: break;
  }
  set_message("field not available for this solver");
  return FSALE_ERR_ARG;
}

// corner reorder: host index c*vpc+k  <->  device index k*nc+c
inline size_t corner_dev_index(const DevMesh &m, size_t host_i) {
  const size_t c = host_i / m.vpc, k = host_i % m.vpc;
  return k * m.nc + c;
}

}  // namespace
}  // namespace fsale

using namespace fsale;

extern "C" {

const char *fsale_status_string(int status) {
  switch (status) {
    case FSALE_OK: return "ok";
    case FSALE_ERR_ARG: return "bad argument";
    case FSALE_ERR_MESH: return "unsupported or invalid mesh";
    case FSALE_ERR_HIP: return "HIP runtime error";
    case FSALE_ERR_OOM: return "out of device memory";
    case FSALE_ERR_NEG_STATE: return "negative density or internal energy";
    case FSALE_ERR_DT: return "infinite delta t";
    case FSALE_ERR_STATE: return "call out of order";
    case FSALE_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

const char *fsale_last_message(void) { return g_message.c_str(); }

int fsale_ctx_create(const fsale_mesh_desc *mesh, int solver, fsale_ctx **out) {
  if (!mesh || !out || (solver != FSALE_SOLVER_EULER && solver != FSALE_SOLVER_MAIRE)) return FSALE_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_message("no HIP device visible");
    return FSALE_ERR_NO_DEVICE;
  }
  fsale_ctx *ctx = new fsale_ctx();
  ctx->solver = solver;
  (void)hipGetDevice(&ctx->device);
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    set_message("hipStreamCreate failed");
    return FSALE_ERR_HIP;
  }
  int rc = create_impl(mesh, solver, ctx);
  if (rc) {
    fsale_ctx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return FSALE_OK;
}

int fsale_ctx_destroy(fsale_ctx *ctx) {
  if (!ctx) return FSALE_ERR_ARG;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &p : ctx->prof) {
    for (hipEvent_t e : p.start) (void)hipEventDestroy(e);
    for (hipEvent_t e : p.stop) (void)hipEventDestroy(e);
  }
  if (ctx->graph.exec) (void)hipGraphExecDestroy(ctx->graph.exec);
  for (void *p : ctx->allocations) (void)hipFree(p);
  if (ctx->h_s) (void)hipHostFree(ctx->h_s);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FSALE_OK;
}

int fsale_sync(fsale_ctx *ctx) {
  if (!ctx) return FSALE_ERR_ARG;
  return check_device_error(ctx);
}

int fsale_mesh_counts(const fsale_ctx *ctx, int32_t *counts) {
  if (!ctx || !counts) return FSALE_ERR_ARG;
  counts[0] = ctx->m.nc;
  counts[1] = ctx->m.nv;
  counts[2] = ctx->m.nf;
  counts[3] = ctx->m.nnbr;
  counts[4] = ctx->m.dim;
  return FSALE_OK;
}

int fsale_get_topology(fsale_ctx *ctx, int32_t *face_cells, int32_t *cell_faces, int32_t *nbr_off,
                       int32_t *nbr) {
  if (!ctx) return FSALE_ERR_ARG;
  const HostTopology &t = ctx->topo;
  if (face_cells) std::memcpy(face_cells, t.fc.data(), t.fc.size() * sizeof(int32_t));
  if (cell_faces) std::memcpy(cell_faces, t.cf.data(), t.cf.size() * sizeof(int32_t));
  if (nbr_off) std::memcpy(nbr_off, t.nbr_off.data(), t.nbr_off.size() * sizeof(int32_t));
  if (nbr) std::memcpy(nbr, t.nbr.data(), t.nbr.size() * sizeof(int32_t));
  return FSALE_OK;
}

int fsale_mesh_connectivity(const fsale_mesh_desc *mesh, int32_t *counts, int32_t *face_cells,
                            int32_t *cell_faces, int32_t *nbr_off, int32_t *nbr) {
  if (!mesh || !counts) return FSALE_ERR_ARG;
  HostTopology t;
  std::string err;
  int rc = build_topology(*mesh, t, err);
  if (rc) {
    set_message(err);
    return rc;
  }
  counts[0] = t.nc;
  counts[1] = t.nv;
  counts[2] = t.nf;
  counts[3] = (int32_t)t.nbr.size();
  counts[4] = t.dim;
  if (face_cells) std::memcpy(face_cells, t.fc.data(), t.fc.size() * sizeof(int32_t));
  if (cell_faces) std::memcpy(cell_faces, t.cf.data(), t.cf.size() * sizeof(int32_t));
  if (nbr_off) std::memcpy(nbr_off, t.nbr_off.data(), t.nbr_off.size() * sizeof(int32_t));
  if (nbr) std::memcpy(nbr, t.nbr.data(), t.nbr.size() * sizeof(int32_t));
  return FSALE_OK;
}

int fsale_get_field(fsale_ctx *ctx, int field, double *host_out) {
  if (!ctx || !host_out) return FSALE_ERR_ARG;
  FieldView fv;
  double *tmp[9];
  int rc = field_view(ctx, field, fv, tmp);
  if (rc) return rc;
  std::vector<double> buf((size_t)fv.n);
  for (int k = 0; k < fv.ncomp; ++k) {
    FSALE_HIP(hipMemcpyAsync(buf.data(), fv.ptrs[k], buf.size() * sizeof(double), hipMemcpyDeviceToHost,
                             ctx->stream));
    FSALE_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < buf.size(); ++i) {
      const size_t src = fv.corner ? corner_dev_index(ctx->m, i) : i;
      host_out[i * fv.ncomp + k] = buf[src];
    }
  }
  return check_device_error(ctx);
}

int fsale_set_field(fsale_ctx *ctx, int field, const double *host_in) {
  if (!ctx || !host_in) return FSALE_ERR_ARG;
  FieldView fv;
  double *tmp[9];
  int rc = field_view(ctx, field, fv, tmp);
  if (rc) return rc;
  if (!fv.writable) {
    set_message("geometry fields are derived from the coordinates");
    return FSALE_ERR_ARG;
  }
  std::vector<double> buf((size_t)fv.n);
  for (int k = 0; k < fv.ncomp; ++k) {
    for (size_t i = 0; i < buf.size(); ++i) {
      const size_t dst = fv.corner ? corner_dev_index(ctx->m, i) : i;
      buf[dst] = host_in[i * fv.ncomp + k];
    }
    FSALE_HIP(hipMemcpyAsync(fv.ptrs[k], buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
    FSALE_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (ctx->solver == FSALE_SOLVER_EULER) ctx->eu.partials_valid = false;
  if (field == FSALE_F_COORDS) return launch_geometry(ctx, true, true);
  return FSALE_OK;
}

int fsale_get_time(fsale_ctx *ctx, double *soln_time, double *time_step, int64_t *steps) {
  if (!ctx) return FSALE_ERR_ARG;
  int rc = read_scalars(ctx);
  if (rc) return rc;
  if (soln_time) *soln_time = ctx->h_s->t;
  if (time_step) *time_step = ctx->h_s->dt;
  if (steps) *steps = ctx->h_s->steps;
  return FSALE_OK;
}

int fsale_set_time(fsale_ctx *ctx, double soln_time, double time_step) {
  if (!ctx) return FSALE_ERR_ARG;
  int rc = read_scalars(ctx);
  if (rc) return rc;
  ctx->h_s->t = soln_time;
  ctx->h_s->dt = time_step;
  FSALE_HIP(hipMemcpyAsync(ctx->s, ctx->h_s, sizeof(DevScalars), hipMemcpyHostToDevice, ctx->stream));
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  return FSALE_OK;
}

int fsale_set_profiling(fsale_ctx *ctx, int enable) {
  if (!ctx) return FSALE_ERR_ARG;
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  ctx->profiling = enable != 0;
  return FSALE_OK;
}

int fsale_get_profile(fsale_ctx *ctx, int kernel, double *total_ms, int64_t *launches) {
  if (!ctx || kernel < 0 || kernel >= K_COUNT) return FSALE_ERR_ARG;
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  ProfSlot &p = ctx->prof[kernel];
  const size_t n = p.stop.size();
  for (size_t i = 0; i < n; ++i) {
    float ms = 0;
    FSALE_HIP(hipEventElapsedTime(&ms, p.start[i], p.stop[i]));
    p.total_ms += ms;
    p.launches += 1;
  }
  for (hipEvent_t e : p.start) (void)hipEventDestroy(e);
  for (hipEvent_t e : p.stop) (void)hipEventDestroy(e);
  p.start.clear();
  p.stop.clear();
  if (total_ms) *total_ms = p.total_ms;
  if (launches) *launches = p.launches;
  return FSALE_OK;
}

int fsale_reset_profile(fsale_ctx *ctx) {
  if (!ctx) return FSALE_ERR_ARG;
  for (int k = 0; k < K_COUNT; ++k) {
    double t;
    int64_t n;
    int rc = fsale_get_profile(ctx, k, &t, &n);
    if (rc) return rc;
    ctx->prof[k].total_ms = 0;
    ctx->prof[k].launches = 0;
  }
  return FSALE_OK;
}

int fsale_error_cell(fsale_ctx *ctx, int64_t *cell) {
  if (!ctx || !cell) return FSALE_ERR_ARG;
  int rc = read_scalars(ctx);
  if (rc) return rc;
  *cell = ctx->h_s->err_code ? ctx->h_s->err_cell : -1;
  return FSALE_OK;
}

}  // extern "C"
