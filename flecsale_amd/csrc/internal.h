// Internal declarations shared by the host and device translation units of
// libflecsale_amd.  Not part of the public ABI (include/flecsale_amd.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/flecsale_amd.h"

namespace fsale {

constexpr int kBlock = 256;     // threads per block for the entity kernels
constexpr int kMaxCornersPerVertex = 8;

void set_message(const std::string &msg);

#define FSALE_HIP(call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::fsale::set_message(std::string(#call) + ": " + hipGetErrorString(e_)); \
      return (e_ == hipErrorOutOfMemory) ? FSALE_ERR_OOM : FSALE_ERR_HIP;      \
    }                                                                          \
  } while (0)

// Host-side connectivity, built once per mesh (mesh_host.cpp).  Conventions
// (the reference's burton mesh is not vendored; these reproduce its goldens):
//  - faces numbered in order of first visit, cell by cell in local face order
//    (2D: edge k = (v_k, v_{k+1}); 3D: Exodus hex sides), fc[0] = first
//    (lowest-id) cell, face vertices oriented out of fc[0];
//  - cell-vertex neighbours: ascending cell id, self excluded;
//  - vertex -> corners in ascending cell id.
struct HostTopology {
  int dim = 0, nc = 0, nv = 0, vpc = 0, fpc = 0, vpf = 0, nf = 0;
  std::vector<int32_t> cv;        // [nc*vpc] cell-major (as given)
  std::vector<int32_t> cf;        // [nc*fpc] cell-major face ids
  std::vector<int32_t> fc;        // [nf*2]   face cells (-1 boundary)
  std::vector<int32_t> fv;        // [nf*vpf] face vertices
  std::vector<int32_t> nbr_off;   // [nc+1]
  std::vector<int32_t> nbr;
  std::vector<int32_t> vc_off;    // [nv+1]
  std::vector<int32_t> vc;        // corner ids c*vpc + k, ascending
};

int build_topology(const fsale_mesh_desc &mesh, HostTopology &t, std::string &err);

// Device-side mesh: SoA geometry, k-major fixed-degree connectivity.
struct DevMesh {
  int dim = 0, nc = 0, nv = 0, nf = 0, vpc = 0, fpc = 0, vpf = 0, nnbr = 0;
  double *x[3] = {nullptr, nullptr, nullptr};      // vertex coords [nv]
  int32_t *cv = nullptr;     // [vpc][nc]
  int32_t *cf = nullptr;     // [fpc][nc]: face*2 + (cell is fc[1] of the face)
  int32_t *fc = nullptr;     // [2][nf]
  int32_t *fv = nullptr;     // [vpf][nf]
  int32_t *nbr_off = nullptr, *nbr = nullptr;
  int32_t *vc_off = nullptr, *vc = nullptr;        // vertex -> corner (c*vpc+k)
  uint8_t *cwtag = nullptr;  // [vpc][nc] Maire wedge boundary tags
  uint8_t *vbnd = nullptr;   // [nv] vertex on a boundary face
  double *ccen[3] = {nullptr, nullptr, nullptr};
  double *cvol = nullptr, *cminlen = nullptr;
  double *fcen[3] = {nullptr, nullptr, nullptr};
  double *fnrm[3] = {nullptr, nullptr, nullptr};
  double *farea = nullptr;
};

// Scalars that stay on the device between tasks (dt future, solution time,
// error word, reduction scratch).
struct DevScalars {
  double dt;            // current time step (the reference's dt future)
  double t;             // solution time
  double red[2];        // max reductions as uint64 bit patterns (>= 0 values)
  unsigned int done;    // finished-block counter of the reduction kernels
  int active;           // step is live (t < t_final, no error)
  int err_code;         // first fsale_status raised on the device
  int pad;
  long long err_cell;   // lowest failing cell
  long long steps;      // completed steps
};

struct EulerFields {
  double *d = nullptr, *v[3] = {nullptr, nullptr, nullptr}, *e = nullptr, *p = nullptr,
         *T = nullptr, *a = nullptr;
  // second state buffer: the fused flux+update kernel reads one and writes
  // the other (its neighbours must see the old state); a step flips twice
  double *alt_d = nullptr, *alt_v[3] = {nullptr, nullptr, nullptr}, *alt_e = nullptr, *alt_p = nullptr,
         *alt_T = nullptr, *alt_a = nullptr;
  double *gd[3] = {}, *gv[9] = {}, *ge[3] = {}, *gp[3] = {}, *ga[3] = {};
  double *F[5] = {};
  double *partials = nullptr;    // per-block max of dt^-1 [grid8(nc)]
  bool partials_valid = false;   // partials describe the current state
  double gamma = 0, cv = 0;
};

struct MaireFields {
  double *V = nullptr, *M = nullptr, *v[2] = {}, *p = nullptr, *d = nullptr, *e = nullptr,
         *T = nullptr, *a = nullptr, *g = nullptr;
  double *v0[2] = {}, *e0 = nullptr, *x0[2] = {};
  double *un[2] = {};
  double *npc[2] = {}, *Fpc[2] = {};   // corner arrays [vpc][nc]
  double *dudt[4] = {};
  double cv = 1.0;
};

struct GraphCache {
  hipGraphExec_t exec = nullptr;
  std::vector<double> key;
  int steps = 0;
};

// Per-kernel timing with HIP events on the context's stream (profiling mode
// runs the step loops eagerly so every launch is bracketed).
enum KernelId {
  K_EUL_DT = 0, K_EUL_GRAD, K_EUL_FLUX, K_EUL_UPDATE,
  K_MAIRE_NODAL, K_MAIRE_RESIDUAL, K_MAIRE_DT, K_MAIRE_MOVE, K_MAIRE_UPDATE, K_MAIRE_COPY,
  K_EUL_FLUXUPD,
  K_COUNT
};

struct ProfSlot {
  std::vector<hipEvent_t> start, stop;
  double total_ms = 0;
  long long launches = 0;
};

}  // namespace fsale

struct fsale_ctx {
  int solver = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  fsale::HostTopology topo;
  fsale::DevMesh m;
  fsale::DevScalars *s = nullptr;      // device
  fsale::DevScalars *h_s = nullptr;    // pinned host mirror
  fsale::EulerFields eu;
  fsale::MaireFields ma;
  bool have_state = false;
  fsale::GraphCache graph;
  std::vector<void *> allocations;
  bool profiling = false;
  fsale::ProfSlot prof[fsale::K_COUNT];
};

namespace fsale {
// RAII bracket of one launch when the context is profiling
struct ProfScope {
  fsale_ctx *ctx;
  int id;
  hipEvent_t stop = nullptr;
  ProfScope(fsale_ctx *c, int kid);
  ~ProfScope();
};

int dev_alloc(fsale_ctx *ctx, void **p, size_t bytes);
template <class T>
int dev_alloc_n(fsale_ctx *ctx, T **p, size_t n) {
  return dev_alloc(ctx, reinterpret_cast<void **>(p), n * sizeof(T) + 16);
}
int read_scalars(fsale_ctx *ctx);      // sync + copy DevScalars to h_s
int check_device_error(fsale_ctx *ctx);

// geometry (geom.hip)
int launch_geometry(fsale_ctx *ctx, bool cells, bool faces);
// Euler (euler.hip)
int euler_alloc(fsale_ctx *ctx);
// Maire (maire.hip)
int maire_alloc(fsale_ctx *ctx);
}  // namespace fsale
