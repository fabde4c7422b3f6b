// Ghost-cell exchange between the contexts of a partitioned mesh.
//
// The reference runs one rank per mesh colour; FleCSI refreshes the ghost
// copies of a dense field after a task has written the owned cells and
// before a task reads the ghosts (apps/hydro/driver.h:329-375: the state
// after apply_update, the gradients between evaluate_gradients and
// evaluate_fluxes, whose shared-cell faces are recomputed on each rank,
// apps/hydro/tasks.h:466-534).  Here each rank's sub-mesh (owned cells
// first, then its ghosts) is a context on the one device, and a halo plan
// copies owner cells' values into the ghost cells of another context with a
// device-side gather / scatter, ordered against both contexts' streams by
// events (no host synchronisation).  On several GPUs the same plan is the
// message of a point-to-point exchange.
#include <vector>

#include "internal.h"

struct fsale_halo {
  fsale_ctx *dst = nullptr;
  fsale_ctx *src = nullptr;
  int n = 0;
  int32_t *d_dst = nullptr, *d_src = nullptr;   // storage positions
  hipEvent_t src_done = nullptr, copy_done = nullptr;
};

namespace fsale {
namespace {

constexpr int kHaloFields = 24;

struct HaloCopy {
  int n, nfield;
  const int32_t *dst_pos, *src_pos;
  const double *src[kHaloFields];
  double *dst[kHaloFields];
};

__global__ void __launch_bounds__(kBlock) k_halo_copy(HaloCopy h) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= h.n) return;
  const int s = h.src_pos[i], d = h.dst_pos[i];
  for (int f = 0; f < h.nfield; ++f) h.dst[f][d] = h.src[f][s];
}

// the fields of an exchange, in the same order for both contexts
int halo_fields(const EulerFields &u, int dim, int what, std::vector<double *> &out) {
  out.clear();
  if (what == FSALE_HALO_STATE) {
    out.push_back(u.d);
    for (int k = 0; k < dim; ++k) out.push_back(u.v[k]);
    for (double *p : {u.e, u.p, u.T, u.a}) out.push_back(p);
  } else if (what == FSALE_HALO_GRADIENTS) {
    for (int k = 0; k < dim; ++k) out.push_back(u.gd[k]);
    for (int k = 0; k < dim * dim; ++k) out.push_back(u.gv[k]);
    for (int k = 0; k < dim; ++k) out.push_back(u.ge[k]);
    for (int k = 0; k < dim; ++k) out.push_back(u.gp[k]);
    for (int k = 0; k < dim; ++k) out.push_back(u.ga[k]);
  } else {
    return FSALE_ERR_ARG;
  }
  for (double *p : out)
    if (!p) return FSALE_ERR_STATE;
  return FSALE_OK;
}

bool plain_euler(const fsale_ctx *c) {
  return c && c->solver == FSALE_SOLVER_EULER && c->single && !c->probs.empty() && c->probs[0].live &&
         c->probs[0].topo;
}

}  // namespace
}  // namespace fsale

using namespace fsale;

extern "C" {

int fsale_halo_create(fsale_ctx *dst, fsale_ctx *src, int32_t n, const int32_t *dst_cells,
                      const int32_t *src_cells, fsale_halo **out) {
  if (!out) return FSALE_ERR_ARG;
  *out = nullptr;
  if (!plain_euler(dst) || !plain_euler(src) || n < 0 || (n > 0 && (!dst_cells || !src_cells))) {
    set_message("a halo joins two Eulerian contexts");
    return FSALE_ERR_ARG;
  }
  if (dst->m.dim != src->m.dim || dst->device != src->device) {
    set_message("halo contexts differ in dimension or device");
    return FSALE_ERR_ARG;
  }
  const Problem &D = dst->probs[0], &S = src->probs[0];
  std::vector<int32_t> dp(n), sp(n);
  for (int i = 0; i < n; ++i) {
    if (dst_cells[i] < 0 || dst_cells[i] >= D.nc || src_cells[i] < 0 || src_cells[i] >= S.nc) {
      set_message("halo cell id out of range");
      return FSALE_ERR_ARG;
    }
    dp[i] = D.topo->cpos[dst_cells[i]] + D.c0;
    sp[i] = S.topo->cpos[src_cells[i]] + S.c0;
  }
  fsale_halo *h = new fsale_halo();
  h->dst = dst;
  h->src = src;
  h->n = n;
  auto fail = [&](hipError_t e) {
    set_message(std::string("halo: ") + hipGetErrorString(e));
    fsale_halo_destroy(h);
    return FSALE_ERR_HIP;
  };
  hipError_t e;
  if (n > 0) {
    if ((e = hipMalloc(&h->d_dst, sizeof(int32_t) * n)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&h->d_src, sizeof(int32_t) * n)) != hipSuccess) return fail(e);
    if ((e = hipMemcpy(h->d_dst, dp.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice)) != hipSuccess)
      return fail(e);
    if ((e = hipMemcpy(h->d_src, sp.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice)) != hipSuccess)
      return fail(e);
  }
  if ((e = hipEventCreateWithFlags(&h->src_done, hipEventDisableTiming)) != hipSuccess) return fail(e);
  if ((e = hipEventCreateWithFlags(&h->copy_done, hipEventDisableTiming)) != hipSuccess) return fail(e);
  *out = h;
  return FSALE_OK;
}

int fsale_halo_exchange(fsale_halo *h, int what) {
  if (!h) return FSALE_ERR_ARG;
  std::vector<double *> fs, fd;
  int rc;
  if ((rc = halo_fields(h->src->eu, h->src->m.dim, what, fs)) ||
      (rc = halo_fields(h->dst->eu, h->dst->m.dim, what, fd))) {
    set_message("halo exchange: unknown field set");
    return rc;
  }
  if (h->n == 0) return FSALE_OK;
  HaloCopy c{};
  c.n = h->n;
  c.nfield = (int)fs.size();
  c.dst_pos = h->d_dst;
  c.src_pos = h->d_src;
  for (int f = 0; f < c.nfield; ++f) {
    c.src[f] = fs[f];
    c.dst[f] = fd[f];
  }
  // the copy waits for what the source's stream has queued, and the source
  // stream's later work (the next write of these fields) waits for the copy
  FSALE_HIP(hipEventRecord(h->src_done, h->src->stream));
  FSALE_HIP(hipStreamWaitEvent(h->dst->stream, h->src_done, 0));
  hipLaunchKernelGGL(k_halo_copy, dim3((h->n + kBlock - 1) / kBlock), dim3(kBlock), 0, h->dst->stream, c);
  FSALE_HIP(hipGetLastError());
  FSALE_HIP(hipEventRecord(h->copy_done, h->dst->stream));
  FSALE_HIP(hipStreamWaitEvent(h->src->stream, h->copy_done, 0));
  return FSALE_OK;
}

int fsale_halo_destroy(fsale_halo *h) {
  if (!h) return FSALE_ERR_ARG;
  if (h->copy_done) (void)hipEventSynchronize(h->copy_done);
  if (h->d_dst) (void)hipFree(h->d_dst);
  if (h->d_src) (void)hipFree(h->d_src);
  if (h->src_done) (void)hipEventDestroy(h->src_done);
  if (h->copy_done) (void)hipEventDestroy(h->copy_done);
  delete h;
  return FSALE_OK;
}

}  // extern "C"
