// Contexts, problems, field I/O and error reporting of the C ABI
// (include/flecsale_amd.h).  A context replaces the FleCSI runtime's field
// storage and mesh client handles (apps/hydro/driver.h:35-158,215-232;
// apps/maire_hydro/driver.h:36-166,237-258): every field of every resident
// problem lives in device memory as SoA FP64 arrays of capacity size.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <type_traits>

#include "internal.h"
#include "kernels_common.h"

namespace fsale {

static thread_local std::string g_message;
void set_message(const std::string &msg) { g_message = msg; }

int RangeAlloc::alloc(int n) {
  if (n <= 0) return 0;
  for (auto it = free.begin(); it != free.end(); ++it) {
    if (it->second >= n) {
      const int start = it->first, len = it->second;
      free.erase(it);
      if (len > n) free[start + n] = len - n;
      return start;
    }
  }
  return -1;
}

bool RangeAlloc::fits(int n) const {
  if (n <= 0) return true;
  for (const auto &kv : free)
    if (kv.second >= n) return true;
  return false;
}

void RangeAlloc::release(int start, int n) {
  if (n <= 0) return;
  auto it = free.emplace(start, n).first;
  auto next = std::next(it);
  if (next != free.end() && it->first + it->second == next->first) {
    it->second += next->second;
    free.erase(next);
  }
  if (it != free.begin()) {
    auto prev = std::prev(it);
    if (prev->first + prev->second == it->first) {
      prev->second += it->second;
      free.erase(it);
    }
  }
}

long long RangeAlloc::available() const {
  long long s = 0;
  for (auto &kv : free) s += kv.second;
  return s;
}

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
  FSALE_HIP(hipMemcpyAsync(ctx->h_s, ctx->s, sizeof(DevScalars) * ctx->max_problems, hipMemcpyDeviceToHost,
                           ctx->stream));
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  // no round is in flight: slots retired during one may be reused
  for (auto &P : ctx->probs) P.quarantined = false;
  return FSALE_OK;
}

int check_device_error(fsale_ctx *ctx, int prob) {
  int rc = read_scalars(ctx);
  if (rc) return rc;
  const DevScalars &S = ctx->h_s[prob];
  if (S.err_code) {
    if (S.err_code == FSALE_ERR_NEG_STATE)
      set_message("Negative density or internal energy encountered! (cell " + std::to_string(S.err_cell) + ")");
    else if (S.err_code == FSALE_ERR_DT)
      set_message("infinite delta t");
    return S.err_code;
  }
  return FSALE_OK;
}

int upload_params(fsale_ctx *ctx, int prob) {
  FSALE_HIP(hipMemcpyAsync(ctx->d_params + prob, &ctx->probs[prob].params, sizeof(ProbParams),
                           hipMemcpyHostToDevice, ctx->stream));
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  return FSALE_OK;
}

int reset_scalars(fsale_ctx *ctx, int prob, double t, double dt) {
  DevScalars h{};
  h.t = t;
  h.dt = dt;
  h.active = 1;
  h.err_code = 0;
  h.err_cell = INT64_MAX;
  h.steps = 0;
  h.err_token = 0;
  ctx->h_s[prob] = h;
  FSALE_HIP(hipMemcpyAsync(ctx->s + prob, &ctx->h_s[prob], sizeof(DevScalars), hipMemcpyHostToDevice,
                           ctx->stream));
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  return FSALE_OK;
}

namespace {

// The 3D fused step over bricks of structured grids: on unless
// FSALE_BRICKS=0 (the gathering kernel then runs, with the same results).
bool brick_enabled() {
  static const bool on = [] {
    const char *e = std::getenv("FSALE_BRICKS");
    return !(e && !std::strcmp(e, "0"));
  }();
  return on;
}

// Grid of a one-block-per-tile launch over `live` tiles: a power of two (so a
// captured graph stays valid while the batch changes within a factor of
// two), at least 8, at most the tile capacity rounded up to 8.
int direct_grid(int live, int cap_tiles) {
  long long g = 8;
  while (g < live) g *= 2;
  const long long cap = ((long long)(cap_tiles + 7) / 8) * 8;
  if (g > cap) g = cap;
  return (int)std::max(8LL, g);
}

// Dispatch list of a one-block-per-tile launch with grid G: block b runs on
// XCD b % 8 and gets tile (b % 8) * chunk + b / 8, so each XCD works on one
// contiguous range of the list; blocks without a tile get count 0.
void build_direct(const std::vector<Tile> &tiles, int G, std::vector<Tile> &out) {
  const int T = (int)tiles.size();
  const int chunk = (T + 7) / 8;
  out.assign(G, Tile{0, 0, 0, 0});
  for (int b = 0; b < G; ++b) {
    const int x = b & 7, i = b >> 3;
    const int t = x * chunk + i;
    if (i < chunk && t < T) out[b] = tiles[t];
  }
}

}  // namespace

// Tile lists of the live problems, problems in id order, each problem's
// tiles contiguous, plus their one-block-per-tile dispatch lists.  Uploaded
// only when the set of problems changed.
int sync_tiles(fsale_ctx *ctx) {
  if (!ctx->tiles_dirty) return FSALE_OK;
  ctx->h_ctiles.clear();
  ctx->h_vtiles.clear();
  ctx->h_btiles.clear();
  ctx->h_bbtiles.clear();
  for (int p = 0; p < (int)ctx->probs.size(); ++p) {
    const Problem &P = ctx->probs[p];
    if (!P.live) continue;
    const std::vector<int32_t> *ord = P.topo && !P.topo->tile_order.empty() ? &P.topo->tile_order : nullptr;
    for (int i = 0; i < P.nct; ++i) {
      const int j = ord ? (*ord)[i] : i;
      ctx->h_ctiles.push_back(Tile{p, P.c0 + j * kBlock, std::min(kBlock, P.nc - j * kBlock), P.ct0 + j});
    }
    for (int i = 0; i < P.nvt; ++i)
      ctx->h_vtiles.push_back(Tile{p, P.v0 + i * kBlock, std::min(kBlock, P.nv - i * kBlock), P.vt0 + i});
    for (int i = 0; i * kBlock < P.nb; ++i)
      ctx->h_btiles.push_back(Tile{p, P.v0 + i * kBlock, std::min(kBlock, P.nb - i * kBlock), P.vt0 + i});
    // one pass of a block each (kBlock / kMaxCornersPerVertex vertices):
    // their solves are long serial chains, so spread over as many blocks
    constexpr int kBig = kBlock / kMaxCornersPerVertex;
    for (int i = 0; i * kBig < P.nb - P.nb2; ++i)
      ctx->h_bbtiles.push_back(Tile{p, P.v0 + P.nb2 + i * kBig, std::min(kBig, P.nb - P.nb2 - i * kBig), P.vt0});
  }
  std::vector<int> plist;
  ctx->lds = ctx->lds_all = ctx->m.lnb != nullptr;
  ctx->s3 = false;
  ctx->s3_all = ctx->solver == FSALE_SOLVER_EULER && ctx->m.dim == 3;
  ctx->fp3 = false;
  ctx->brick_all = ctx->solver == FSALE_SOLVER_EULER && ctx->m.dim == 3 && ctx->m.brick && brick_enabled();
  for (int p = 0; p < (int)ctx->probs.size(); ++p)
    if (ctx->probs[p].live) {
      plist.push_back(p);
      if (!ctx->probs[p].brick) ctx->brick_all = false;
      if (!ctx->probs[p].lds) ctx->lds = false;
      if (!ctx->probs[p].lds_all) ctx->lds_all = false;
      if (ctx->probs[p].params.s_fp) ctx->fp3 = true;
      if (ctx->probs[p].params.s_nx)
        ctx->s3 = true;
      else
        ctx->s3_all = false;
    }
  if (!ctx->lds) ctx->lds_all = false;
  if (!ctx->s3) ctx->s3_all = false;
  if (plist.empty()) ctx->brick_all = false;
  const int counts[5] = {(int)ctx->h_ctiles.size(), (int)ctx->h_vtiles.size(), (int)plist.size(),
                         (int)ctx->h_btiles.size(), (int)ctx->h_bbtiles.size()};
  ctx->n_bbtiles = counts[4];
  if (!plist.empty())
    FSALE_HIP(hipMemcpyAsync(ctx->d_plist, plist.data(), sizeof(int) * plist.size(), hipMemcpyHostToDevice,
                             ctx->stream));
  int pg = 1;
  while (pg < (int)plist.size() && pg < 1024) pg *= 2;
  ctx->pgrid = pg;
  if (!ctx->h_ctiles.empty())
    FSALE_HIP(hipMemcpyAsync(ctx->d_ctiles, ctx->h_ctiles.data(), sizeof(Tile) * ctx->h_ctiles.size(),
                             hipMemcpyHostToDevice, ctx->stream));
  if (!ctx->h_vtiles.empty())
    FSALE_HIP(hipMemcpyAsync(ctx->d_vtiles, ctx->h_vtiles.data(), sizeof(Tile) * ctx->h_vtiles.size(),
                             hipMemcpyHostToDevice, ctx->stream));
  FSALE_HIP(hipMemcpyAsync(ctx->d_counts, counts, sizeof(counts), hipMemcpyHostToDevice, ctx->stream));
  std::vector<Tile> dc, dv;
  ctx->n_cdirect = direct_grid(counts[0], ctx->ra_ct.capacity);
  build_direct(ctx->h_ctiles, ctx->n_cdirect, dc);
  FSALE_HIP(hipMemcpyAsync(ctx->d_cdirect, dc.data(), sizeof(Tile) * dc.size(), hipMemcpyHostToDevice, ctx->stream));
  std::vector<Tile> db;
  if (ctx->d_vdirect) {
    ctx->n_vdirect = direct_grid(counts[1], ctx->ra_vt.capacity);
    build_direct(ctx->h_vtiles, ctx->n_vdirect, dv);
    FSALE_HIP(hipMemcpyAsync(ctx->d_vdirect, dv.data(), sizeof(Tile) * dv.size(), hipMemcpyHostToDevice,
                             ctx->stream));
    if (!ctx->h_btiles.empty())
      FSALE_HIP(hipMemcpyAsync(ctx->d_btiles, ctx->h_btiles.data(), sizeof(Tile) * ctx->h_btiles.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    if (!ctx->h_bbtiles.empty())
      FSALE_HIP(hipMemcpyAsync(ctx->d_bbtiles, ctx->h_bbtiles.data(), sizeof(Tile) * ctx->h_bbtiles.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    ctx->n_bdirect = direct_grid(counts[3], ctx->ra_vt.capacity);
    build_direct(ctx->h_btiles, ctx->n_bdirect, db);
    FSALE_HIP(hipMemcpyAsync(ctx->d_bdirect, db.data(), sizeof(Tile) * db.size(), hipMemcpyHostToDevice,
                             ctx->stream));
  }
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  ctx->tiles_dirty = false;
  return FSALE_OK;
}

// Arguments of a launch of kernel `kid`: one-block-per-tile kernels (shape
// 0) read the dispatch lists, persistent ones loop over the tile lists.
Batch batch_args(const fsale_ctx *ctx, int kid) {
  Batch b;
  b.ct.tiles = ctx->d_ctiles;
  b.ct.n = ctx->d_counts;
  b.vt.tiles = ctx->d_vtiles;
  b.vt.n = ctx->d_counts + 1;
  b.bt.tiles = ctx->d_btiles;
  b.bt.n = ctx->d_counts + 3;
  if (kid >= 0 && kid < K_COUNT && ctx->shape.blocks_per_cu[kid] <= 0) {
    b.ct.direct = ctx->d_cdirect;
    b.ct.ndirect = ctx->n_cdirect;
    b.vt.direct = ctx->d_vdirect;
    b.vt.ndirect = ctx->d_vdirect ? ctx->n_vdirect : 0;
    b.bt.direct = ctx->d_bdirect;
    b.bt.ndirect = ctx->d_bdirect ? ctx->n_bdirect : 0;
  }
  b.pp = ctx->d_params;
  b.s = ctx->s;
  b.cap_c_tiles = ctx->ra_ct.capacity;
  b.plist = ctx->d_plist;
  b.np = ctx->d_counts + 2;
  return b;
}

// Grid of a tile-list kernel, a multiple of 8 (one slice per XCD): shape 0
// is one block per tile (the dispatch lists' grid), shape n a persistent
// grid of CUs x n blocks, never more blocks than the capacity has tiles.
int persistent_grid(const fsale_ctx *ctx, int kid) {
  const bool vertex_list = kid == K_MAIRE_NODAL || kid == K_MAIRE_MOVE || kid == K_MAIRE_NODAL_BND;
  // the Lagrangian update also moves the vertices (a 2D driver step's
  // k_maire_move_update walks both lists)
  const bool both = kid == K_MAIRE_COPY || (kid == K_MAIRE_UPDATE && ctx->solver == FSALE_SOLVER_MAIRE);
  const int bpc = (kid >= 0 && kid < K_COUNT) ? ctx->shape.blocks_per_cu[kid] : 0;
  if (bpc <= 0) {
    const int g = kid == K_MAIRE_NODAL_BND ? ctx->n_bdirect
                  : vertex_list           ? ctx->n_vdirect
                  : both                  ? std::max(ctx->n_cdirect, ctx->n_vdirect)
                                          : ctx->n_cdirect;
    return std::max(8, g);
  }
  const int cap_tiles = vertex_list ? ctx->ra_vt.capacity
                        : both      ? std::max(ctx->ra_ct.capacity, ctx->ra_vt.capacity)
                                    : ctx->ra_ct.capacity;
  long long g = (long long)ctx->num_cus * bpc;
  const long long cap = ((long long)(cap_tiles + 7) / 8) * 8;
  if (g > cap) g = cap;
  return (int)std::max(8LL, (g / 8) * 8);
}

namespace {

// Uploads of one admission: every array is copied into a region of the
// context's pinned staging buffer and its transfer is queued at once on the
// upload stream (stream order keeps it ahead of the geometry and initial-state
// kernels queued there), so an admission overlaps a running step round; one
// synchronisation of the upload stream at the end makes the new problem's
// data visible to later work and releases the buffer.  A running round
// cannot see the new problem: its tiles and its entry in the problem list are
// uploaded by the next round's launch.
// Host threads that fill the staging rows of an admission in parallel; the
// calling thread works too.  Created on first use, kept for the process.
class RowPool {
 public:
  static RowPool &get() {
    static RowPool *pool = new RowPool();
    return *pool;
  }
  void run(std::vector<std::function<void()>> &jobs) {
    if (jobs.empty()) return;
    if (workers_ == 0 || jobs.size() == 1) {
      for (auto &j : jobs) j();
      return;
    }
    {
      std::lock_guard<std::mutex> lock(mu_);
      jobs_ = &jobs;
      next_ = 0;
      pending_ = jobs.size();
      ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lock(mu_);
    done_.wait(lock, [&] { return pending_ == 0; });
    jobs_ = nullptr;
  }

 private:
  RowPool() {
    const unsigned hw = std::thread::hardware_concurrency();
    workers_ = (int)std::min<unsigned>(hw > 1 ? hw - 1 : 0, 7);
    for (int i = 0; i < workers_; ++i) std::thread([this] { loop(); }).detach();
  }
  void work() {
    for (;;) {
      size_t i;
      std::vector<std::function<void()>> *jobs;
      {
        std::lock_guard<std::mutex> lock(mu_);
        if (!jobs_ || next_ >= jobs_->size()) return;
        i = next_++;
        jobs = jobs_;
      }
      (*jobs)[i]();
      std::lock_guard<std::mutex> lock(mu_);
      if (--pending_ == 0) done_.notify_all();
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [&] { return gen_ != seen; });
        seen = gen_;
      }
      work();
    }
  }
  int workers_ = 0;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  std::vector<std::function<void()>> *jobs_ = nullptr;
  size_t next_ = 0, pending_ = 0;
  uint64_t gen_ = 0;
};

struct Stager {
  fsale_ctx *ctx;
  size_t used = 0;
  // rows whose contents are computed on the row pool before their transfer
  struct Deferred {
    void *dst;
    const void *src;
    size_t bytes;
  };
  std::vector<Deferred> deferred;
  std::vector<std::function<void()>> fills;
  explicit Stager(fsale_ctx *c) : ctx(c) {}
  int reserve(size_t bytes) {
    bytes += 64 * 64;
    if (bytes <= ctx->stage_bytes) return FSALE_OK;
    FSALE_HIP(hipStreamSynchronize(ctx->upload));
    if (ctx->stage) (void)hipHostFree(ctx->stage);
    ctx->stage = nullptr;
    ctx->stage_bytes = 0;
    FSALE_HIP(hipHostMalloc(&ctx->stage, bytes, hipHostMallocDefault));
    ctx->stage_bytes = bytes;
    return FSALE_OK;
  }
  // a region of n elements, to be filled by the caller before put()
  template <class T>
  T *slot(size_t n) {
    const size_t off = (used + 63) & ~size_t(63);
    if (off + n * sizeof(T) > ctx->stage_bytes) return nullptr;
    used = off + n * sizeof(T);
    return reinterpret_cast<T *>(static_cast<char *>(ctx->stage) + off);
  }
  template <class T>
  int put(T *dst, const T *region, size_t n) {
    if (n) FSALE_HIP(hipMemcpyAsync(dst, region, n * sizeof(T), hipMemcpyHostToDevice, ctx->upload));
    return FSALE_OK;
  }
  template <class T>
  int copy(T *dst, const T *src, size_t n) {
    T *r = slot<T>(n);
    if (!r) {
      set_message("staging buffer overflow");
      return FSALE_ERR_ARG;
    }
    std::memcpy(r, src, n * sizeof(T));
    return put(dst, r, n);
  }
  // a region of n elements filled later by fill(region, n), then sent to dst
  template <class T, class F>
  int defer(T *dst, size_t n, F fill) {
    T *r = slot<T>(n);
    if (!r) {
      set_message("staging buffer overflow");
      return FSALE_ERR_ARG;
    }
    deferred.push_back(Deferred{dst, r, n * sizeof(T)});
    fills.emplace_back([r, n, fill]() { fill(r, n); });
    return FSALE_OK;
  }
  // drops rows not yet filled (their fill functions may refer to locals of a
  // caller that failed)
  void discard() {
    fills.clear();
    deferred.clear();
  }
  int run_deferred() {
    RowPool::get().run(fills);
    fills.clear();
    for (size_t i = 0; i < deferred.size(); ++i) {
      const Deferred &d = deferred[i];
      if (!d.bytes) continue;
      const hipError_t e = hipMemcpyAsync(d.dst, d.src, d.bytes, hipMemcpyHostToDevice, ctx->upload);
      if (e != hipSuccess) {
        set_message(std::string("upload ") + std::to_string(i) + " of " + std::to_string(deferred.size()) + " (" +
                    std::to_string(d.bytes) + " bytes): " + hipGetErrorString(e));
        deferred.clear();
        return e == hipErrorOutOfMemory ? FSALE_ERR_OOM : FSALE_ERR_HIP;
      }
    }
    deferred.clear();
    return FSALE_OK;
  }
  int finish() {
    int rc = run_deferred();
    FSALE_HIP(hipStreamSynchronize(ctx->upload));
    used = 0;
    return rc;
  }
};

inline int vt_popcount(unsigned x) { return (int)(x & 1) + (int)((x >> 1) & 1) + (int)((x >> 2) & 1); }

int alloc_batch(fsale_ctx *ctx, int solver, int dim, int vpc, int max_problems, long long cap_c, long long cap_f,
                long long cap_v, long long cap_n) {
  if (!((dim == 2 && vpc == 4) || (dim == 3 && vpc == 8))) {
    set_message("unsupported cell type (quads in 2D, hexes in 3D)");
    return FSALE_ERR_MESH;
  }
  // entity capacities below 2^28: kernels address an 8-byte element of an
  // entity array with a 32-bit byte offset
  if (max_problems <= 0 || cap_c <= 0 || cap_f <= 0 || cap_v <= 0 || cap_n < 0 || cap_c > (1 << 28) ||
      cap_f > (1 << 28) || cap_v > (1 << 28) || cap_n > INT32_MAX / 2) {
    set_message("bad batch capacity");
    return FSALE_ERR_ARG;
  }
  ctx->solver = solver;
  ctx->max_problems = max_problems;
  DevMesh &m = ctx->m;
  m.dim = dim;
  m.vpc = vpc;
  m.fpc = (dim == 2) ? vpc : 6;
  m.vpf = (dim == 2) ? 2 : 4;
  m.cap_c = (int)cap_c;
  m.cap_f = (int)cap_f;
  m.cap_v = (int)cap_v;
  ctx->ra_c.init((int)cap_c);
  ctx->ra_f.init((int)cap_f);
  ctx->ra_v.init((int)cap_v);
  ctx->ra_n.init((int)cap_n);
  const int cap_ct = (int)((cap_c + kBlock - 1) / kBlock) + max_problems;
  const int cap_vt = (int)((cap_v + kBlock - 1) / kBlock) + max_problems;
  ctx->ra_ct.init(cap_ct);
  ctx->ra_vt.init(cap_vt);
  int rc;
  for (int d = 0; d < dim; ++d) {
    if ((rc = dev_alloc_n(ctx, &m.x[d], cap_v))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.ccen[d], cap_c))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.fcen[d], cap_f))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.fnrm[d], cap_f))) return rc;
  }
  if ((rc = dev_alloc_n(ctx, &m.cv, (size_t)cap_c * vpc))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.cf, (size_t)cap_c * m.fpc))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.fc, (size_t)cap_f * 2))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.fv, (size_t)cap_f * m.vpf))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.nbr_start, cap_c))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.nbr_cnt, cap_c))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.nbr, cap_n + 1))) return rc;
  m.nbk_deg = dim == 2 ? kNbk2 : kNbk3;
  if ((rc = dev_alloc_n(ctx, &m.nbk, (size_t)cap_c * m.nbk_deg))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.cfn, (size_t)cap_c * m.fpc))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.corig, cap_c))) return rc;
  if (dim == 3 && solver == FSALE_SOLVER_EULER) {
    for (int i = 0; i < 3; ++i)
      if ((rc = dev_alloc_n(ctx, &m.kfp[i], (size_t)cap_c * m.fpc))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.kfa, (size_t)cap_c * m.fpc))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.brick, (size_t)cap_ct))) return rc;
  }
  if (dim == 2)
    if ((rc = dev_alloc_n(ctx, &m.xy, cap_v))) return rc;
  if (dim == 2 && vpc == 4 && solver == FSALE_SOLVER_EULER) {
    if ((rc = dev_alloc_n(ctx, &m.halo_cnt, (size_t)cap_ct))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.halo, (size_t)cap_ct * kHalo2))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.lnb, (size_t)cap_c * (kNbk2 / 2)))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.lfn, (size_t)cap_c * 2))) return rc;
  }
  if ((rc = dev_alloc_n(ctx, &m.cvol, cap_c))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.cminlen, cap_c))) return rc;
  if ((rc = dev_alloc_n(ctx, &m.farea, cap_f))) return rc;
  if (solver == FSALE_SOLVER_MAIRE) {
    if ((rc = dev_alloc_n(ctx, &m.vc_cnt, cap_v))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.vck, (size_t)cap_v * kMaxCornersPerVertex))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.cwtag, (size_t)cap_c * vpc))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.vbnd, cap_v))) return rc;
    if ((rc = dev_alloc_n(ctx, &m.bv, cap_v))) return rc;
  }
  if ((rc = dev_alloc_n(ctx, &ctx->d_ctiles, cap_ct))) return rc;
  if ((rc = dev_alloc_n(ctx, &ctx->d_vtiles, cap_vt))) return rc;
  if ((rc = dev_alloc_n(ctx, &ctx->d_cdirect, ((size_t)cap_ct + 7) / 8 * 8))) return rc;
  if (solver == FSALE_SOLVER_MAIRE) {
    if ((rc = dev_alloc_n(ctx, &ctx->d_vdirect, ((size_t)cap_vt + 7) / 8 * 8))) return rc;
    if ((rc = dev_alloc_n(ctx, &ctx->d_btiles, cap_vt))) return rc;
    // tiles of kBlock / kMaxCornersPerVertex boundary vertices (sync_tiles)
    const size_t cap_bb = ((size_t)cap_v + 31) / 32 + (size_t)max_problems;
    static_assert(kBlock / kMaxCornersPerVertex == 32, "one block pass of 3D vertices");
    if ((rc = dev_alloc_n(ctx, &ctx->d_bbtiles, cap_bb))) return rc;
    if ((rc = dev_alloc_n(ctx, &ctx->d_bdirect, ((size_t)cap_vt + 7) / 8 * 8))) return rc;
  }
  if ((rc = dev_alloc_n(ctx, &ctx->d_counts, 5))) return rc;
  if ((rc = dev_alloc_n(ctx, &ctx->d_plist, max_problems))) return rc;
  if ((rc = dev_alloc_n(ctx, &ctx->d_params, max_problems))) return rc;
  if ((rc = dev_alloc_n(ctx, &ctx->s, max_problems))) return rc;
  FSALE_HIP(hipHostMalloc((void **)&ctx->h_s, sizeof(DevScalars) * max_problems, hipHostMallocDefault));
  for (int p = 0; p < max_problems; ++p) {
    ctx->h_s[p] = DevScalars{};
    ctx->h_s[p].err_cell = INT64_MAX;
  }
  FSALE_HIP(hipMemcpy(ctx->s, ctx->h_s, sizeof(DevScalars) * max_problems, hipMemcpyHostToDevice));
  std::vector<ProbParams> zero(max_problems);
  std::memset(zero.data(), 0, sizeof(ProbParams) * max_problems);
  FSALE_HIP(hipMemcpy(ctx->d_params, zero.data(), sizeof(ProbParams) * max_problems, hipMemcpyHostToDevice));
  const int counts[5] = {0, 0, 0, 0, 0};
  FSALE_HIP(hipMemcpy(ctx->d_counts, counts, sizeof(counts), hipMemcpyHostToDevice));
  ctx->probs.assign(max_problems, Problem{});
  if (solver == FSALE_SOLVER_MAIRE) return maire_alloc(ctx);
  return euler_alloc(ctx);
}

// places one mesh in the batch: ranges, connectivity (offset to batch ids),
// coordinates, geometry, boundary tags, parameters and scalars, all through
// `st` (the caller finishes it)
int place_mesh(fsale_ctx *ctx, const fsale_mesh_desc &mesh, int *out_prob, Stager &st) {
  DevMesh &m = ctx->m;
  if (mesh.dim != m.dim || mesh.verts_per_cell != m.vpc) {
    set_message("mesh type differs from the batch's");
    return FSALE_ERR_MESH;
  }
  if (!mesh.coords || !mesh.cells) return FSALE_ERR_ARG;
  int prob = -1;
  for (int p = 0; p < ctx->max_problems; ++p)
    if (!ctx->probs[p].live && !ctx->probs[p].quarantined) {
      prob = p;
      break;
    }
  if (prob < 0) {
    set_message("no free problem slot");
    return FSALE_ERR_OOM;
  }
  // capacity known from the descriptor alone: fail before any connectivity work
  if (!ctx->ra_c.fits(mesh.num_cells) || !ctx->ra_v.fits(mesh.num_vertices) ||
      !ctx->ra_ct.fits((mesh.num_cells + kBlock - 1) / kBlock) ||
      !ctx->ra_vt.fits((mesh.num_vertices + kBlock - 1) / kBlock)) {
    set_message("batch capacity exhausted");
    return FSALE_ERR_OOM;
  }
  int rc;
  std::string err;
  std::shared_ptr<const HostTopology> topo = cached_topology(mesh, rc, err);
  if (rc) {
    set_message(err);
    return rc;
  }
  const HostTopology &t = *topo;
  if (t.max_corners > kMaxCornersPerVertex) {
    set_message("more than 8 cells around a vertex");
    return FSALE_ERR_MESH;
  }
  Problem P;
  P.nc = t.nc;
  P.nf = t.nf;
  P.nv = t.nv;
  P.nn = (int)t.nbr.size();
  P.nct = (t.nc + kBlock - 1) / kBlock;
  P.nvt = (t.nv + kBlock - 1) / kBlock;
  P.c0 = ctx->ra_c.alloc(P.nc);
  P.f0 = ctx->ra_f.alloc(P.nf);
  P.v0 = ctx->ra_v.alloc(P.nv);
  P.n0 = ctx->ra_n.alloc(P.nn);
  P.ct0 = ctx->ra_ct.alloc(P.nct);
  P.vt0 = ctx->ra_vt.alloc(P.nvt);
  if (P.c0 < 0 || P.f0 < 0 || P.v0 < 0 || P.n0 < 0 || P.ct0 < 0 || P.vt0 < 0) {
    if (P.c0 >= 0) ctx->ra_c.release(P.c0, P.nc);
    if (P.f0 >= 0) ctx->ra_f.release(P.f0, P.nf);
    if (P.v0 >= 0) ctx->ra_v.release(P.v0, P.nv);
    if (P.n0 >= 0) ctx->ra_n.release(P.n0, P.nn);
    if (P.ct0 >= 0) ctx->ra_ct.release(P.ct0, P.nct);
    if (P.vt0 >= 0) ctx->ra_vt.release(P.vt0, P.nvt);
    set_message("batch capacity exhausted");
    return FSALE_ERR_OOM;
  }
  P.topo = topo;
  // mark the slot taken before any failure so that the cleanup releases it
  P.live = true;
  P.params = ProbParams{};
  P.params.cell_first = P.c0;
  P.params.cell_count = P.nc;
  P.params.tile_first = P.ct0;
  P.params.tile_count = P.nct;
  if (ctx->solver == FSALE_SOLVER_EULER) {
    P.params.s_nx = topo->s_nx;
    P.params.s_np = topo->s_np;
    P.params.s_fp = topo->s_fp;
  }
  ctx->probs[prob] = P;
  ctx->tiles_dirty = true;
  *out_prob = prob;
  const int nc = t.nc, nf = t.nf, nv = t.nv;
  const bool maire = ctx->solver == FSALE_SOLVER_MAIRE;
  // what the kernels of this context read: batched 2D Eulerian and Lagrangian
  // steps rebuild face geometry from the vertices (the plain drop-in context
  // keeps every face array for its task-level entry points and field reads);
  // the CSR neighbour list only for cells past the k-major head
  const bool faces = ctx->single || m.dim == 3;
  const bool cell_faces = faces || !maire;
  const bool nbr_tail = !maire && (ctx->single || t.max_neighbours > m.nbk_deg);
  ctx->probs[prob].faces_on_device = faces;
  // staging: connectivity rows, coordinates, Lagrangian lists and tags, and
  // room for the caller's state arrays, parameters and scalars
  size_t bytes = 4 * ((size_t)nc * (t.vpc + 2 * t.fpc + m.nbk_deg + 2) + (size_t)nf * (2 + t.vpf) + P.nn) +
                 8 * (size_t)nv * m.dim + 8 * (size_t)nc * (m.dim + 3) + sizeof(ProbParams) + sizeof(DevScalars);
  bytes += 4 * (size_t)nc;   // original ids
  const bool lds = m.lnb && !t.halo_cnt.empty();
  if (lds) bytes += 4 * (size_t)P.nct * (1 + kHalo2);
  const bool brick = m.brick && !t.brick.empty() && (int)t.brick.size() == P.nct;
  if (brick) bytes += 4 * (size_t)P.nct;
  if (maire) bytes += 4 * (size_t)nv * (2 + kMaxCornersPerVertex) + 2 * (size_t)nc * t.vpc + nv + 64 * t.vpc;
  if (maire && m.dim == 3) bytes += 32 * (size_t)nc + 64;
  if ((rc = st.reserve(bytes + 64 * 64))) return rc;
  // k-rows of connectivity, filled on the row pool
  auto rows = [&](int32_t *dst_base, int rows_n, size_t stride, int n, auto value) -> int {
    for (int k = 0; k < rows_n; ++k) {
      int rc2 = st.defer(dst_base + (size_t)k * stride, (size_t)n, [value, k](int32_t *r, size_t cnt) {
        for (size_t i = 0; i < cnt; ++i) r[i] = value(k, (int)i);
      });
      if (rc2) return rc2;
    }
    return FSALE_OK;
  };
  // connectivity with batch ids in storage order (row position s holds the
  // entity t.?idx[s]; every id it stores is a storage position), one k-row
  // segment at a time
  const int32_t *cidx = t.cidx.data(), *cpos = t.cpos.data(), *vidx = t.vidx.data(), *vpos = t.vpos.data(),
                *fidx = t.fidx.data(), *fpos = t.fpos.data();
  const int c0 = P.c0, v0 = P.v0, f0 = P.f0, n0 = P.n0;
  if ((rc = rows(m.cv + c0, t.vpc, m.cap_c, nc,
                 [&t, cidx, vpos, v0](int k, int s) { return vpos[t.cv[(size_t)cidx[s] * t.vpc + k]] + v0; })))
    return rc;
  if (faces && (rc = rows(m.cf + c0, t.fpc, m.cap_c, nc, [&t, cidx, fpos, f0](int k, int s) {
                  const int c = cidx[s];
                  const int f = t.cf[(size_t)c * t.fpc + k];
                  return 2 * (fpos[f] + f0) + (t.fc[2 * (size_t)f] == c ? 0 : 1);
                })))
    return rc;
  if (cell_faces && (rc = rows(m.cfn + c0, t.fpc, m.cap_c, nc, [&t, cidx, cpos, c0](int k, int s) {
                       const int c = cidx[s];
                       const int f = t.cf[(size_t)c * t.fpc + k];
                       const bool second = t.fc[2 * (size_t)f] != c;
                       const int other = second ? t.fc[2 * (size_t)f] : t.fc[2 * (size_t)f + 1];
                       return (other >= 0 ? cpos[other] + c0 + 1 : 0) * 2 + (second ? 1 : 0);
                     })))
    return rc;
  if (ctx->single && (rc = rows(m.fc + f0, 2, m.cap_f, nf, [&t, fidx, cpos, c0](int k, int s) {
                        const int c = t.fc[2 * (size_t)fidx[s] + k];
                        return c >= 0 ? cpos[c] + c0 : -1;
                      })))
    return rc;
  if (faces && (rc = rows(m.fv + f0, t.vpf, m.cap_f, nf, [&t, fidx, vpos, v0](int k, int s) {
                  return vpos[t.fv[(size_t)fidx[s] * t.vpf + k]] + v0;
                })))
    return rc;
  if ((rc = rows(m.corig + c0, 1, 0, nc, [cidx](int, int s) { return cidx[s]; }))) return rc;
  if (!maire) {
    if ((rc = rows(m.nbr_cnt + c0, 1, 0, nc,
                   [&t, cidx](int, int s) { return t.nbr_off[cidx[s] + 1] - t.nbr_off[cidx[s]]; })))
      return rc;
    // right-aligned: a cell with fewer neighbours than slots has its -1
    // padding first (the gradient kernel reads the cell itself there, which
    // leaves its min / max and sums exactly as they start)
    const int deg = m.nbk_deg;
    if ((rc = rows(m.nbk + c0, deg, m.cap_c, nc, [&t, cidx, cpos, c0, deg](int k, int s) {
           const int c = cidx[s];
           const int cnt = t.nbr_off[c + 1] - t.nbr_off[c];
           const int pad = cnt < deg ? deg - cnt : 0;
           return k >= pad ? cpos[t.nbr[t.nbr_off[c] + k - pad]] + c0 : -1;
         })))
      return rc;
  }
  ctx->probs[prob].lds = lds;
  ctx->probs[prob].lds_all =
      lds && std::all_of(t.halo_cnt.begin(), t.halo_cnt.end(), [](int32_t h) { return h >= 0; });
  if (lds) {
    // the tiles' LDS staging plan (tile slots from P.ct0), halo cells as batch ids
    const int ct0 = P.ct0;
    if ((rc = rows(m.halo_cnt + ct0, 1, 0, P.nct, [&t](int, int j) { return t.halo_cnt[j]; }))) return rc;
    if ((rc = rows(m.halo + (size_t)ct0 * kHalo2, 1, 0, P.nct * kHalo2,
                   [&t, c0](int, int i) { return t.halo[i] + c0; })))
      return rc;
  }
  ctx->probs[prob].brick = brick;
  if (brick && (rc = rows(m.brick + P.ct0, 1, 0, P.nct, [&t, c0](int, int j) { return t.brick[j] + c0; })))
    return rc;
  if (nbr_tail) {
    if ((rc = rows(m.nbr_start + c0, 1, 0, nc, [&t, cidx, n0](int, int s) { return t.nbr_off[cidx[s]] + n0; })))
      return rc;
    if ((rc = rows(m.nbr + n0, 1, 0, P.nn, [&t, cpos, c0](int, int i) { return cpos[t.nbr[i]] + c0; }))) return rc;
  }
  for (int d = 0; d < m.dim; ++d) {
    const double *xs = mesh.coords;
    const int dim = m.dim;
    if ((rc = st.defer(m.x[d] + v0, (size_t)nv, [xs, dim, d, vidx](double *x, size_t n) {
           for (size_t s = 0; s < n; ++s) x[s] = xs[(size_t)vidx[s] * dim + d];
         })))
      return rc;
  }
  if ((rc = st.run_deferred())) return rc;
  if (lds && (rc = euler_lds_slots(ctx, prob, ctx->upload))) return rc;
  if ((rc = launch_geometry(ctx, prob, true, faces, ctx->upload))) return rc;
  if ((rc = maire_pack_problem(ctx, prob, ctx->upload))) return rc;
  if (maire) {
    // vertex -> corner lists and boundary tags from the initial face normals:
    // tag d when |n_d| is 1 within 1e-14 (apps/maire_hydro/2d/inputs_regression.cc:134-162);
    // the normals are evaluated here with the device's formula
    if ((rc = rows(m.vc_cnt + v0, 1, 0, nv,
                   [&t, vidx](int, int s) { return t.vc_off[vidx[s] + 1] - t.vc_off[vidx[s]]; })))
      return rc;
    if (m.dim == 3) {
      // the 3D nodal solve's packed cell -> vertex rows, [cell][8]
      if ((rc = st.defer(ctx->ma.cv8 + 8 * (size_t)c0, 8 * (size_t)nc, [&t, cidx, vpos, v0](int32_t *r, size_t n) {
             for (size_t i = 0; i < n; ++i) r[i] = vpos[t.cv[(size_t)cidx[i / 8] * 8 + i % 8]] + v0;
           })))
        return rc;
    }
    if ((rc = rows(m.vck + v0, kMaxCornersPerVertex, m.cap_v, nv, [&t, vidx, cpos, c0](int j, int s) {
           const int v = vidx[s];
           if (j >= t.vc_off[v + 1] - t.vc_off[v]) return -1;
           const int corner = t.vc[t.vc_off[v] + j];
           const int c = corner / t.vpc, k = corner - c * t.vpc;
           return (cpos[c] + c0) * t.vpc + k;
         })))
      return rc;
    // face tags: bit 0-2 = |n_x|, |n_y|, |n_z| is 1, bit 3 = boundary face
    std::vector<uint8_t> ftag(nf, 0), vb(nv, 0), vtags(nv, 0);
    const double *xs = mesh.coords;
    for (int f = 0; f < nf; ++f) {
      if (t.fc[2 * (size_t)f + 1] >= 0) continue;
      double n[3] = {0, 0, 0};
      if (m.dim == 2) {
        const int a = t.fv[2 * (size_t)f], b = t.fv[2 * (size_t)f + 1];
        double cx, cy, area;
        face_geom_2d(xs[2 * (size_t)a], xs[2 * (size_t)a + 1], xs[2 * (size_t)b], xs[2 * (size_t)b + 1], cx, cy,
                     n[0], n[1], area);
      } else {
        double q[4][3], cen[3], area;
        for (int i = 0; i < 4; ++i)
          for (int d = 0; d < 3; ++d) q[i][d] = xs[3 * (size_t)t.fv[4 * (size_t)f + i] + d];
        quad3_geom(q, cen, n, area);
      }
      uint8_t tag = 0x8;
      for (int d = 0; d < m.dim; ++d)
        if (std::fabs(std::fabs(n[d]) - 1.0) <= 1e-14) tag |= (uint8_t)(1u << d);
      ftag[f] = tag;
      for (int i = 0; i < t.vpf; ++i) {
        vb[t.fv[(size_t)f * t.vpf + i]] = 1;
        vtags[t.fv[(size_t)f * t.vpf + i]] |= (uint8_t)(tag & 7);
      }
    }
    // corner nibbles: 2D (face of edge k-1 -> k, face of edge k -> k+1) with
    // the boundary bit at 2; 3D the three sides holding k, in side order
    for (int k = 0; k < t.vpc; ++k) {
      uint16_t *cw = st.slot<uint16_t>(nc);
      if (!cw) return FSALE_ERR_ARG;
      int sides[3], ns = 0;
      if (m.dim == 2) {
        sides[0] = (k + t.vpc - 1) % t.vpc;
        sides[1] = k;
        ns = 2;
      } else {
        for (int lf = 0; lf < 6; ++lf)
          for (int i = 0; i < 4; ++i)
            if (hex_face_vertex(lf, i) == k) sides[ns++] = lf;
      }
      for (int sp = 0; sp < nc; ++sp) {
        const int c = cidx[sp];
        uint16_t w = 0;
        for (int j = 0; j < ns; ++j) {
          uint8_t tg = ftag[t.cf[(size_t)c * t.fpc + sides[j]]];
          if (m.dim == 2 && tg) tg = (uint8_t)((tg & 3) | 4);
          w |= (uint16_t)(tg << (4 * j));
        }
        cw[sp] = w;
      }
      if ((rc = st.put(m.cwtag + (size_t)k * m.cap_c + c0, cw, nc))) return rc;
    }
    uint8_t *vbnd = st.slot<uint8_t>(nv);
    if (!vbnd) return FSALE_ERR_ARG;
    int nb = 0;
    for (int sp = 0; sp < nv; ++sp) {
      vbnd[sp] = vb[vidx[sp]];
      nb += vbnd[sp];
    }
    if ((rc = st.put(m.vbnd + v0, vbnd, nv))) return rc;
    // the boundary vertices (storage positions) for the symmetry-solve
    // launch, grouped by their number of symmetry tags (the size of their QR
    // system), ascending within a group: a wave then takes one solve path
    int32_t *bvl = st.slot<int32_t>(nb > 0 ? nb : 1);
    if (!bvl) return FSALE_ERR_ARG;
    nb = 0;
    int nb2 = 0;
    for (int ns = 0; ns <= 3; ++ns) {
      if (ns == 2) nb2 = nb;
      for (int sp = 0; sp < nv; ++sp) {
        const int vt = vt_popcount(vtags[vidx[sp]]);
        if (vbnd[sp] && vt == ns) bvl[nb++] = sp + v0;
      }
    }
    if ((rc = st.put(m.bv + v0, bvl, nb))) return rc;
    ctx->probs[prob].nb = nb;
    ctx->probs[prob].nb2 = nb2;
  }
  // the deferred rows read this function's locals (P, the row lambdas): fill
  // and send them before returning
  return st.run_deferred();
}

// A problem's parameters and fresh scalars (time t, time step dt), staged.
int stage_scalars(fsale_ctx *ctx, int prob, Stager &st, double t, double dt) {
  DevScalars h{};
  h.t = t;
  h.dt = dt;
  h.active = 1;
  h.err_cell = INT64_MAX;
  ctx->h_s[prob] = h;
  int rc = st.copy(ctx->d_params + prob, &ctx->probs[prob].params, 1);
  return rc ? rc : st.copy(ctx->s + prob, &ctx->h_s[prob], 1);
}

// Initial density, velocity (interleaved [nc*dim]), pressure and, for the
// Lagrangian solver, gamma of one problem, staged.
int stage_state(fsale_ctx *ctx, int prob, Stager &st, const double *d, const double *v, const double *p,
                const double *gamma) {
  const Problem &P = ctx->probs[prob];
  const int N = ctx->m.dim, nc = P.nc;
  const int32_t *cidx = P.topo->cidx.data();
  const bool maire = ctx->solver == FSALE_SOLVER_MAIRE;
  double *dd = maire ? ctx->ma.d : ctx->eu.d;
  double *pp = maire ? ctx->ma.p : ctx->eu.p;
  // one row per component, in storage order
  auto row = [&](double *dst, const double *src, int stride, int comp) -> int {
    double *r = st.slot<double>(nc);
    if (!r) return FSALE_ERR_ARG;
    for (int sp = 0; sp < nc; ++sp) r[sp] = src[(size_t)cidx[sp] * stride + comp];
    return st.put(dst + P.c0, r, nc);
  };
  int rc;
  if ((rc = row(dd, d, 1, 0)) || (rc = row(pp, p, 1, 0))) return rc;
  for (int k = 0; k < N; ++k)
    if ((rc = row(maire ? ctx->ma.v[k] : ctx->eu.v[k], v, N, k))) return rc;
  if (maire && (rc = row(ctx->ma.g, gamma, 1, 0))) return rc;
  return FSALE_OK;
}

// host <-> device field conversions for one problem
struct FieldView {
  double *ptrs[9] = {};   // component arrays on the device (batch ids)
  int ncomp = 0;          // interleaved components on the host
  int n = 0;              // entities of the problem
  int first = 0;          // first batch id
  const int32_t *pos = nullptr;   // storage position of host entity i (cell for corner fields)
  bool corner = false;    // host corner id c*4+k, device [k][cap_c] + c0 + pos[c]
  bool writable = true;
};

int field_view(fsale_ctx *ctx, int prob, int field, FieldView &fv) {
  DevMesh &m = ctx->m;
  const Problem &P = ctx->probs[prob];
  const HostTopology &t = *P.topo;
  const int N = m.dim;
  // kind: 'c' cells, 'v' vertices, 'f' faces (their storage positions)
  auto order = [&](char kind) {
    return kind == 'c' ? t.cpos.data() : kind == 'v' ? t.vpos.data() : t.fpos.data();
  };
  auto one = [&](double *p, int n, int first, char kind) {
    fv.ptrs[0] = p;
    fv.ncomp = 1;
    fv.n = n;
    fv.first = first;
    fv.pos = order(kind);
  };
  auto vec = [&](double *const *p, int ncomp, int n, int first, char kind) {
    for (int i = 0; i < ncomp; ++i) fv.ptrs[i] = p[i];
    fv.ncomp = ncomp;
    fv.n = n;
    fv.first = first;
    fv.pos = order(kind);
  };
  if (ctx->solver == FSALE_SOLVER_EULER) {
    EulerFields &u = ctx->eu;
    switch (field) {
      case FSALE_F_DENSITY: one(u.d, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_VELOCITY: vec(u.v, N, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_INTERNAL_ENERGY: one(u.e, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_PRESSURE: one(u.p, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_TEMPERATURE: one(u.T, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_SOUND_SPEED: one(u.a, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_GRAD_DENSITY: vec(u.gd, N, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_GRAD_VELOCITY: vec(u.gv, N * N, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_GRAD_ENERGY: vec(u.ge, N, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_GRAD_PRESSURE: vec(u.gp, N, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_GRAD_SOUND_SPEED: vec(u.ga, N, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_FLUX:
        if (!u.F[0]) break;
        vec(u.F, N + 2, P.nf, P.f0, 'f');
        return FSALE_OK;
      default: break;
    }
  } else {
    MaireFields &u = ctx->ma;
    switch (field) {
      case FSALE_F_DENSITY: one(u.d, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_VELOCITY: vec(u.v, N, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_INTERNAL_ENERGY: one(u.e, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_PRESSURE: one(u.p, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_TEMPERATURE: one(u.T, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_SOUND_SPEED: one(u.a, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_MASS: one(u.M, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_VOLUME: one(u.V, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_GAMMA: one(u.g, P.nc, P.c0, 'c'); return FSALE_OK;
      case FSALE_F_NODE_VELOCITY: vec(u.un, N, P.nv, P.v0, 'v'); return FSALE_OK;
      case FSALE_F_CORNER_NORMAL:
        vec(u.npc, N, P.nc * m.vpc, P.c0, 'c');
        fv.corner = true;
        return FSALE_OK;
      case FSALE_F_CORNER_FORCE:
        vec(u.Fpc, N, P.nc * m.vpc, P.c0, 'c');
        fv.corner = true;
        return FSALE_OK;
      case FSALE_F_RESIDUAL: vec(u.dudt, N + 2, P.nc, P.c0, 'c'); return FSALE_OK;
      default: break;
    }
  }
  switch (field) {
    case FSALE_F_CELL_CENTROID: vec(m.ccen, N, P.nc, P.c0, 'c'); fv.writable = false; return FSALE_OK;
    case FSALE_F_CELL_VOLUME: one(m.cvol, P.nc, P.c0, 'c'); fv.writable = false; return FSALE_OK;
    case FSALE_F_CELL_MIN_LENGTH: one(m.cminlen, P.nc, P.c0, 'c'); fv.writable = false; return FSALE_OK;
    case FSALE_F_FACE_CENTROID:
    case FSALE_F_FACE_NORMAL:
    case FSALE_F_FACE_AREA:
      if (!P.faces_on_device) break;
      if (field == FSALE_F_FACE_CENTROID) vec(m.fcen, N, P.nf, P.f0, 'f');
      if (field == FSALE_F_FACE_NORMAL) vec(m.fnrm, N, P.nf, P.f0, 'f');
      if (field == FSALE_F_FACE_AREA) one(m.farea, P.nf, P.f0, 'f');
      fv.writable = false;
      return FSALE_OK;
    case FSALE_F_COORDS: vec(m.x, N, P.nv, P.v0, 'v'); return FSALE_OK;
    default: break;
  }
  set_message("field not available for this solver");
  return FSALE_ERR_ARG;
}

bool valid_prob(const fsale_ctx *ctx, int prob) {
  return ctx && prob >= 0 && prob < ctx->max_problems && ctx->probs[prob].live;
}

}  // namespace

int get_problem_field(fsale_ctx *ctx, int prob, int field, double *host_out) {
  if (!valid_prob(ctx, prob) || !host_out) return FSALE_ERR_ARG;
  FieldView fv;
  int rc = field_view(ctx, prob, field, fv);
  if (rc) return rc;
  const DevMesh &m = ctx->m;
  std::vector<double> buf;
  for (int k = 0; k < fv.ncomp; ++k) {
    if (fv.corner) {
      // corner fields: device [kk][cap_c] rows, host c*vpc + kk
      const int nc = fv.n / m.vpc;
      buf.resize(nc);
      for (int kk = 0; kk < m.vpc; ++kk) {
        FSALE_HIP(hipMemcpyAsync(buf.data(), fv.ptrs[k] + (size_t)kk * m.cap_c + fv.first, nc * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
        FSALE_HIP(hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < nc; ++c) host_out[((size_t)c * m.vpc + kk) * fv.ncomp + k] = buf[fv.pos[c]];
      }
    } else {
      buf.resize(fv.n);
      FSALE_HIP(hipMemcpyAsync(buf.data(), fv.ptrs[k] + fv.first, fv.n * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
      FSALE_HIP(hipStreamSynchronize(ctx->stream));
      for (int i = 0; i < fv.n; ++i) host_out[(size_t)i * fv.ncomp + k] = buf[fv.pos[i]];
    }
  }
  return check_device_error(ctx, prob);
}

// Several fields of one problem with one synchronisation: the arrays and the
// problem's scalars go through one pinned buffer.
int get_problem_fields(fsale_ctx *ctx, int prob, int n, const int32_t *fields, double *const *outs) {
  if (!valid_prob(ctx, prob) || n <= 0 || !fields || !outs) return FSALE_ERR_ARG;
  std::vector<FieldView> views(n);
  size_t total = 0;
  for (int i = 0; i < n; ++i) {
    if (!outs[i]) return FSALE_ERR_ARG;
    int rc = field_view(ctx, prob, fields[i], views[i]);
    if (rc) return rc;
    if (views[i].corner) {   // corner fields: one at a time
      for (int j = 0; j < n; ++j)
        if ((rc = get_problem_field(ctx, prob, fields[j], outs[j]))) return rc;
      return FSALE_OK;
    }
    total += (size_t)views[i].ncomp * views[i].n;
  }
  const size_t bytes = total * sizeof(double) + sizeof(DevScalars) + 64;
  if (bytes > ctx->dl_bytes) {
    FSALE_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->dl) (void)hipHostFree(ctx->dl);
    ctx->dl = nullptr;
    ctx->dl_bytes = 0;
    FSALE_HIP(hipHostMalloc(&ctx->dl, bytes, hipHostMallocDefault));
    ctx->dl_bytes = bytes;
  }
  double *buf = static_cast<double *>(ctx->dl);
  size_t off = 0;
  for (const FieldView &fv : views)
    for (int k = 0; k < fv.ncomp; ++k) {
      FSALE_HIP(hipMemcpyAsync(buf + off, fv.ptrs[k] + fv.first, fv.n * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
      off += fv.n;
    }
  DevScalars *sc = reinterpret_cast<DevScalars *>(buf + off);
  FSALE_HIP(hipMemcpyAsync(sc, ctx->s + prob, sizeof(DevScalars), hipMemcpyDeviceToHost, ctx->stream));
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  ctx->h_s[prob] = *sc;
  off = 0;
  for (int i = 0; i < n; ++i) {
    const FieldView &fv = views[i];
    for (int k = 0; k < fv.ncomp; ++k) {
      for (int e = 0; e < fv.n; ++e) outs[i][(size_t)e * fv.ncomp + k] = buf[off + fv.pos[e]];
      off += fv.n;
    }
  }
  if (sc->err_code) return check_device_error(ctx, prob);
  return FSALE_OK;
}

__global__ void k_zero_params(ProbParams *p) {
  constexpr int words = (int)(sizeof(ProbParams) / sizeof(int));
  static_assert(sizeof(ProbParams) % sizeof(int) == 0, "ProbParams is a whole number of words");
  for (int i = threadIdx.x; i < words; i += blockDim.x) reinterpret_cast<int *>(p)[i] = 0;
}

int stage_field_downloads(fsale_ctx *ctx, const int *probs, int nprobs, int n, const int32_t *fields,
                          double *const *const *outs, std::vector<PendingFields> &pend) {
  if (n <= 0 || n > 8 || nprobs < 0) return FSALE_ERR_ARG;
  pend.clear();
  pend.resize(nprobs);
  size_t total = 0;
  for (int q = 0; q < nprobs; ++q) {
    const int prob = probs[q];
    if (!valid_prob(ctx, prob)) return FSALE_ERR_ARG;
    PendingFields &pf = pend[q];
    pf.topo = ctx->probs[prob].topo;
    pf.n = n;
    pf.off = total;
    for (int i = 0; i < n; ++i) {
      FieldView fv;
      int rc = field_view(ctx, prob, fields[i], fv);
      if (rc) return rc;
      if (fv.corner || !outs[q][i]) return FSALE_ERR_ARG;
      pf.ncomp[i] = fv.ncomp;
      pf.count[i] = fv.n;
      pf.pos[i] = fv.pos;
      pf.outs[i] = outs[q][i];
      total += (size_t)fv.ncomp * fv.n;
    }
  }
  const size_t bytes = total * sizeof(double) + 64;
  if (bytes > ctx->dl_bytes) {
    FSALE_HIP(hipStreamSynchronize(ctx->upload));
    if (ctx->dl) (void)hipHostFree(ctx->dl);
    ctx->dl = nullptr;
    ctx->dl_bytes = 0;
    FSALE_HIP(hipHostMalloc(&ctx->dl, bytes + bytes / 2, hipHostMallocDefault));
    ctx->dl_bytes = bytes + bytes / 2;
  }
  if (!ctx->ev_dl) FSALE_HIP(hipEventCreateWithFlags(&ctx->ev_dl, hipEventDisableTiming));
  FSALE_HIP(hipEventRecord(ctx->ev_dl, ctx->stream));
  FSALE_HIP(hipStreamWaitEvent(ctx->upload, ctx->ev_dl, 0));
  double *buf = static_cast<double *>(ctx->dl);
  for (int q = 0; q < nprobs; ++q) {
    size_t off = pend[q].off;
    for (int i = 0; i < n; ++i) {
      FieldView fv;
      (void)field_view(ctx, probs[q], fields[i], fv);
      for (int k = 0; k < fv.ncomp; ++k) {
        FSALE_HIP(hipMemcpyAsync(buf + off, fv.ptrs[k] + fv.first, fv.n * sizeof(double), hipMemcpyDeviceToHost,
                                 ctx->upload));
        off += fv.n;
      }
    }
  }
  return FSALE_OK;
}

int finish_field_downloads(fsale_ctx *ctx, std::vector<PendingFields> &pend) {
  if (pend.empty()) return FSALE_OK;
  FSALE_HIP(hipStreamSynchronize(ctx->upload));
  const double *buf = static_cast<const double *>(ctx->dl);
  for (PendingFields &pf : pend) {
    size_t off = pf.off;
    for (int i = 0; i < pf.n; ++i)
      for (int k = 0; k < pf.ncomp[i]; ++k) {
        double *out = pf.outs[i];
        const int32_t *pos = pf.pos[i];
        const int nc = pf.ncomp[i];
        for (int e = 0; e < pf.count[i]; ++e) out[(size_t)e * nc + k] = buf[off + pos[e]];
        off += pf.count[i];
      }
    pf.topo.reset();
  }
  pend.clear();
  return FSALE_OK;
}

int retire_deferred(fsale_ctx *ctx, int prob) {
  if (!ctx || ctx->single || !valid_prob(ctx, prob)) return FSALE_ERR_ARG;
  Problem &P = ctx->probs[prob];
  ctx->ra_c.release(P.c0, P.nc);
  ctx->ra_f.release(P.f0, P.nf);
  ctx->ra_v.release(P.v0, P.nv);
  ctx->ra_n.release(P.n0, P.nn);
  ctx->ra_ct.release(P.ct0, P.nct);
  ctx->ra_vt.release(P.vt0, P.nvt);
  P = Problem{};
  P.quarantined = true;
  ctx->tiles_dirty = true;
  // after the round in flight (stream order, no host wait): the per-problem
  // kernels skip a slot that owns no cells
  hipLaunchKernelGGL(k_zero_params, dim3(1), dim3(64), 0, ctx->stream, ctx->d_params + prob);
  FSALE_HIP(hipGetLastError());
  return FSALE_OK;
}

int set_problem_field(fsale_ctx *ctx, int prob, int field, const double *host_in) {
  if (!valid_prob(ctx, prob) || !host_in) return FSALE_ERR_ARG;
  FieldView fv;
  int rc = field_view(ctx, prob, field, fv);
  if (rc) return rc;
  if (!fv.writable) {
    set_message("geometry fields are derived from the coordinates");
    return FSALE_ERR_ARG;
  }
  const DevMesh &m = ctx->m;
  std::vector<double> buf;
  for (int k = 0; k < fv.ncomp; ++k) {
    if (fv.corner) {
      const int nc = fv.n / m.vpc;
      buf.resize(nc);
      for (int kk = 0; kk < m.vpc; ++kk) {
        for (int c = 0; c < nc; ++c) buf[fv.pos[c]] = host_in[((size_t)c * m.vpc + kk) * fv.ncomp + k];
        FSALE_HIP(hipMemcpyAsync(fv.ptrs[k] + (size_t)kk * m.cap_c + fv.first, buf.data(), nc * sizeof(double),
                                 hipMemcpyHostToDevice, ctx->stream));
        FSALE_HIP(hipStreamSynchronize(ctx->stream));
      }
    } else {
      buf.resize(fv.n);
      for (int i = 0; i < fv.n; ++i) buf[fv.pos[i]] = host_in[(size_t)i * fv.ncomp + k];
      FSALE_HIP(hipMemcpyAsync(fv.ptrs[k] + fv.first, buf.data(), fv.n * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
      FSALE_HIP(hipStreamSynchronize(ctx->stream));
    }
  }
  ctx->probs[prob].partials_valid = false;
  if (field == FSALE_F_COORDS &&
      (rc = launch_geometry(ctx, prob, true, ctx->probs[prob].faces_on_device, ctx->stream)))
    return rc;
  return maire_pack_problem(ctx, prob, ctx->stream);
}

// d, v, p of one problem from host arrays (v interleaved)
int set_problem_state(fsale_ctx *ctx, int prob, const double *d, const double *v, const double *p) {
  int rc;
  if ((rc = set_problem_field(ctx, prob, FSALE_F_DENSITY, d))) return rc;
  if ((rc = set_problem_field(ctx, prob, FSALE_F_VELOCITY, v))) return rc;
  return set_problem_field(ctx, prob, FSALE_F_PRESSURE, p);
}


namespace {
int run_steps(fsale_ctx *ctx, int nsteps, int use_graph);
}

// Driver steps of every live problem; chunks of up to 10 steps are replayed
// from a captured graph (the graph reads tiles, parameters and scalars from
// device memory, so it stays valid while problems come and go).  Eulerian
// steps work on packed copies of the state (EulerFields::gs), made before
// the steps and unpacked after them.
__global__ void k_task_open(DevScalars *s) {
  if (threadIdx.x == 0 && s->err_token == 0) s->active = 1;
}

int open_tasks(fsale_ctx *ctx) {
  if (ctx->tasks_open) return FSALE_OK;
  hipLaunchKernelGGL(k_task_open, dim3(1), dim3(64), 0, ctx->stream, ctx->s);
  FSALE_HIP(hipGetLastError());
  ctx->tasks_open = true;
  return FSALE_OK;
}

int batch_steps(fsale_ctx *ctx, int nsteps, int use_graph) {
  int rc;
  if ((rc = sync_tiles(ctx))) return rc;
  ctx->tasks_open = false;
  if (ctx->h_ctiles.empty() || nsteps <= 0) return FSALE_OK;
  if (ctx->solver == FSALE_SOLVER_EULER) {
    bool need = false;
    for (auto &P : ctx->probs)
      if (P.live && !P.partials_valid) need = true;
    if (need) {
      if ((rc = euler_dt_partials(ctx))) return rc;
      for (auto &P : ctx->probs)
        if (P.live) P.partials_valid = true;
    }
  }
  const bool euler = ctx->solver == FSALE_SOLVER_EULER;
  if (euler && (rc = euler_pack(ctx))) return rc;
  if ((rc = run_steps(ctx, nsteps, use_graph))) return rc;
  return euler ? euler_unpack(ctx) : FSALE_OK;
}

namespace {

// the steps of batch_steps, eager or from the captured graph
int run_steps(fsale_ctx *ctx, int nsteps, int use_graph) {
  int rc;
  // first / last: the step opens / closes a run of consecutive steps (the
  // Lagrangian step defers its close to the next step of the run)
  auto step = [&](bool first, bool last) {
    return ctx->solver == FSALE_SOLVER_EULER ? euler_step(ctx) : maire_step(ctx, first, last);
  };
  if (!use_graph || nsteps < 2 || ctx->profiling) {
    for (int n = 0; n < nsteps; ++n)
      if ((rc = step(n == 0, n == nsteps - 1))) return rc;
    return FSALE_OK;
  }
  const int chunk = nsteps < 10 ? nsteps : 10;
  std::vector<double> key = {(double)ctx->solver, (double)chunk};
  for (int k = 0; k < K_COUNT; ++k) key.push_back((double)persistent_grid(ctx, k));
  key.push_back((double)ctx->pgrid);
  key.push_back((double)ctx->n_bbtiles);
  key.push_back((double)ctx->lds + 2.0 * ctx->lds_all + 4.0 * ctx->s3 + 8.0 * ctx->s3_all + 16.0 * ctx->fp3 +
                32.0 * ctx->brick_all);
  GraphCache &gc = ctx->graph;
  GraphEntry *g = nullptr;
  for (auto &e : gc.entries)
    if (e.key == key) g = &e;
  if (!g) {
    if (gc.entries.size() >= GraphCache::kMaxEntries) {
      // evict the least recently used, once its last launch has completed
      auto lru = std::min_element(gc.entries.begin(), gc.entries.end(),
                                  [](const GraphEntry &a, const GraphEntry &b) { return a.used < b.used; });
      FSALE_HIP(hipEventSynchronize(lru->done));
      (void)hipGraphExecDestroy(lru->exec);
      (void)hipEventDestroy(lru->done);
      gc.entries.erase(lru);
    }
    GraphEntry e;
    e.key = key;
    FSALE_HIP(hipEventCreateWithFlags(&e.done, hipEventDisableTiming));
    hipGraph_t graph;
    hipError_t err = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (err != hipSuccess) {
      (void)hipEventDestroy(e.done);
      set_message(std::string("hipStreamBeginCapture: ") + hipGetErrorString(err));
      return FSALE_ERR_HIP;
    }
    for (int n = 0; n < chunk; ++n) {
      rc = step(n == 0, n == chunk - 1);
      if (rc) {
        (void)hipStreamEndCapture(ctx->stream, &graph);
        (void)hipEventDestroy(e.done);
        return rc;
      }
    }
    err = hipStreamEndCapture(ctx->stream, &graph);
    if (err == hipSuccess) {
      err = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
    }
    if (err != hipSuccess) {
      (void)hipEventDestroy(e.done);
      set_message(std::string("graph capture: ") + hipGetErrorString(err));
      return FSALE_ERR_HIP;
    }
    gc.entries.push_back(e);
    g = &gc.entries.back();
  }
  g->used = ++gc.clock;
  int n = 0;
  for (; n + chunk <= nsteps; n += chunk) FSALE_HIP(hipGraphLaunch(g->exec, ctx->stream));
  FSALE_HIP(hipEventRecord(g->done, ctx->stream));
  for (const int n0 = n; n < nsteps; ++n)
    if ((rc = step(n == n0, n == nsteps - 1))) return rc;
  return FSALE_OK;
}

// Device arrays that carry a batch from one step to the next (everything
// else a step writes is rewritten before it is read): (pointer, bytes).
void step_state(fsale_ctx *ctx, std::vector<std::pair<void *, size_t>> &out) {
  const DevMesh &m = ctx->m;
  const size_t C = (size_t)m.cap_c * sizeof(double), V = (size_t)m.cap_v * sizeof(double);
  if (ctx->solver == FSALE_SOLVER_EULER) {
    EulerFields &u = ctx->eu;
    for (double *p : {u.d, u.e, u.p, u.T, u.a}) out.push_back({p, C});
    for (int d = 0; d < m.dim; ++d) out.push_back({u.v[d], C});
    out.push_back({u.partials, sizeof(double) * kWaves * ((size_t)ctx->ra_ct.capacity + 1)});
  } else {
    MaireFields &u = ctx->ma;
    for (double *p : {u.V, u.p, u.d, u.e, u.T, u.a, m.cvol, m.cminlen}) out.push_back({p, C});
    for (int d = 0; d < m.dim; ++d) {
      out.push_back({u.v[d], C});
      out.push_back({m.ccen[d], C});
      out.push_back({m.x[d], V});
    }
  }
  out.push_back({ctx->s, sizeof(DevScalars) * ctx->max_problems});
}

// The packed copies the step kernels read, rebuilt from the arrays a
// snapshot restores: the Eulerian state records, the Lagrangian nodal
// solve's cell state and coordinates.
int refresh_packed(fsale_ctx *ctx) {
  if (ctx->solver == FSALE_SOLVER_EULER) return euler_pack(ctx);
  int rc;
  for (int p = 0; p < ctx->max_problems; ++p)
    if (ctx->probs[p].live && (rc = maire_pack_problem(ctx, p, ctx->stream))) return rc;
  return FSALE_OK;
}

struct Snapshot {
  std::vector<std::pair<void *, size_t>> live;
  std::vector<void *> copy;
  ~Snapshot() {
    for (void *p : copy)
      if (p) (void)hipFree(p);
  }
  int take(fsale_ctx *ctx) {
    step_state(ctx, live);
    copy.assign(live.size(), nullptr);
    for (size_t i = 0; i < live.size(); ++i) {
      FSALE_HIP(hipMalloc(&copy[i], live[i].second));
      FSALE_HIP(hipMemcpyAsync(copy[i], live[i].first, live[i].second, hipMemcpyDeviceToDevice, ctx->stream));
    }
    return FSALE_OK;
  }
  int restore(fsale_ctx *ctx) {
    for (size_t i = 0; i < live.size(); ++i)
      FSALE_HIP(hipMemcpyAsync(live[i].first, copy[i], live[i].second, hipMemcpyDeviceToDevice, ctx->stream));
    return FSALE_OK;
  }
};

}  // namespace

// Launch-shape autotuner: every candidate shape runs `nsteps` real eager
// steps of the resident batch from the same snapshot, each launch timed
// with HIP events; each kernel keeps its fastest shape.  The batch's state,
// step budgets and profiling counters are left as they were.
// A candidate shape replaces the best so far only when it is faster by this
// fraction: per-launch timings of a few steps vary by a few percent, and one
// block per tile (the first candidate) keeps each XCD's tiles contiguous.
constexpr double kTuneMargin = 0.02;

int autotune(fsale_ctx *ctx, int nsteps, double *best_ms) {
  int rc;
  if ((rc = sync_tiles(ctx))) return rc;
  if (ctx->h_ctiles.empty()) return FSALE_OK;
  if (ctx->solver == FSALE_SOLVER_EULER) {
    if ((rc = euler_dt_partials(ctx))) return rc;
    for (auto &P : ctx->probs)
      if (P.live) P.partials_valid = true;
  }
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  Snapshot snap;
  if ((rc = snap.take(ctx))) return rc;
  // no problem runs out of step budget while the candidates are timed
  for (int p = 0; p < ctx->max_problems; ++p)
    if (ctx->probs[p].live) {
      ProbParams pp = ctx->probs[p].params;
      pp.max_steps = INT64_MAX;
      FSALE_HIP(hipMemcpyAsync(ctx->d_params + p, &pp, sizeof(pp), hipMemcpyHostToDevice, ctx->stream));
      FSALE_HIP(hipStreamSynchronize(ctx->stream));
    }
  ProfSlot saved[K_COUNT];
  for (int k = 0; k < K_COUNT; ++k) std::swap(saved[k], ctx->prof[k]);
  const bool was_profiling = ctx->profiling;
  ctx->profiling = true;
  const int cands[] = {0, 1, 2, 3, 4, 6, 8, 12, 16};
  double best[K_COUNT];
  long long per_step[K_COUNT] = {};
  LaunchShape choice = ctx->shape;
  for (int k = 0; k < K_COUNT; ++k) best[k] = 1e300;
  rc = FSALE_OK;
  for (int c : cands) {
    for (int k = 0; k < K_COUNT; ++k) ctx->shape.blocks_per_cu[k] = c;
    if ((rc = snap.restore(ctx))) break;
    if ((rc = refresh_packed(ctx))) break;
    for (int n = 0; n < nsteps && !rc; ++n)
      rc = ctx->solver == FSALE_SOLVER_EULER ? euler_step(ctx) : maire_step(ctx, n == 0, n == nsteps - 1);
    if (rc) break;
    // the two nodal kernels are one launch at one block per tile: they are
    // chosen together, by their summed time per step
    double pair_ms = 0;
    int64_t pair_nl = 0;
    for (int k = 0; k < K_COUNT; ++k) {
      double tot = 0;
      int64_t nl = 0;
      if ((rc = fsale_get_profile(ctx, k, &tot, &nl))) break;
      if (k == K_MAIRE_NODAL || k == K_MAIRE_NODAL_BND) {
        pair_ms += tot;
        if (k == K_MAIRE_NODAL) pair_nl = nl;
      } else if (nl > 0 && tot / nl < best[k] * (c ? 1.0 - kTuneMargin : 1.0)) {
        best[k] = tot / nl;
        per_step[k] = nl / nsteps;
        choice.blocks_per_cu[k] = c;
      }
      ctx->prof[k].total_ms = 0;
      ctx->prof[k].launches = 0;
    }
    if (rc) break;
    if (pair_nl > 0 && pair_ms / pair_nl < best[K_MAIRE_NODAL] * (c ? 1.0 - kTuneMargin : 1.0)) {
      best[K_MAIRE_NODAL] = pair_ms / pair_nl;   // per nodal evaluation, both kinds
      per_step[K_MAIRE_NODAL] = pair_nl / nsteps;
      choice.blocks_per_cu[K_MAIRE_NODAL] = c;
      choice.blocks_per_cu[K_MAIRE_NODAL_BND] = c;
    }
  }
  ctx->profiling = was_profiling;
  for (int k = 0; k < K_COUNT; ++k) std::swap(saved[k], ctx->prof[k]);
  int rc2 = snap.restore(ctx);
  if (rc2 == FSALE_OK) rc2 = refresh_packed(ctx);
  for (int p = 0; p < ctx->max_problems && !rc2; ++p)
    if (ctx->probs[p].live) rc2 = upload_params(ctx, p);
  if (rc2 == FSALE_OK) rc2 = read_scalars(ctx);
  if (rc) return rc;
  if (rc2) return rc2;
  ctx->shape = choice;
  if (best_ms) {
    double t = 0;
    for (int k = 0; k < K_COUNT; ++k)
      if (per_step[k]) t += best[k] * per_step[k];
    *best_ms = t;
  }
  return FSALE_OK;
}

}  // namespace fsale

using namespace fsale;

extern "C" {

const char *fsale_status_string(int status) {
  switch (status) {
    case FSALE_OK: return "ok";
    case FSALE_ERR_ARG: return "bad argument";
    case FSALE_ERR_MESH: return "unsupported or invalid mesh";
    case FSALE_ERR_HIP: return "HIP runtime error";
    case FSALE_ERR_OOM: return "out of device memory or batch capacity";
    case FSALE_ERR_NEG_STATE: return "negative density or internal energy";
    case FSALE_ERR_DT: return "infinite delta t";
    case FSALE_ERR_STATE: return "call out of order";
    case FSALE_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

const char *fsale_last_message(void) { return g_message.c_str(); }

int fsale_qr_solve(int32_t n, const double *A, const double *b, double *x) {
  if (!A || !b || !x || n < 1 || n > 4) return FSALE_ERR_ARG;
  auto run = [&](auto tag) {
    constexpr int NR = decltype(tag)::value;
    double M[NR][NR], B[NR];
    for (int r = 0; r < NR; ++r) {
      B[r] = b[r];
      for (int c = 0; c < NR; ++c) M[r][c] = A[r * NR + c];
    }
    qr_solve_fixed<NR>(M, B);
    for (int r = 0; r < NR; ++r) x[r] = B[r];
  };
  switch (n) {
    case 1: run(std::integral_constant<int, 1>()); break;
    case 2: run(std::integral_constant<int, 2>()); break;
    case 3: run(std::integral_constant<int, 3>()); break;
    default: run(std::integral_constant<int, 4>()); break;
  }
  return FSALE_OK;
}

int fsale_batch_create(int solver, int dim, int verts_per_cell, int32_t max_problems, int64_t cell_capacity,
                       int64_t face_capacity, int64_t vertex_capacity, int64_t neighbour_capacity,
                       fsale_ctx **out) {
  if (!out || (solver != FSALE_SOLVER_EULER && solver != FSALE_SOLVER_MAIRE)) return FSALE_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_message("no HIP device visible");
    return FSALE_ERR_NO_DEVICE;
  }
  fsale_ctx *ctx = new fsale_ctx();
  (void)hipGetDevice(&ctx->device);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->upload, hipStreamNonBlocking) != hipSuccess) {
    fsale_ctx_destroy(ctx);
    set_message("hipStreamCreate failed");
    return FSALE_ERR_HIP;
  }
  int rc = alloc_batch(ctx, solver, dim, verts_per_cell, max_problems, cell_capacity, face_capacity,
                       vertex_capacity, neighbour_capacity);
  if (rc) {
    fsale_ctx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return FSALE_OK;
}

int fsale_ctx_create(const fsale_mesh_desc *mesh, int solver, fsale_ctx **out) {
  if (!mesh || !out) return FSALE_ERR_ARG;
  *out = nullptr;
  // exact capacity from the mesh's own connectivity
  std::string err;
  int rc;
  std::shared_ptr<const HostTopology> t = cached_topology(*mesh, rc, err);
  if (rc) {
    set_message(err);
    return rc;
  }
  fsale_ctx *ctx = nullptr;
  if ((rc = fsale_batch_create(solver, mesh->dim, mesh->verts_per_cell, 1, t->nc, t->nf, t->nv,
                               (int64_t)t->nbr.size(), &ctx)))
    return rc;
  ctx->single = true;
  if (solver == FSALE_SOLVER_EULER) {
    // the task-level evaluate_fluxes keeps a face flux array
    for (int i = 0; i < mesh->dim + 2; ++i)
      if ((rc = dev_alloc_n(ctx, &ctx->eu.F[i], t->nf))) {
        fsale_ctx_destroy(ctx);
        return rc;
      }
  }
  int prob;
  Stager st(ctx);
  if ((rc = place_mesh(ctx, *mesh, &prob, st)) || (rc = stage_scalars(ctx, prob, st, 0.0, 0.0)) ||
      (rc = st.finish())) {
    st.discard();
    (void)hipStreamSynchronize(ctx->upload);
    fsale_ctx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return FSALE_OK;
}

int fsale_ctx_destroy(fsale_ctx *ctx) {
  if (!ctx) return FSALE_ERR_ARG;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->upload) (void)hipStreamSynchronize(ctx->upload);
  for (auto &p : ctx->prof) {
    for (hipEvent_t e : p.start) (void)hipEventDestroy(e);
    for (hipEvent_t e : p.stop) (void)hipEventDestroy(e);
  }
  for (auto &g : ctx->graph.entries) {   // the stream is idle: no launch in flight
    (void)hipGraphExecDestroy(g.exec);
    (void)hipEventDestroy(g.done);
  }
  for (void *p : ctx->allocations) (void)hipFree(p);
  if (ctx->h_s) (void)hipHostFree(ctx->h_s);
  if (ctx->stage) (void)hipHostFree(ctx->stage);
  if (ctx->dl) (void)hipHostFree(ctx->dl);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->upload) (void)hipStreamDestroy(ctx->upload);
  if (ctx->ev_dl) (void)hipEventDestroy(ctx->ev_dl);
  delete ctx;
  return FSALE_OK;
}

int fsale_dump(fsale_ctx *ctx, const char *path) {
  if (!ctx || !ctx->single || ctx->probs.empty() || !ctx->probs[0].live || !path) return FSALE_ERR_ARG;
  const Problem &P = ctx->probs[0];
  if (!P.have_state) return FSALE_ERR_STATE;
  const int N = ctx->m.dim, nc = P.nc, nv = P.nv;
  std::vector<double> cen((size_t)nc * N), d(nc), e(nc), p(nc), v((size_t)nc * N), x((size_t)nv * N);
  const int32_t ids[6] = {FSALE_F_CELL_CENTROID, FSALE_F_DENSITY, FSALE_F_INTERNAL_ENERGY, FSALE_F_PRESSURE,
                          FSALE_F_VELOCITY, FSALE_F_COORDS};
  double *outs[6] = {cen.data(), d.data(), e.data(), p.data(), v.data(), x.data()};
  int rc = get_problem_fields(ctx, 0, 6, ids, outs);
  if (rc) return rc;
  if ((rc = read_scalars(ctx))) return rc;
  FILE *f = std::fopen(path, "w");
  if (!f) {
    set_message(std::string("cannot write ") + path);
    return FSALE_ERR_ARG;
  }
  std::fprintf(f, "# Solution time: %.14e\n# Number iterations: %lld\n", ctx->h_s[0].t,
               (long long)ctx->h_s[0].steps);
  std::fprintf(f, "# BEGIN CELLS\n# Total number: %d\n# local_id global_id ", nc);
  for (int k = 0; k < N; ++k) std::fprintf(f, "centroid(%d) ", k);
  std::fprintf(f, "density internal_energy pressure ");
  for (int k = 0; k < N; ++k) std::fprintf(f, "velocity(%d) ", k);
  std::fprintf(f, "\n");
  for (int c = 0; c < nc; ++c) {
    std::fprintf(f, "%d %d ", c, c);
    for (int k = 0; k < N; ++k) std::fprintf(f, "%.14e ", cen[(size_t)c * N + k]);
    std::fprintf(f, "%.14e %.14e %.14e ", d[c], e[c], p[c]);
    for (int k = 0; k < N; ++k) std::fprintf(f, "%.14e ", v[(size_t)c * N + k]);
    std::fprintf(f, "\n");
  }
  std::fprintf(f, "# END CELLS\n# BEGIN VERTICES\n# Total number: %d\n# local_id global_id ", nv);
  for (int k = 0; k < N; ++k) std::fprintf(f, "coordinate(%d) ", k);
  std::fprintf(f, "\n");
  for (int i = 0; i < nv; ++i) {
    std::fprintf(f, "%d %d ", i, i);
    for (int k = 0; k < N; ++k) std::fprintf(f, "%.14e ", x[(size_t)i * N + k]);
    std::fprintf(f, "\n");
  }
  std::fprintf(f, "# END VERTICES\n");
  const bool ok = std::fclose(f) == 0;
  if (!ok) {
    set_message(std::string("cannot write ") + path);
    return FSALE_ERR_ARG;
  }
  return FSALE_OK;
}

int fsale_output(fsale_ctx *ctx, const char *prefix, const char *postfix, int64_t iteration, int32_t rank) {
  if (!ctx || !ctx->single || ctx->probs.empty() || !ctx->probs[0].live || !prefix || !postfix || iteration < 0 ||
      rank < 0)
    return FSALE_ERR_ARG;
  const Problem &P = ctx->probs[0];
  if (!P.have_state) return FSALE_ERR_STATE;
  const int N = ctx->m.dim, nc = P.nc, nv = P.nv;
  const bool euler = ctx->solver == FSALE_SOLVER_EULER;
  std::vector<double> x((size_t)nv * N), d(nc), p(nc);
  const int32_t ids[3] = {FSALE_F_COORDS, FSALE_F_DENSITY, FSALE_F_PRESSURE};
  double *outs[3] = {x.data(), d.data(), p.data()};
  int rc = get_problem_fields(ctx, 0, euler ? 3 : 2, ids, outs);
  if (rc) return rc;
  if ((rc = read_scalars(ctx))) return rc;
  // apps/hydro/tasks.h:723-725: prefix_rankRRRRRR.IIIIII.postfix;
  // apps/maire_hydro/tasks.h:681-683: prefix_rankRRRRRR_IIIIII.postfix
  char tail[64];
  std::snprintf(tail, sizeof(tail), "_rank%06d%c%06lld.", (int)rank, euler ? '.' : '_', (long long)iteration);
  const std::string path = std::string(prefix) + tail + postfix;
  const char *names[2] = {"density", "pressure"};
  const double *vals[2] = {d.data(), p.data()};
  return exodus_write(path.c_str(), prefix, N, nv, x.data(), nc, ctx->m.vpc, P.topo->cv.data(), ctx->h_s[0].t,
                      euler ? 2 : 1, names, vals);
}

int fsale_sync(fsale_ctx *ctx) {
  if (!ctx) return FSALE_ERR_ARG;
  int rc = read_scalars(ctx);
  if (rc) return rc;
  for (int p = 0; p < ctx->max_problems; ++p)
    if (ctx->probs[p].live && ctx->h_s[p].err_code) return check_device_error(ctx, p);
  return FSALE_OK;
}

int fsale_mesh_counts(const fsale_ctx *ctx, int32_t *counts) {
  if (!valid_prob(ctx, 0) || !counts) return FSALE_ERR_ARG;
  const Problem &P = ctx->probs[0];
  counts[0] = P.nc;
  counts[1] = P.nv;
  counts[2] = P.nf;
  counts[3] = P.nn;
  counts[4] = ctx->m.dim;
  return FSALE_OK;
}

int fsale_get_topology(fsale_ctx *ctx, int32_t *face_cells, int32_t *cell_faces, int32_t *nbr_off,
                       int32_t *nbr) {
  if (!valid_prob(ctx, 0)) return FSALE_ERR_ARG;
  const HostTopology &t = *ctx->probs[0].topo;
  if (face_cells) std::memcpy(face_cells, t.fc.data(), t.fc.size() * sizeof(int32_t));
  if (cell_faces) std::memcpy(cell_faces, t.cf.data(), t.cf.size() * sizeof(int32_t));
  if (nbr_off) std::memcpy(nbr_off, t.nbr_off.data(), t.nbr_off.size() * sizeof(int32_t));
  if (nbr) std::memcpy(nbr, t.nbr.data(), t.nbr.size() * sizeof(int32_t));
  return FSALE_OK;
}

int fsale_mesh_connectivity(const fsale_mesh_desc *mesh, int32_t *counts, int32_t *face_cells,
                            int32_t *cell_faces, int32_t *nbr_off, int32_t *nbr) {
  if (!mesh || !counts) return FSALE_ERR_ARG;
  std::string err;
  int rc;
  std::shared_ptr<const HostTopology> tp = cached_topology(*mesh, rc, err);
  if (rc) {
    set_message(err);
    return rc;
  }
  const HostTopology &t = *tp;
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
  return get_problem_field(ctx, 0, field, host_out);
}

int fsale_set_field(fsale_ctx *ctx, int field, const double *host_in) {
  return set_problem_field(ctx, 0, field, host_in);
}

int fsale_get_time(fsale_ctx *ctx, double *soln_time, double *time_step, int64_t *steps) {
  if (!valid_prob(ctx, 0)) return FSALE_ERR_ARG;
  int rc = read_scalars(ctx);
  if (rc) return rc;
  if (soln_time) *soln_time = ctx->h_s[0].t;
  if (time_step) *time_step = ctx->h_s[0].dt;
  if (steps) *steps = ctx->h_s[0].steps;
  return FSALE_OK;
}

int fsale_set_time(fsale_ctx *ctx, double soln_time, double time_step) {
  if (!valid_prob(ctx, 0)) return FSALE_ERR_ARG;
  int rc = read_scalars(ctx);
  if (rc) return rc;
  ctx->h_s[0].t = soln_time;
  ctx->h_s[0].dt = time_step;
  FSALE_HIP(hipMemcpyAsync(ctx->s, ctx->h_s, sizeof(DevScalars), hipMemcpyHostToDevice, ctx->stream));
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  return FSALE_OK;
}

int fsale_error_cell(fsale_ctx *ctx, int64_t *cell) {
  if (!valid_prob(ctx, 0) || !cell) return FSALE_ERR_ARG;
  int rc = read_scalars(ctx);
  if (rc) return rc;
  *cell = ctx->h_s[0].err_code ? ctx->h_s[0].err_cell : -1;
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

// ----------------------------------------------------------------- batches

int fsale_batch_add_euler(fsale_ctx *ctx, const fsale_mesh_desc *mesh, fsale_eos eos,
                          const fsale_euler_params *prm, const double *d, const double *v, const double *p,
                          int32_t *problem) {
  if (!ctx || ctx->single || ctx->solver != FSALE_SOLVER_EULER || !mesh || !prm || !d || !v || !p || !problem)
    return FSALE_ERR_ARG;
  if (!(eos.gamma > 1) || !(eos.cv > 0) || prm->flux_kind < 0 || (prm->flux_kind & 0xf) > 2 ||
      (prm->flux_kind >> 4) > 1)
    return FSALE_ERR_ARG;
  int prob = -1, rc;
  Stager st(ctx);
  if ((rc = place_mesh(ctx, *mesh, &prob, st))) {
    st.discard();
    (void)st.finish();
    if (prob >= 0 && prob < ctx->max_problems && ctx->probs[prob].live) fsale_batch_remove(ctx, prob);
    return rc;
  }
  Problem &P = ctx->probs[prob];
  P.params.gamma = eos.gamma;
  P.params.cv = eos.cv;
  P.params.cfl = prm->cfl;
  P.params.final_time = prm->final_time;
  P.params.kind = prm->flux_kind;
  P.params.first_order = prm->first_order != 0;
  P.params.stop_at_final = prm->stop_at_final != 0;
  P.params.max_steps = prm->max_steps;
  if ((rc = stage_scalars(ctx, prob, st, 0.0, 0.0)) || (rc = stage_state(ctx, prob, st, d, v, p, nullptr)) ||
      (rc = euler_initial_state(ctx, prob, ctx->upload)) || (rc = st.finish())) {
    (void)st.finish();
    fsale_batch_remove(ctx, prob);
    return rc;
  }
  P.have_state = true;
  P.partials_valid = false;
  *problem = prob;
  return FSALE_OK;
}

int fsale_batch_add_maire(fsale_ctx *ctx, const fsale_mesh_desc *mesh, double cv, const double *gamma,
                          const fsale_maire_params *prm, const double *d, const double *v, const double *p,
                          int32_t *problem) {
  if (!ctx || ctx->single || ctx->solver != FSALE_SOLVER_MAIRE || !mesh || !prm || !gamma || !d || !v || !p ||
      !problem || !(cv > 0))
    return FSALE_ERR_ARG;
  for (int c = 0; c < mesh->num_cells; ++c)
    if (!(gamma[c] > 1)) return FSALE_ERR_ARG;
  int prob = -1, rc;
  Stager st(ctx);
  if ((rc = place_mesh(ctx, *mesh, &prob, st))) {
    st.discard();
    (void)st.finish();
    if (prob >= 0 && prob < ctx->max_problems && ctx->probs[prob].live) fsale_batch_remove(ctx, prob);
    return rc;
  }
  Problem &P = ctx->probs[prob];
  P.params.cv = cv;
  P.params.cfl_acc = prm->cfl_accoustic;
  P.params.cfl_vol = prm->cfl_volume;
  P.params.cfl_growth = prm->cfl_growth;
  P.params.final_time = prm->final_time;
  P.params.max_steps = prm->max_steps;
  if ((rc = stage_scalars(ctx, prob, st, 0.0, prm->initial_time_step)) ||
      (rc = stage_state(ctx, prob, st, d, v, p, gamma)) || (rc = maire_initial_state(ctx, prob, ctx->upload)) ||
      (rc = st.finish())) {
    (void)st.finish();
    fsale_batch_remove(ctx, prob);
    return rc;
  }
  P.have_state = true;
  *problem = prob;
  return FSALE_OK;
}

int fsale_batch_set_maire_boundary(fsale_ctx *ctx, int32_t problem, int tag, int kind, const double *value) {
  if (!ctx || ctx->single || ctx->solver != FSALE_SOLVER_MAIRE || !valid_prob(ctx, problem)) return FSALE_ERR_ARG;
  int rc = set_maire_boundary(ctx, problem, tag, kind, value);
  return rc ? rc : upload_params(ctx, problem);
}

int fsale_batch_remove(fsale_ctx *ctx, int32_t problem) {
  if (!ctx || ctx->single || !valid_prob(ctx, problem)) return FSALE_ERR_ARG;
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  Problem &P = ctx->probs[problem];
  ctx->ra_c.release(P.c0, P.nc);
  ctx->ra_f.release(P.f0, P.nf);
  ctx->ra_v.release(P.v0, P.nv);
  ctx->ra_n.release(P.n0, P.nn);
  ctx->ra_ct.release(P.ct0, P.nct);
  ctx->ra_vt.release(P.vt0, P.nvt);
  P = Problem{};
  ctx->tiles_dirty = true;
  // a dead problem owns no cells: the per-problem kernels skip it
  ProbParams zero;
  std::memset(&zero, 0, sizeof(zero));
  FSALE_HIP(hipMemcpyAsync(ctx->d_params + problem, &zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream));
  FSALE_HIP(hipStreamSynchronize(ctx->stream));
  return FSALE_OK;
}

int fsale_batch_step(fsale_ctx *ctx, int nsteps, int use_graph) {
  if (!ctx || nsteps < 0) return FSALE_ERR_ARG;
  for (auto &P : ctx->probs)
    if (P.live && !P.have_state) return FSALE_ERR_STATE;
  return batch_steps(ctx, nsteps, use_graph);
}

int fsale_batch_status(fsale_ctx *ctx, int32_t max_out, fsale_problem_status *out, int32_t *count) {
  if (!ctx || !out || !count || max_out < 0) return FSALE_ERR_ARG;
  int rc = read_scalars(ctx);
  if (rc) return rc;
  int n = 0;
  for (int p = 0; p < ctx->max_problems && n < max_out; ++p) {
    const Problem &P = ctx->probs[p];
    if (!P.live) continue;
    const DevScalars &S = ctx->h_s[p];
    fsale_problem_status st;
    st.problem = p;
    st.error = S.err_code;
    st.steps = S.steps;
    st.time = S.t;
    st.time_step = S.dt;
    bool done = S.err_code != 0 || S.steps >= P.params.max_steps;
    if (ctx->solver == FSALE_SOLVER_MAIRE || P.params.stop_at_final) done = done || !(S.t < P.params.final_time);
    st.done = done ? 1 : 0;
    out[n++] = st;
  }
  *count = n;
  return FSALE_OK;
}

int fsale_batch_get_field(fsale_ctx *ctx, int32_t problem, int field, double *host_out) {
  return get_problem_field(ctx, problem, field, host_out);
}

int fsale_batch_get_fields(fsale_ctx *ctx, int32_t problem, int32_t n, const int32_t *fields, double *const *host_out) {
  return get_problem_fields(ctx, problem, n, fields, host_out);
}

int fsale_batch_capacity(fsale_ctx *ctx, int64_t *free_cells, int64_t *free_faces, int64_t *free_vertices,
                         int64_t *free_neighbours, int32_t *free_problems) {
  if (!ctx) return FSALE_ERR_ARG;
  if (free_cells) *free_cells = ctx->ra_c.available();
  if (free_faces) *free_faces = ctx->ra_f.available();
  if (free_vertices) *free_vertices = ctx->ra_v.available();
  if (free_neighbours) *free_neighbours = ctx->ra_n.available();
  if (free_problems) {
    int n = 0;
    for (auto &P : ctx->probs)
      if (!P.live && !P.quarantined) ++n;
    *free_problems = n;
  }
  return FSALE_OK;
}

int fsale_launch_shape(fsale_ctx *ctx, int32_t kernel, int32_t blocks_per_cu) {
  if (!ctx || kernel < -1 || kernel >= K_COUNT || blocks_per_cu < 0 || blocks_per_cu > 32) return FSALE_ERR_ARG;
  for (int k = 0; k < K_COUNT; ++k)
    if (kernel < 0 || k == kernel) ctx->shape.blocks_per_cu[k] = blocks_per_cu;
  return FSALE_OK;
}

int fsale_get_launch_shape(fsale_ctx *ctx, int32_t *blocks_per_cu) {
  if (!ctx || !blocks_per_cu) return FSALE_ERR_ARG;
  for (int k = 0; k < K_COUNT; ++k) blocks_per_cu[k] = ctx->shape.blocks_per_cu[k];
  return FSALE_OK;
}

int fsale_lds_staging(fsale_ctx *ctx, int32_t *on) {
  if (!ctx || !on) return FSALE_ERR_ARG;
  int rc;
  if ((rc = sync_tiles(ctx))) return rc;
  if (ctx->m.dim == 3)
    *on = (ctx->s3_all ? 2 : ctx->s3 ? 1 : 0) + (ctx->brick_all ? 4 : 0);
  else
    *on = ctx->lds_all ? 2 : ctx->lds ? 1 : 0;
  return FSALE_OK;
}

int fsale_autotune(fsale_ctx *ctx, int32_t nsteps, double *best_ms) {
  if (!ctx || nsteps < 1 || nsteps > 1000) return FSALE_ERR_ARG;
  for (auto &P : ctx->probs)
    if (P.live && !P.have_state) return FSALE_ERR_STATE;
  return autotune(ctx, nsteps, best_ms);
}

}  // extern "C"
