// Internal declarations shared by the host and device translation units of
// libflecsale_amd.  Not part of the public ABI (include/flecsale_amd.h).
//
// Every context is a ragged batch: problems (one mesh + state each) occupy
// disjoint ranges of capacity-sized device arrays, and every step kernel
// walks a device-side list of tiles (a tile = up to kBlock consecutive cells,
// or vertices, of one problem), one block per tile or in a persistent loop.  The plain drop-in API is
// a batch holding exactly one problem.  Problem-specific scalars (EOS, CFL,
// time, dt, step budget, error word) live in per-problem device records, so a
// problem's arithmetic is the same alone and in any mix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/flecsale_amd.h"

namespace fsale {

constexpr int kBlock = 256;     // threads per block = entities per tile
constexpr int kWaves = kBlock / 64;   // waves per block = dt partials per tile
constexpr int kMaxCornersPerVertex = 8;
// neighbours held k-major per cell: all cell-vertex neighbours of an interior
// cell of a structured quad (8) or hex (26) mesh
constexpr int kNbk2 = 8, kNbk3 = 28;   // 3D: 26 hex neighbours, padded to groups of 4
// 2D quad tiles staged in LDS by the Eulerian driver kernels: at most this
// many cells outside a tile (a 16 x 16 block of a structured mesh has 68),
// sized so that three blocks of the fused kernel's staging fit one CU's LDS
constexpr int kHalo2 = 80;
// 3D hex tiles of a structured mesh (cells numbered x fastest, then y, then
// z) staged in LDS by the Eulerian gradient kernel one neighbour plane at a
// time: a tile's neighbours in the plane dz away lie in kBlock + 2 nx + 2
// consecutive cells, so meshes up to nx = 191 cells per row are staged
constexpr int kLds3 = 640;
// bricks of cells of the 3D fused step on structured grids (one cell per thread)
constexpr int kBrickX = 8, kBrickY = 8, kBrickZ = 4;
static_assert(kBrickX * kBrickY * kBrickZ == kBlock, "one brick cell per thread");

void set_message(const std::string &msg);

#define FSALE_HIP(call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::fsale::set_message(std::string(#call) + ": " + hipGetErrorString(e_)); \
      return (e_ == hipErrorOutOfMemory) ? FSALE_ERR_OOM : FSALE_ERR_HIP;      \
    }                                                                          \
  } while (0)

// Host-side connectivity of one mesh (mesh_host.cpp), local ids.  Conventions
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
  int max_corners = 0;            // most corners around one vertex
  int max_neighbours = 0;         // most cell-vertex neighbours of one cell
  // device storage order (a Z-order curve, for locality of the gathers):
  // xidx[s] = the entity stored at position s, xpos its inverse.  No result
  // depends on it: every ordered sum follows the original ids.
  std::vector<int32_t> cidx, cpos, vidx, vpos, fidx, fpos;
  // 3D: the order in which the cell tiles (kBlock consecutive storage
  // positions) are dealt to blocks.  The list is cut into 8 contiguous parts,
  // one per XCD; each part is a slab along the second-longest axis, swept
  // along the longest, so an XCD's blocks in flight cover a compact region
  // whose neighbours stay in its L2 (a whole plane of a 128^3 mesh would
  // not).  Empty: storage order.
  std::vector<int32_t> tile_order;
  // 2D quads: LDS staging plan of the cell tiles (kBlock consecutive storage
  // positions each).  halo_cnt[tile]: the number of cells outside the tile
  // that its cells' neighbour lists name (-1: more than kHalo2, or a cell
  // with more neighbours than the k-major head holds); halo[tile*kHalo2 + i]:
  // their storage positions, ascending.  The cells' LDS slots are derived
  // from these on the device (k_lds_slots).
  std::vector<int32_t> halo_cnt, halo;
  // 3D: the mesh is a structured nx x ny x nz hex grid numbered x fastest
  // (every cell's cell-vertex neighbours are the grid's, in that numbering)
  // and is stored in that order: s_nx = nx and s_np = nx * ny, else 0
  int s_nx = 0, s_np = 0;
  // 3D structured grid whose hex sides pair up: for every cell, its -x, -y
  // and -z sides (local faces 3, 0, 4) are, where a neighbour lies across
  // them, that neighbour's +x, +y and +z sides (local faces 1, 2, 5)
  int s_fp = 0;
  // With s_fp and nx, ny multiples of 8 and nz of 4: the fused step covers
  // the grid in 8 x 8 x 4 bricks of cells, one per cell tile; brick[j] is
  // the first cell of the brick of tile j.  Tiles are dispatched in
  // tile_order; the bricks are dealt to them so that the k-th tile
  // dispatched gets the k-th brick of a sweep in which each eighth of the
  // dispatch list (one XCD's) covers a slab of whole brick rows along y,
  // swept along z in layers of brick rows along x.  Empty: no bricks.
  std::vector<int32_t> brick;
};

int build_topology(const fsale_mesh_desc &mesh, HostTopology &t, std::string &err);
// Exodus-II output of one time step (exodus_write.cpp); cells 0-based [nc][vpc]
int exodus_write(const char *path, const char *title, int dim, int nv, const double *coords, int nc, int vpc,
                 const int32_t *cells, double time, int nvars, const char *const *names,
                 const double *const *values);
// connectivity of a mesh from the process-wide cache (built on a miss)
std::shared_ptr<const HostTopology> cached_topology(const fsale_mesh_desc &mesh, int &rc, std::string &err);

// Device-side mesh storage of a batch: SoA geometry, k-major fixed-degree
// connectivity with row stride = capacity, global (batch) entity ids.
struct DevMesh {
  int dim = 0, vpc = 0, fpc = 0, vpf = 0;
  int cap_c = 0, cap_f = 0, cap_v = 0;           // row strides
  double *x[3] = {nullptr, nullptr, nullptr};    // vertex coords [cap_v]
  int32_t *cv = nullptr;        // [vpc][cap_c]
  int32_t *cf = nullptr;        // [fpc][cap_c]: face*2 + (cell is fc[1] of the face)
  int32_t *fc = nullptr;        // [2][cap_f]
  int32_t *fv = nullptr;        // [vpf][cap_f]
  int32_t *nbr_start = nullptr; // [cap_c] cell-vertex neighbour list start
  int32_t *nbr_cnt = nullptr;   // [cap_c]
  int32_t *nbr = nullptr;       // [cap_n]
  // the first nbk_deg neighbours of each cell again, k-major, so a gradient
  // work-item issues all its neighbour loads at once (cells with more
  // neighbours continue in the CSR list); a shorter list is right-aligned,
  // -1 padding first
  int nbk_deg = 0;
  int32_t *nbk = nullptr;       // [nbk_deg][cap_c]
  // cell across local face k: (cell + 1) * 2 + (this cell is the face's
  // cells[1]); boundary faces hold 0
  int32_t *cfn = nullptr;       // [fpc][cap_c]
  int32_t *corig = nullptr;     // [cap_c] the problem-local id of the cell stored at each position
  int32_t *vc_cnt = nullptr;    // [cap_v] corners of each vertex
  int32_t *vck = nullptr;       // [kMaxCornersPerVertex][cap_v] corner c*vpc + k (batch cell
                                // id), ascending cell id, -1 padded
  // Maire wedge boundary tags of each corner, one nibble per face holding
  // the corner (2D: the faces of edges (k-1, k) and (k, k+1); 3D: the three
  // hex sides holding k, in side order): bit 2 (2D) / bit 3 (3D) = boundary
  // face, the low bits = its x / y / z symmetry tags
  uint16_t *cwtag = nullptr;    // [vpc][cap_c]
  uint8_t *vbnd = nullptr;      // [cap_v] vertex on a boundary face
  int32_t *bv = nullptr;        // [cap_v] a problem's boundary vertices (batch ids, ascending)
                                // at positions [v0, v0 + nb) of its vertex range
  double *ccen[3] = {nullptr, nullptr, nullptr};
  double *cvol = nullptr, *cminlen = nullptr;
  double *fcen[3] = {nullptr, nullptr, nullptr};
  double *fnrm[3] = {nullptr, nullptr, nullptr};
  double *farea = nullptr;
  // 3D: the geometry of each cell's local faces again, k-major [fpc][cap_c],
  // so the fused step reads it coalesced instead of gathering 8 bytes per
  // 128-byte line: pairs (centroid x, y), (centroid z, normal x), (normal y,
  // z) as stored for the face, and the area; four loads per face
  double2 *kfp[3] = {};
  double *kfa = nullptr;
  // 2D: vertex coordinates again as (x, y) pairs [cap_v], one 16-byte
  // gather per vertex (the Lagrangian kernels that move vertices write both)
  double2 *xy = nullptr;
  // 2D quads, Eulerian: the cell tiles' LDS staging plan (HostTopology::
  // halo_cnt, halo) by tile slot, in batch ids, and per cell the LDS slots of
  // its neighbour slots (lnb) and face neighbours (lfn) as 16-bit pairs,
  // k-major: row k holds slots 2k (low half) and 2k + 1
  int32_t *halo_cnt = nullptr;   // [tile capacity]
  int32_t *halo = nullptr;       // [tile capacity * kHalo2]
  int32_t *lnb = nullptr;        // [kNbk2 / 2][cap_c]
  int32_t *lfn = nullptr;        // [2][cap_c]
  // 3D Eulerian: the first cell of the 8 x 8 x 4 brick of cells that the
  // fused step covers for each cell tile slot (HostTopology::brick)
  int32_t *brick = nullptr;      // [tile capacity]
};

// A tile: `count` (<= kBlock) consecutive entities of problem `prob`
// starting at global id `first`; `slot` indexes per-tile scratch (dt partials).
struct Tile {
  int prob, first, count, slot;
};

// Device tile list: tiles[0..*n); or, for one-block-per-tile launches,
// direct[blockIdx] for blocks below ndirect (see for_each_tile).
struct TileList {
  const Tile *tiles = nullptr;
  const int *n = nullptr;
  const Tile *direct = nullptr;
  int ndirect = 0;
};

// Per-problem parameters (read-only in a step).
struct ProbParams {
  double gamma, cv;                        // Eulerian EOS (Maire: per-cell gamma, cv here)
  double cfl, final_time;                  // Eulerian CFL / final time
  double cfl_acc, cfl_vol, cfl_growth;     // Maire time_constants_t
  long long max_steps;                     // step budget: active while steps < max_steps
  int kind, first_order, stop_at_final;
  int owned;                               // Eulerian: cells with original id >= owned are ghosts
                                           // of another rank (0: every cell is owned)
  int tile_first, tile_count;              // the problem's cell tiles (dt partials)
  int cell_first, cell_count;
  int s_nx, s_np;                          // 3D structured grid strides (HostTopology), 0 if none
  // Maire boundary conditions of the faces of each symmetry tag (x, y, z):
  // FSALE_BC_* kind, pressure in bc_val[t][0] or velocity in bc_val[t][:dim]
  int bc_kind[3];
  int s_fp;                                // HostTopology::s_fp: sides pair up across the grid
  double bc_val[3][3];
};

// Per-problem scalars that stay on the device between tasks (the dt future,
// the solution time, the error word).
struct DevScalars {
  double dt;            // current time step (the reference's dt future)
  double t;             // solution time
  int active;           // the current step is live (written between launches only)
  int err_code;         // first fsale_status raised on the device
  long long err_cell;   // lowest failing cell (local id of the problem)
  long long steps;      // completed steps
  int err_token;        // launch that raised err_code (0: none)
  int packed;           // Eulerian: steps in the current batch of steps (its state was packed)
  // Lagrangian 2D driver step: the step as its predictor update sees it
  // (written by the predictor's residual launch: the previous step closed,
  // and the dt it closed with), so that the update launch can finalize the
  // time step and close the previous step itself without reading what it
  // writes (k_maire_move_update FOLD)
  double snap_t, snap_dt;
  long long snap_steps;
  int snap_active;
  // Lagrangian 2D driver steps ran since the last k_maire_settle: the fields
  // no kernel of a step reads are not current (k_maire_settle writes them)
  int lazy;
};

struct EulerFields {
  double *d = nullptr, *v[3] = {nullptr, nullptr, nullptr}, *e = nullptr, *p = nullptr,
         *T = nullptr, *a = nullptr;
  double *gd[3] = {}, *gv[9] = {}, *ge[3] = {}, *gp[3] = {}, *ga[3] = {};
  // The driver loop's working copies, one record per cell holding what a
  // neighbour gathers, stored as pairs of doubles in planes of capacity
  // pairs, so a gather takes one 16-byte load per pair instead of one 8-byte
  // load per field:
  //   gs: {d, v[N], e, p, a, centroid[N]} (2D: 8 doubles; 3D: 10),
  //       twice: the fused flux+update kernel reads one and writes the other
  //       (its neighbours must see the old state), so a step flips twice;
  //   gg: gradients {d[N], v[N][N], e[N], p[N], a[N]} (2D: 12; 3D: 21 + 1 pad).
  // Packed from the fields above when a batch of steps starts, unpacked when
  // it ends.
  // Each plane's base is its own pointer (a scalar base for every load).
  static constexpr int kPlanesS = 5, kPlanesG = 11;   // the 3D counts; 2D uses 4 and 6
  double2 *gs[kPlanesS] = {}, *alt_gs[kPlanesS] = {}, *gg[kPlanesG] = {};
  double *F[5] = {};
  double *partials = nullptr;    // per cell tile and wave: max of dt^-1
};

struct MaireFields {
  double *V = nullptr, *M = nullptr, *v[3] = {}, *p = nullptr, *d = nullptr, *e = nullptr,
         *T = nullptr, *a = nullptr, *g = nullptr;
  double *v0[3] = {}, *e0 = nullptr, *x0[3] = {};
  double *un[3] = {};
  double *npc[3] = {}, *Fpc[3] = {};   // corner arrays [vpc][cap_c]
  double *dudt[5] = {};                // volume, momentum[dim], energy
  double *partials = nullptr;          // per cell tile and wave: max a/min_length, max |dV/dt|/V
  // 3D: what the nodal solve gathers, packed per entity so that a corner's
  // gathers touch one cache line per cell or vertex instead of one per array
  // (xa and cs written by every kernel that writes their sources, and by
  // maire_pack_problem after host writes)
  int32_t *cv8 = nullptr;              // [cap_c][8] cell vertices (batch ids), written with the mesh
  double *xa = nullptr;                // [cap_v][4] coordinates (x, y, z, 0)
  double *cs = nullptr;                // [cap_c][6] (rho a, p, v, 0)
  // 2D: the same as two planes of pairs, (rho a, p) and (v_x, v_y) [cap_c]
  // (the coordinates are DevMesh::xy)
  double2 *cs2[2] = {};
  // 2D driver steps: the coordinates at the start of the step, as pairs,
  // saved by the step's first nodal solve (x0 is written from them at the
  // end of a batch of steps, k_maire_settle); the predictor and corrector
  // updates move the vertices from these and rebuild each cell's geometry
  // from the same values
  double2 *x0p = nullptr;
};

// Kernel arguments common to all step kernels
struct Batch {
  TileList ct;                // cell tiles
  TileList vt;                // vertex tiles (Maire)
  TileList bt;                // boundary-vertex tiles (Maire): entities index DevMesh::bv
  const ProbParams *pp = nullptr;
  DevScalars *s = nullptr;
  int cap_c_tiles = 0;        // partials stride
  const int *plist = nullptr; // problems of the tile lists, in id order
  const int *np = nullptr;    // their count
};

// Instantiated step graphs of a context, keyed by what capture bakes in
// (chunk length, launch grids, staging mode).  An executable graph is
// destroyed only after an event recorded behind its last launch has
// completed: a graph launch is asynchronous, and a key change between two
// calls (serving rounds change the persistent grid with the live-problem
// count) used to destroy an exec whose previous launch could still be
// running.  A few are kept, so alternating keys do not recapture.
struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;   // recorded on the context's stream after each launch
  std::vector<double> key;
  uint64_t used = 0;
};

struct GraphCache {
  static constexpr size_t kMaxEntries = 6;
  std::vector<GraphEntry> entries;
  uint64_t clock = 0;
};

// Per-kernel timing with HIP events on the context's stream (profiling mode
// runs the step loops eagerly so every launch is bracketed).
enum KernelId {
  K_EUL_DT = 0, K_EUL_GRAD, K_EUL_FLUX, K_EUL_UPDATE,
  K_MAIRE_NODAL, K_MAIRE_RESIDUAL, K_MAIRE_DT, K_MAIRE_MOVE, K_MAIRE_UPDATE, K_MAIRE_COPY,
  K_EUL_FLUXUPD, K_MAIRE_NODAL_BND,
  K_COUNT
};

struct ProfSlot {
  std::vector<hipEvent_t> start, stop;
  double total_ms = 0;
  long long launches = 0;
};

// Host record of one resident problem
struct Problem {
  bool live = false;
  // removed while a round that still lists it may be running (a deferred
  // retirement): the slot is not reused until the stream has been waited for
  bool quarantined = false;
  int c0 = 0, nc = 0, f0 = 0, nf = 0, v0 = 0, nv = 0, n0 = 0, nn = 0;   // ranges
  int ct0 = 0, nct = 0, vt0 = 0, nvt = 0;                               // tile ranges
  int nb = 0;                                                           // boundary vertices (Maire)
  int nb2 = 0;   // of these, the ones on fewer than two tagged faces (listed first)
  std::shared_ptr<const HostTopology> topo;
  ProbParams params{};
  bool faces_on_device = true;   // face connectivity and geometry uploaded
  bool lds = false;              // the cell tiles' LDS staging plan is uploaded
  bool lds_all = false;          // ... and every tile has one (no tile gathers from memory)
  bool brick = false;            // the fused step runs over bricks (HostTopology::brick uploaded)
  bool have_state = false;
  bool partials_valid = false;
};

// first-fit range allocator over [0, capacity)
struct RangeAlloc {
  int capacity = 0;
  std::map<int, int> free;   // start -> length
  void init(int cap) {
    capacity = cap;
    free.clear();
    if (cap > 0) free[0] = cap;
  }
  int alloc(int n);          // -1 if no room
  bool fits(int n) const;    // a free range of n exists
  void release(int start, int n);
  long long available() const;
};

// Launch shape of each persistent kernel (autotuned per context):
// blocks_per_cu[k] = 0: one block per live tile (grid = live tiles rounded
// up to a power of two, so graphs survive most admissions); n > 0: a
// persistent grid of CUs * n blocks that loops over the tiles.
struct LaunchShape {
  int blocks_per_cu[K_COUNT] = {};
};

}  // namespace fsale

struct fsale_ctx {
  int solver = 0;
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  bool single = false;                 // plain drop-in context (one problem)
  fsale::DevMesh m;
  fsale::EulerFields eu;
  fsale::MaireFields ma;
  // problems, tiles and their device mirrors
  std::vector<fsale::Problem> probs;   // index = problem id
  int max_problems = 0;
  fsale::RangeAlloc ra_c, ra_f, ra_v, ra_n, ra_ct, ra_vt;
  fsale::Tile *d_ctiles = nullptr, *d_vtiles = nullptr;
  fsale::Tile *d_cdirect = nullptr, *d_vdirect = nullptr;   // one-block-per-tile dispatch lists
  int n_cdirect = 0, n_vdirect = 0;
  fsale::Tile *d_btiles = nullptr, *d_bdirect = nullptr;    // boundary-vertex tiles (Maire)
  int n_bdirect = 0;
  int *d_counts = nullptr;             // [0] cell tiles, [1] vertex tiles, [2] problems, [3] boundary tiles,
                                       // [4] boundary tiles of vertices on two or more tagged faces
  // 3D Maire: the boundary vertices on two or three tagged faces (each
  // problem's tail of DevMesh::bv from nb2), the only ones whose symmetry
  // solve can have two or three rows, in tiles of one block pass each
  fsale::Tile *d_bbtiles = nullptr;
  int n_bbtiles = 0;
  int *d_plist = nullptr;              // problems with tiles in the lists
  int pgrid = 1;                       // blocks of the per-problem kernels
  hipStream_t upload = nullptr;        // admissions: uploads, geometry, initial state
  fsale::ProbParams *d_params = nullptr;
  fsale::DevScalars *s = nullptr;      // [max_problems] device
  fsale::DevScalars *h_s = nullptr;    // pinned host mirror
  std::vector<fsale::Tile> h_ctiles, h_vtiles, h_btiles, h_bbtiles;
  bool tiles_dirty = true;
  bool lds = false;                    // every live problem has LDS staging plans (sync_tiles)
  bool lds_all = false;                // ... for every one of its tiles
  bool s3 = false, s3_all = false;     // some / every live problem is a structured 3D grid (s_nx)
  bool fp3 = false;                    // some live problem's hex sides pair up (s_fp)
  bool brick_all = false;              // every live problem's fused step runs over bricks
  bool tasks_open = true;              // no driver run since the step flag was last set for task calls
  void *stage = nullptr;               // pinned staging buffer for uploads
  size_t stage_bytes = 0;
  void *dl = nullptr;                  // pinned buffer for multi-field downloads
  hipEvent_t ev_dl = nullptr;          // deferred downloads: the compute stream's work so far
  size_t dl_bytes = 0;
  fsale::GraphCache graph;
  fsale::LaunchShape shape;
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
int read_scalars(fsale_ctx *ctx);      // sync + copy DevScalars to h_s (and lift slot quarantines)
// Task-level entry points of a single-mesh context: a driver run leaves the
// problem's step flag (DevScalars::active) cleared once its step budget or
// final time is reached, and every step kernel skips a problem whose flag is
// clear; the first task call after a run sets it again (an error the run
// latched still stops the kernels, as it stops the run).
int open_tasks(fsale_ctx *ctx);
int set_maire_boundary(fsale_ctx *ctx, int prob, int tag, int kind, const double *value);
// Lagrangian contexts: refreshes the nodal solve's packed copies of one
// problem's cell state and coordinates (no-op otherwise)
int maire_pack_problem(fsale_ctx *ctx, int prob, hipStream_t stream);
// 2D Eulerian: the LDS slots of each cell's neighbours and face neighbours
// (DevMesh::lnb, lfn) from its k-major rows and its tile's halo list
int euler_lds_slots(fsale_ctx *ctx, int prob, hipStream_t stream);

// Retirement of finished problems while the next round runs (serving loop).
// stage_field_downloads enqueues, on the upload stream and after all work
// enqueued so far on the compute stream, the copies of `fields` of each
// listed problem into the pinned download buffer; finish_field_downloads
// waits for them and writes host_out (host order).  retire_deferred then
// releases the problem's ranges without waiting for the compute stream: the
// round in flight still lists the problem's tiles, but the problem is done,
// so every kernel skips them; its slot stays unused until read_scalars has
// waited for the stream.  Downloads precede, on the upload stream, every
// later admission's upload into the released ranges.
struct PendingFields {
  std::shared_ptr<const HostTopology> topo;   // keeps the views' positions alive
  int n = 0;
  int ncomp[8] = {}, count[8] = {};
  const int32_t *pos[8] = {};
  double *outs[8] = {};
  size_t off = 0;                              // doubles into the download buffer
};
int stage_field_downloads(fsale_ctx *ctx, const int *probs, int nprobs, int n, const int32_t *fields,
                          double *const *const *outs, std::vector<PendingFields> &pend);
int finish_field_downloads(fsale_ctx *ctx, std::vector<PendingFields> &pend);
int retire_deferred(fsale_ctx *ctx, int prob);
int check_device_error(fsale_ctx *ctx, int prob);
int sync_tiles(fsale_ctx *ctx);        // upload the tile lists if they changed
int upload_params(fsale_ctx *ctx, int prob);
Batch batch_args(const fsale_ctx *ctx, int kid);
int persistent_grid(const fsale_ctx *ctx, int kid);

// geometry (geom.hip): cells/faces of one problem, on `stream`
int launch_geometry(fsale_ctx *ctx, int prob, bool cells, bool faces, hipStream_t stream);
int euler_initial_state(fsale_ctx *ctx, int prob, hipStream_t stream);
int maire_initial_state(fsale_ctx *ctx, int prob, hipStream_t stream);
// Euler (euler.hip)
int euler_alloc(fsale_ctx *ctx);
int euler_dt_partials(fsale_ctx *ctx);
int euler_step(fsale_ctx *ctx);       // one driver step over every live problem
int euler_pack(fsale_ctx *ctx);       // per-field state -> the driver loop's packed records
int euler_unpack(fsale_ctx *ctx);     // and back
// Maire (maire.hip)
int maire_alloc(fsale_ctx *ctx);
int maire_step(fsale_ctx *ctx, bool first, bool last);
}  // namespace fsale
