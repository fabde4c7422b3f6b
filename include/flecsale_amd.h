/*
 * flecsale_amd.h -- C ABI of the MI355X implementation of FleCSALE's two hydro
 * time-step paths (apps/hydro Eulerian MUSCL, apps/maire_hydro cell-centred
 * Lagrangian).
 *
 * Each entry point replaces one FleCSI task body of the reference; the task's
 * accessor arguments become fields owned by an opaque context that lives in
 * device memory (SoA, FP64; int32 connectivity).  Plain pointers and sizes
 * only.  No entry point throws; every one returns an fsale_status.
 *
 * Threading: one host thread per context.  Work is queued on the context's
 * HIP stream and is asynchronous until fsale_sync() or a host read
 * (fsale_get_field / fsale_get_time).  Device-side failures (negative density
 * or energy, non-positive inverse time step) are latched in a per-context
 * error word and reported by the next synchronising call.
 */
#ifndef FLECSALE_AMD_H
#define FLECSALE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
typedef enum {
  FSALE_OK = 0,
  FSALE_ERR_ARG = 1,          /* bad argument / size                                */
  FSALE_ERR_MESH = 2,         /* mesh is not a conforming quad or hex mesh          */
  FSALE_ERR_HIP = 3,          /* HIP runtime failure (message: fsale_last_message)  */
  FSALE_ERR_OOM = 4,          /* device allocation failed                           */
  FSALE_ERR_NEG_STATE = 5,    /* "Negative density or internal energy encountered!"
                                 apps/hydro/tasks.h:654-655,
                                 flecsale/eqns/lagrange_eqns.h:337-341,361-365      */
  FSALE_ERR_DT = 6,           /* "infinite delta t" apps/hydro/tasks.h:366-367      */
  FSALE_ERR_STATE = 7,        /* call out of order (e.g. no initial conditions)     */
  FSALE_ERR_NO_DEVICE = 8     /* no HIP device visible                              */
} fsale_status;

/* ------------------------------------------------------------------- types */
typedef struct fsale_ctx fsale_ctx;

/* A conforming mesh: quads (dim 2, CCW vertex order as in the Exodus files of
 * data/meshes) or hexes (dim 3, Exodus hex8 order).  Replaces the burton mesh
 * client handle client_handle_r<mesh_t> (apps/hydro/types.h:33,87-88). */
typedef struct {
  int32_t dim;               /* 2 or 3                                       */
  int32_t num_cells;
  int32_t num_vertices;
  int32_t verts_per_cell;    /* 4 (dim 2) or 8 (dim 3)                       */
  const double *coords;      /* [num_vertices * dim], interleaved            */
  const int32_t *cells;      /* [num_cells * verts_per_cell], 0-based        */
} fsale_mesh_desc;

/* flecsale::eos::ideal_gas_t<double>(gamma, cv), passed by value like the
 * reference's eos_t (flecsale/eos/ideal_gas.h:47-52, apps/hydro/tasks.h:48). */
typedef struct { double gamma; double cv; } fsale_eos;

/* Face flux functions: flux_function<eqns_t> (apps/hydro/types.h:118-125)
 * selects hlle_flux; rusanov_flux is flecsale/eqns/flux.h:52-65; HLLC is a
 * build-side addition (Toro ch. 10, same wave-speed bounds as HLLE). */
enum { FSALE_FLUX_HLLE = 0, FSALE_FLUX_RUSANOV = 1, FSALE_FLUX_HLLC = 2 };
/* Boundary flux, OR-ed into flux_kind: boundary_flux<eqns_t> = wall_flux
 * (apps/hydro/types.h:132-136, euler_eqns.h:313-324) is the default; the
 * transmissive variant (physical flux of the reconstructed interior state) is
 * a build-side addition for outflow problems. */
enum { FSALE_BOUNDARY_WALL = 0x00, FSALE_BOUNDARY_TRANSMISSIVE = 0x10 };

enum { FSALE_SOLVER_EULER = 0, FSALE_SOLVER_MAIRE = 1 };

/* Field ids for fsale_get_field / fsale_set_field.  Vectors are interleaved
 * [n * dim] on the host side; gradients are [num_cells * dim] (grad_v is
 * [num_cells * dim * dim], row d = dv_d/dx). */
enum {
  FSALE_F_DENSITY = 0, FSALE_F_VELOCITY, FSALE_F_INTERNAL_ENERGY, FSALE_F_PRESSURE,
  FSALE_F_TEMPERATURE, FSALE_F_SOUND_SPEED,
  FSALE_F_GRAD_DENSITY, FSALE_F_GRAD_VELOCITY, FSALE_F_GRAD_ENERGY, FSALE_F_GRAD_PRESSURE,
  FSALE_F_GRAD_SOUND_SPEED,
  FSALE_F_FLUX,              /* [num_faces * (dim+2)]: mass, momentum, energy  */
  FSALE_F_CELL_CENTROID, FSALE_F_CELL_VOLUME, FSALE_F_CELL_MIN_LENGTH,
  FSALE_F_FACE_CENTROID, FSALE_F_FACE_NORMAL, FSALE_F_FACE_AREA,
  FSALE_F_COORDS,            /* vertex coordinates (moving for Maire)           */
  FSALE_F_MASS, FSALE_F_VOLUME,         /* Maire cell mass and volume           */
  FSALE_F_NODE_VELOCITY,     /* Maire u_p [num_vertices * dim]                  */
  FSALE_F_CORNER_NORMAL,     /* Maire (l n)_pc [num_cells*4 * dim], corner c*4+k */
  FSALE_F_CORNER_FORCE,      /* Maire F_pc, same layout                         */
  FSALE_F_RESIDUAL,          /* Maire dU/dt [num_cells * 4]: dV, mom x, mom y, E */
  FSALE_F_GAMMA,             /* per-cell ideal-gas gamma                        */
  FSALE_F_COUNT
};

/* flecsale::linalg::qr (flecsale/linalg/qr.h:44-94, detail/qr_impl.h:38-288):
 * column-pivoted Householder solve of A x = b, n in 1..4, A row-major, on the
 * host with the same arithmetic as the Lagrangian nodal kernel's symmetry
 * solve (apps/maire_hydro/tasks.h:383-418). */
int fsale_qr_solve(int32_t n, const double *A, const double *b, double *x);

/* ------------------------------------------------------------ context life */
/* Builds connectivity (faces, face->cell with cells[0] the lowest cell id and
 * the normal out of it, cell-vertex neighbours, vertex->corner) on the host,
 * uploads it with the coordinates and computes geometry on the device.
 * Replaces specialization_tlt_init + the update_geometry task
 * (apps/hydro/tasks.h:32-36, apps/maire_hydro/tasks.h:59-61). */
int fsale_ctx_create(const fsale_mesh_desc *mesh, int solver, fsale_ctx **out);
int fsale_ctx_destroy(fsale_ctx *ctx);
int fsale_sync(fsale_ctx *ctx);
/* counts = {num_cells, num_vertices, num_faces, num_neighbour_entries, dim} */
int fsale_mesh_counts(const fsale_ctx *ctx, int32_t *counts);
/* face_cells [num_faces*2] (-1 on the boundary), cell_faces [num_cells*fpc],
 * nbr_off [num_cells+1], nbr [counts[3]]; any pointer may be NULL. */
int fsale_get_topology(fsale_ctx *ctx, int32_t *face_cells, int32_t *cell_faces,
                       int32_t *nbr_off, int32_t *nbr);
/* Host-only connectivity of a mesh, without a device (the reference's
 * connectivity dump, the print task of apps/hydro/tasks.h:741-766).  Call
 * with NULL arrays to get counts, then again with arrays of those sizes.
 * Connectivity is kept in a process-wide cache keyed by the cell table
 * (budget FSALE_TOPOLOGY_CACHE_MB, default 8192), so this also registers a
 * mesh ahead of the contexts and batches that will use it. */
int fsale_mesh_connectivity(const fsale_mesh_desc *mesh, int32_t *counts, int32_t *face_cells,
                            int32_t *cell_faces, int32_t *nbr_off, int32_t *nbr);
int fsale_get_field(fsale_ctx *ctx, int field, double *host_out);
int fsale_set_field(fsale_ctx *ctx, int field, const double *host_in);
/* solution time, last time step and completed steps, as the reference's
 * solution_time color handle (apps/hydro/driver.h:274) and time_cnt. */
int fsale_get_time(fsale_ctx *ctx, double *soln_time, double *time_step, int64_t *steps);
int fsale_set_time(fsale_ctx *ctx, double soln_time, double time_step);
const char *fsale_status_string(int status);
const char *fsale_last_message(void);
/* first cell that failed (negative density/energy), or -1 */
int fsale_error_cell(fsale_ctx *ctx, int64_t *cell);

/* Exodus-II mesh files: what the apps' -m option reads (apps/hydro/2d/
 * CMakeLists.txt:38-48 runs hydro_2d -m data/meshes/square_32x32.g; the
 * reader itself is flecsi-sp's, not vendored).  Vertex coordinates and the
 * element blocks' connectivity in file order (blocks concatenated, indices
 * made 0-based); quad4 in 2D, hex8 in 3D.  netCDF classic encodings are read
 * directly; netCDF-4 (HDF5) files through an HDF5 library loaded at run time
 * (FSALE_HDF5_LIB or libhdf5.so).  fsale_exodus_info gives the sizes of the
 * arrays fsale_exodus_read fills: coords[num_vertices*dim],
 * cells[num_cells*verts_per_cell]. */
int fsale_exodus_info(const char *path, int32_t *dim, int32_t *num_cells, int32_t *num_vertices,
                      int32_t *verts_per_cell);
int fsale_exodus_read(const char *path, double *coords, int32_t *cells);

/* The dump task (apps/hydro/tasks.h:771-849; apps/maire_hydro/tasks.h:728):
 * the regression text file of a plain context, one rank: solution time and
 * iteration count, then per cell its local and global id, centroid,
 * density, internal energy, pressure and velocity, then per vertex its ids
 * and coordinates, 15 significant digits (scientific, precision 14), as the
 * reference's .std goldens. */
int fsale_dump(fsale_ctx *ctx, const char *path);
/* The output task (apps/hydro/tasks.h:703-734; apps/maire_hydro/tasks.h:
 * 661-692) of a plain context: an Exodus-II file of the mesh as it is now
 * (moved, for a Lagrangian context), its solution time and the task's cell
 * variables, Eulerian "density" and "pressure", Lagrangian "density".  The
 * file is named as the task names it, Eulerian
 * <prefix>_rank<rank:06>.<iteration:06>.<postfix>, Lagrangian
 * <prefix>_rank<rank:06>_<iteration:06>.<postfix> (the drivers pass postfix
 * "exo", apps/hydro/driver.h:285); prefix may hold a directory.  netCDF
 * classic with 64-bit offsets, read back by fsale_exodus_read. */
int fsale_output(fsale_ctx *ctx, const char *prefix, const char *postfix, int64_t iteration, int32_t rank);
/* The writer under fsale_output, for host arrays: one element block of
 * quad4 (dim 2, verts_per_cell 4) or hex8 (dim 3, 8) cells, 0-based
 * vertex ids [num_cells * verts_per_cell], coords [num_vertices * dim], one
 * time step holding num_vars cell variables values[i][num_cells] named
 * names[i] (at most 32 characters).  No counterpart in the reference beyond
 * the output task's call into flecsi-sp's io_exodus::write (not vendored). */
int fsale_exodus_write(const char *path, const char *title, int32_t dim, int32_t num_vertices,
                       const double *coords, int32_t num_cells, int32_t verts_per_cell, const int32_t *cells,
                       double time, int32_t num_vars, const char *const *names, const double *const *values);
/* Per-kernel timing (the reference's only timer is the wall-clock print
 * around the step loop, apps/hydro/driver.h:311,419).  While enabled, the
 * step loops run without graphs and each launch is bracketed by HIP events on
 * the context's stream.  Kernel ids: 0 dt, 1 gradients, 2 fluxes, 3 update
 * (Eulerian); 4 nodal, 5 residual, 6 dt, 7 move, 8 update, 9 save/restore
 * (Lagrangian); 10 fused fluxes + update (the Eulerian step loop's kernel);
 * 11 nodal solve of the boundary vertices (Lagrangian symmetry QR).
 * total_ms / launches accumulate until fsale_reset_profile. */
int fsale_set_profiling(fsale_ctx *ctx, int enable);
int fsale_get_profile(fsale_ctx *ctx, int kernel, double *total_ms, int64_t *launches);
int fsale_reset_profile(fsale_ctx *ctx);

/* ------------------------------------------------- Eulerian (apps/hydro) */
/* initial_conditions task (apps/hydro/tasks.h:46-68): sets d, v, p from the
 * host arrays (inputs_t::initial_conditions evaluated by the caller) and
 * completes the state with euler_eqns_t::update_state_from_pressure
 * (flecsale/eqns/euler_eqns.h:333-350).  v is [num_cells*dim]. */
int fsale_eul_initial_conditions(fsale_ctx *ctx, fsale_eos eos, const double *d,
                                 const double *v, const double *p);
/* evaluate_time_step (apps/hydro/tasks.h:331-376), a min-reduction task:
 * dt = min(CFL / max_{c,f} (a_c + |v_c.n_f|) A_f / V_c, t_final - t).  The
 * result stays on the device for apply_update; dt_out may be NULL (no sync). */
int fsale_eul_evaluate_time_step(fsale_ctx *ctx, double cfl, double final_time, double *dt_out);
/* No counterpart in the reference.  The same time step for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): every argument but the scalars is
 * a device pointer the caller owns -- cell_faces [num_cells*faces_per_cell]
 * int32 (mesh.faces(c) order), face_normal [num_faces*dim], face_area,
 * cell_volume, velocity [num_cells*dim], sound_speed -- and the result goes
 * to *dt_out with FSALE_OK or FSALE_ERR_DT (no finite time step) in
 * *status_out, both on the device, asynchronously on `stream` (a
 * hipStream_t; NULL = the null stream).  The partial maxima live in a fixed
 * device array of the library (2 MiB per GPU, kept until the process ends):
 * at most 2^24 cells, larger sizes return FSALE_ERR_ARG before any launch.
 * Calls that overlap on one GPU share that array and race: serialise them. */
int fsale_eul_time_step_noalloc(int32_t dim, int32_t num_cells, int32_t faces_per_cell,
                                const int32_t *cell_faces, const double *face_normal, const double *face_area,
                                const double *cell_volume, const double *velocity, const double *sound_speed,
                                double cfl, double final_time, double soln_time, double *dt_out,
                                int32_t *status_out, void *stream);
/* evaluate_gradients (apps/hydro/tasks.h:145-321): LSQ gradients with the
 * vertex limiter.  first_order != 0 zeroes them (build-side switch). */
int fsale_eul_evaluate_gradients(fsale_ctx *ctx, int first_order);
/* evaluate_fluxes (apps/hydro/tasks.h:385-536) with flux_function =
 * flux_kind and boundary_flux = wall_flux (apps/hydro/types.h:132-136). */
int fsale_eul_evaluate_fluxes(fsale_ctx *ctx, int flux_kind);
/* ------------------------------------------------------------ ranks
 * A partitioned mesh (apps/hydro run on several ranks): each rank's cells --
 * its owned cells first, then its ghosts, one layer of cells sharing a vertex
 * with an owned cell -- are a context of their own.  The owned/ghost split
 * replaces FleCSI's index spaces (mesh.cells(flecsi::owned) in
 * evaluate_time_step and apply_update, apps/hydro/tasks.h:350,614): with
 * num_owned < num_cells the context's time step and update skip its ghost
 * cells (their state arrives from the owner), and the whole-driver
 * fsale_eul_advance is refused (FSALE_ERR_STATE). */
int fsale_eul_set_owned_cells(fsale_ctx *ctx, int32_t num_owned);
/* Ghost refresh between two ranks' contexts on one device, FleCSI's ghost
 * copy of a dense field (apps/hydro/driver.h:329-375): cell dst_cells[i] of
 * `dst` receives the values of cell src_cells[i] of `src` (original cell ids
 * of each context).  An exchange is asynchronous, ordered after the work
 * queued on src's stream and before src's later work. */
typedef struct fsale_halo fsale_halo;
enum { FSALE_HALO_STATE = 0,        /* d, v, e, p, T, a (after apply_update)          */
       FSALE_HALO_GRADIENTS = 1 };  /* the five gradients (after evaluate_gradients) */
int fsale_halo_create(fsale_ctx *dst, fsale_ctx *src, int32_t n, const int32_t *dst_cells,
                      const int32_t *src_cells, fsale_halo **out);
int fsale_halo_exchange(fsale_halo *halo, int what);
int fsale_halo_destroy(fsale_halo *halo);
/* apply_update (apps/hydro/tasks.h:589-659): dt * factor from the device dt,
 * advances the solution time like update_soln_time (:545-576). */
int fsale_eul_apply_update(fsale_ctx *ctx, fsale_eos eos, double factor);
/* The driver step loop (apps/hydro/driver.h:317-413): nsteps x { dt;
 * 2 x [gradients, fluxes, apply_update(0.5)] }, all on the device, captured
 * once as a hipGraph when use_graph != 0. stop_at_final != 0 turns steps after
 * t >= final_time into no-ops on the device (the reference loop keeps
 * stepping with dt = 0).  The loop keeps its gradients and face fluxes
 * internal: the FSALE_F_GRAD_* and FSALE_F_FLUX fields hold what the last
 * task-level evaluate_gradients / evaluate_fluxes call wrote. */
int fsale_eul_advance(fsale_ctx *ctx, fsale_eos eos, double cfl, double final_time, int nsteps,
                      int flux_kind, int first_order, int stop_at_final, int use_graph);

/* --------------------------------------- Maire Lagrangian (apps/maire_hydro) */
/* initial_conditions (apps/maire_hydro/tasks.h:71-101): M = rho V, V = cell
 * volume, then lagrange_eqns_t::update_state_from_pressure
 * (flecsale/eqns/lagrange_eqns.h:245-271).  gamma is per cell (the reference
 * has one eos_t; a uniform array reproduces it). */
int fsale_maire_initial_conditions(fsale_ctx *ctx, double cv, const double *gamma,
                                   const double *d, const double *v, const double *p);
/* evaluate_nodal_state (apps/maire_hydro/tasks.h:221-453) with the boundary
 * conditions of fsale_maire_set_boundary (symmetry by default) on the
 * boundary faces whose normal is an axis (tags of
 * apps/maire_hydro/2d/inputs_regression.cc:134-162,
 * apps/maire_hydro/3d/inputs.cc:116-176). */
int fsale_maire_evaluate_nodal_state(fsale_ctx *ctx);
/* Boundary condition classes (apps/maire_hydro/types.h:88-126,
 * boundary_condition_t: has_prescribed_velocity / has_prescribed_pressure /
 * has_symmetry with velocity(x, t) / pressure(x, t)) with constant values,
 * per tag: the boundary faces whose unit normal has |n_tag| = 1 within 1e-14
 * (tag 0 x, 1 y, 2 z).  FSALE_BC_SYMMETRY (symmetry_boundary_condition_t,
 * the default), FSALE_BC_PRESSURE (value[0] = p: each boundary wedge of the
 * face subtracts l p n from its vertex's right-hand side, tasks.h:349-356),
 * FSALE_BC_VELOCITY (value[0..dim-1] = u: a vertex of the face takes u as its
 * velocity, no solve and, as the task's `continue` does, no corner-force
 * scatter, tasks.h:322-336), FSALE_BC_NONE (the face adds no condition). */
enum { FSALE_BC_SYMMETRY = 0, FSALE_BC_PRESSURE = 1, FSALE_BC_VELOCITY = 2, FSALE_BC_NONE = 3 };
int fsale_maire_set_boundary(fsale_ctx *ctx, int tag, int kind, const double *value);
/* evaluate_residual (apps/maire_hydro/tasks.h:461-489). */
int fsale_maire_evaluate_residual(fsale_ctx *ctx);
/* evaluate_time_step (apps/maire_hydro/tasks.h:142-190) followed by the
 * driver's cap min(dt, t_final - t) (apps/maire_hydro/driver.h:391); the
 * previous step is the device-held one.  dt_out may be NULL. */
int fsale_maire_evaluate_time_step(fsale_ctx *ctx, double cfl_accoustic, double cfl_volume,
                                   double cfl_growth, double final_time, double *dt_out);
/* apply_update (apps/maire_hydro/tasks.h:497-548) with delta_t = factor * dt
 * from the saved coordinates, then update_state_from_energy (:110-133). */
int fsale_maire_apply_update(fsale_ctx *ctx, double factor);
/* save_coordinates + save_solution / restore_coordinates + restore_solution
 * (apps/maire_hydro/tasks.h:557-654). */
int fsale_maire_save(fsale_ctx *ctx);
int fsale_maire_restore(fsale_ctx *ctx);
/* The driver loop (apps/maire_hydro/driver.h:330-538): midpoint predictor /
 * corrector; stops when t >= final_time (checked on the device). */
int fsale_maire_advance(fsale_ctx *ctx, double cfl_accoustic, double cfl_volume,
                        double cfl_growth, double final_time, int nsteps, int use_graph);

/* ------------------------------------------------------- ragged batches */
/* Many independent problems (each its own mesh, EOS, CFL and clock) resident
 * in one context and stepped together by the same persistent kernels.  The
 * reference runs one problem per MPI job (apps/hydro/driver.h:237-420,
 * apps/maire_hydro/driver.h:330-538 are its step loops); a batch runs each
 * resident problem through exactly that loop, with the same arithmetic as a
 * context of its own: a problem's results do not depend on what else is
 * resident.  All problems of a batch share the solver, dimension and cell
 * type.  Problems are admitted and retired between steps; the capacities fix
 * the device footprint at creation. */

/* Per-problem driver parameters of the Eulerian loop (apps/hydro/driver.h
 * inputs: CFL, final_time, max_steps; flux_kind as fsale_eul_advance). */
typedef struct {
  double cfl;
  double final_time;
  int32_t flux_kind;
  int32_t first_order;
  int32_t stop_at_final;
  int64_t max_steps;
} fsale_euler_params;

/* Per-problem parameters of the Lagrangian loop (time_constants_t of
 * apps/maire_hydro/tasks.h:142-190, initial_time_step and final_time of
 * apps/maire_hydro/driver.h:330-338). */
typedef struct {
  double cfl_accoustic;
  double cfl_volume;
  double cfl_growth;
  double final_time;
  double initial_time_step;
  int64_t max_steps;
} fsale_maire_params;

typedef struct {
  int32_t problem;
  int32_t error;             /* fsale_status latched on the device             */
  int64_t steps;             /* completed steps                                */
  double time;               /* solution time                                  */
  double time_step;          /* last time step                                 */
  int32_t done;              /* max_steps reached, final time reached (Lagrangian,
                                or stop_at_final) or error                     */
} fsale_problem_status;

/* An empty batch.  Capacities bound the sum over resident problems of cells,
 * faces, vertices and cell-vertex neighbour entries. */
int fsale_batch_create(int solver, int dim, int verts_per_cell, int32_t max_problems,
                       int64_t cell_capacity, int64_t face_capacity, int64_t vertex_capacity,
                       int64_t neighbour_capacity, fsale_ctx **out);
/* Admit a problem: connectivity, geometry and the initial_conditions task of
 * the solver; *problem receives its id.  FSALE_ERR_OOM when it does not fit. */
int fsale_batch_add_euler(fsale_ctx *ctx, const fsale_mesh_desc *mesh, fsale_eos eos,
                          const fsale_euler_params *params, const double *d, const double *v,
                          const double *p, int32_t *problem);
int fsale_batch_add_maire(fsale_ctx *ctx, const fsale_mesh_desc *mesh, double cv, const double *gamma,
                          const fsale_maire_params *params, const double *d, const double *v,
                          const double *p, int32_t *problem);
/* fsale_maire_set_boundary for one resident Lagrangian problem of a batch. */
int fsale_batch_set_maire_boundary(fsale_ctx *ctx, int32_t problem, int tag, int kind, const double *value);
/* Retire a problem and release its ranges (synchronises the stream). */
int fsale_batch_remove(fsale_ctx *ctx, int32_t problem);
/* nsteps driver steps of every resident problem that is not done; captured
 * as hipGraph chunks when use_graph != 0. */
int fsale_batch_step(fsale_ctx *ctx, int nsteps, int use_graph);
/* Status of up to max_out resident problems in id order. */
int fsale_batch_status(fsale_ctx *ctx, int32_t max_out, fsale_problem_status *out, int32_t *count);
int fsale_batch_get_field(fsale_ctx *ctx, int32_t problem, int field, double *host_out);
/* n fields of one problem (fields[i] into host_out[i]) with one device
 * synchronisation. */
int fsale_batch_get_fields(fsale_ctx *ctx, int32_t problem, int32_t n, const int32_t *fields,
                           double *const *host_out);
/* Free capacity (sums; a range still has to fit one contiguous gap). */
int fsale_batch_capacity(fsale_ctx *ctx, int64_t *free_cells, int64_t *free_faces,
                         int64_t *free_vertices, int64_t *free_neighbours, int32_t *free_problems);
/* Launch shape of the tile-list kernels (kernel ids as fsale_get_profile;
 * kernel -1 sets all): blocks_per_cu 0 launches one block per resident
 * tile of 256 entities, 1-32 a persistent grid of (compute units x
 * blocks_per_cu) blocks looping over the tiles.  Shapes change timing only,
 * never results.  Any context (a plain one is a batch of one). */
int fsale_launch_shape(fsale_ctx *ctx, int32_t kernel, int32_t blocks_per_cu);
/* blocks_per_cu[12] receives the current shape of every kernel id. */
int fsale_get_launch_shape(fsale_ctx *ctx, int32_t *blocks_per_cu);
/* LDS staging of the Eulerian driver kernels.  2D quads: each tile of 256
 * cells with at most 80 cells around it is staged in LDS with them, the
 * others gather from memory; the default storage order of 2D meshes above
 * 512^2 cells, spatial blocks of ~256 cells, makes a structured mesh's
 * tiles 16 x 16 blocks with 68 cells around each.  3D hexes: the gradient
 * kernel stages the tiles of a structured grid numbered x fastest (rows of
 * at most 191 cells) one neighbour plane at a time.  *on = 2: every
 * resident problem's tiles are staged; 1: some are; 0: none (or not an
 * Eulerian context).  3D, plus 4: the fused flux + update kernel runs over
 * 8 x 8 x 4 bricks of cells whose neighbours' face states meet in LDS (every
 * resident problem a structured grid with nx, ny multiples of 8 and nz of
 * 4).  No counterpart in the reference (a property of this
 * implementation's kernels); results are the same either way. */
int fsale_lds_staging(fsale_ctx *ctx, int32_t *on);
/* Autotuner: times nsteps eager steps of the resident problems under each
 * candidate shape (HIP events per launch) from one snapshot of the state,
 * keeps the fastest shape per kernel and restores the state; *best_ms (may be
 * NULL) receives the summed kernel time of one step at the chosen shapes.
 * Allocates a transient snapshot of the state arrays. */
int fsale_autotune(fsale_ctx *ctx, int32_t nsteps, double *best_ms);

/* ------------------------------------------------------------ serving */
/* Hydro solves arriving as requests (no counterpart in the reference, which
 * runs one problem per job; each request runs the reference's step loop of
 * its solver).  A server holds one Eulerian and one Lagrangian batch, each on
 * its own HIP stream, so the two solvers' work overlaps on the device.  Every
 * scheduling round (fsale_server_poll) admits queued requests in arrival
 * order while they fit the batches' free capacity (continuous batching: a
 * request that does not fit waits; up to `lookahead` later ones may pass it),
 * launches chunk_steps steps of both batches, and retires finished requests
 * after copying their results to host memory.  Mode 1 serves the same
 * requests one at a time, each in a plain context of its own (the
 * per-problem GPU path, for comparison). */
typedef struct fsale_server fsale_server;

typedef struct {
  int32_t mode;              /* 0 continuous batching, 1 one plain context per request  */
  int32_t dim;               /* mesh type of every request                               */
  int32_t verts_per_cell;
  int32_t max_problems;      /* per batch                                                */
  int64_t euler_cells, euler_faces, euler_vertices, euler_neighbours;   /* batch capacity */
  int64_t maire_cells, maire_faces, maire_vertices, maire_neighbours;
  int32_t chunk_steps;       /* steps per scheduling round                               */
  int32_t lookahead;         /* queued requests that may pass a blocked one              */
  int32_t eager;             /* 1: launch each step's kernels instead of replaying graphs */
  int32_t reserved;          /* 0                                                        */
} fsale_server_config;

typedef struct {
  int32_t solver;            /* FSALE_SOLVER_EULER or FSALE_SOLVER_MAIRE                 */
  fsale_mesh_desc mesh;
  fsale_eos eos;             /* Eulerian                                                 */
  fsale_euler_params euler;
  double cv;                 /* Lagrangian: cv and per-cell gamma                        */
  const double *gamma;
  fsale_maire_params maire;
  const double *d, *v, *p;   /* initial density, velocity [num_cells*dim], pressure     */
} fsale_request;

/* Capacities of 0 are sized from the free device memory: frac_hbm of it,
 * split between the two batches in proportion to their bytes per cell. */
int fsale_server_create(const fsale_server_config *config, double frac_hbm, fsale_server **out);
int fsale_server_destroy(fsale_server *server);
/* Seconds on the server's monotonic clock (arrival and completion times). */
double fsale_server_clock(void);
/* Queue a request (inputs are copied); arrival on fsale_server_clock. */
int fsale_server_submit(fsale_server *server, const fsale_request *request, double arrival, int64_t *id);
/* One scheduling round: retire the requests that finished in the round the
 * previous call launched (waits for it), admit queued requests, launch the
 * next round and return while it runs.  *completed receives the number of
 * requests retired, *pending the number queued or running afterwards.
 * Mode 1 serves the oldest queued request start to finish instead. */
int fsale_server_poll(fsale_server *server, int32_t *completed, int32_t *pending);
/* Accounting of a batched server: stats[0..7] = host seconds spent
 * admitting, waiting for rounds, retiring (result download + removal) and
 * launching rounds; rounds run; requests admitted; GPU seconds of the
 * Eulerian and of the Lagrangian batch's rounds (HIP events). */
int fsale_server_stats(fsale_server *server, double *stats);
/* Timing of a finished request: arrival, admission and completion times. */
int fsale_server_times(fsale_server *server, int64_t id, double *arrival, double *admitted,
                       double *finished);
/* Results of a finished request (status, then d, v, e, p and, Lagrangian,
 * the vertex coordinates; any array may be NULL); the server forgets the
 * request afterwards unless keep != 0. */
int fsale_server_result(fsale_server *server, int64_t id, int keep, fsale_problem_status *status,
                        double *d, double *v, double *e, double *p, double *coords);

#ifdef __cplusplus
}
#endif

#endif /* FLECSALE_AMD_H */
