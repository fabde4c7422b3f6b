// Serving loop: hydro solves arrive as requests, are admitted into a resident
// batch of their solver while its capacity allows, advance together with the
// other resident problems, and leave with their results (fsale_server_*).
//
// No counterpart in the reference (one problem per MPI job); each request
// runs the step loop of its solver (apps/hydro/driver.h:317-413,
// apps/maire_hydro/driver.h:330-538) through the batch ABI, so its result is
// the one a plain context gives.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "internal.h"

using fsale::PendingFields;
using fsale::finish_field_downloads;
using fsale::retire_deferred;
using fsale::stage_field_downloads;

namespace {

enum { Q_QUEUED = 0, Q_RUNNING = 1, Q_DONE = 2 };

struct Request {
  int solver = 0;
  int state = Q_QUEUED;
  double arrival = 0, admitted = 0, finished = 0;
  // owned inputs (released on admission)
  std::vector<double> coords, gamma, d, v, p;
  std::vector<int32_t> cells;
  fsale_mesh_desc mesh{};
  fsale_eos eos{};
  fsale_euler_params ep{};
  double cv = 1;
  fsale_maire_params mp{};
  int nc = 0, nv = 0, dim = 0;
  int problem = -1;
  int64_t blocked_epoch = -1;   // capacity epoch of the last failed admission
  fsale_problem_status status{};
  // results
  std::vector<double> rd, rv, re, rp, rx;
};

double now_s() {
  using clk = std::chrono::steady_clock;
  return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

}  // namespace

struct fsale_server {
  fsale_server_config cfg{};
  fsale_ctx *batch[2] = {nullptr, nullptr};   // by solver
  std::map<int64_t, Request> reqs;
  std::deque<int64_t> queue;                  // arrival order
  std::map<int, int64_t> running[2];          // batch problem -> request
  int64_t next_id = 0;
  int64_t epoch = 0;            // retirements so far: capacity can only grow with them
  // a batch retired problems in the last poll: their slots become reusable
  // once the round that was running then has been waited for
  bool retired_last[2] = {false, false};
  // GPU duration of each batch's rounds (events around its launches): the
  // last round's sets the next poll's admission budget
  hipEvent_t ev_start[2] = {}, ev_end[2] = {};
  bool timed[2] = {false, false};
  double gpu_round_s = 0;       // last round (longest batch)
  double gpu_s[2] = {0, 0};     // all rounds, per batch
  // host seconds spent per phase: admission, waiting for a round, result
  // download + retirement, launching rounds; rounds and admissions counted
  double t_admit = 0, t_wait = 0, t_retire = 0, t_launch = 0;
  int64_t rounds = 0, admissions = 0;
};

namespace {

int fail(int rc, const std::string &msg) {
  fsale::set_message(msg);
  return rc;
}

// copies the finished problem's fields out of its batch, then retires it
// Result arrays of a finished request and the batch fields they receive.
int result_targets(Request &R, int solver, int32_t *ids, double **outs) {
  R.rd.resize(R.nc);
  R.rv.resize((size_t)R.nc * R.dim);
  R.re.resize(R.nc);
  R.rp.resize(R.nc);
  ids[0] = FSALE_F_DENSITY;
  ids[1] = FSALE_F_VELOCITY;
  ids[2] = FSALE_F_INTERNAL_ENERGY;
  ids[3] = FSALE_F_PRESSURE;
  outs[0] = R.rd.data();
  outs[1] = R.rv.data();
  outs[2] = R.re.data();
  outs[3] = R.rp.data();
  if (solver != FSALE_SOLVER_MAIRE) return 4;
  R.rx.resize((size_t)R.nv * R.dim);
  ids[4] = FSALE_F_COORDS;
  outs[4] = R.rx.data();
  return 5;
}

// bookkeeping of a retired request (its results are on the host)
void mark_retired(fsale_server *S, int solver, int problem, const fsale_problem_status &st) {
  const int64_t id = S->running[solver][problem];
  Request &R = S->reqs[id];
  R.status = st;
  R.status.problem = -1;
  S->running[solver].erase(problem);
  ++S->epoch;
  R.state = Q_DONE;
  R.finished = now_s();
}

void release_inputs(Request &R) {
  std::vector<double>().swap(R.coords);
  std::vector<double>().swap(R.gamma);
  std::vector<double>().swap(R.d);
  std::vector<double>().swap(R.v);
  std::vector<double>().swap(R.p);
  std::vector<int32_t>().swap(R.cells);
}

// admission into the request's batch; FSALE_ERR_OOM: does not fit now
int admit(fsale_server *S, int64_t id) {
  Request &R = S->reqs[id];
  R.mesh.coords = R.coords.data();
  R.mesh.cells = R.cells.data();
  int32_t prob = -1;
  int rc;
  if (R.solver == FSALE_SOLVER_EULER)
    rc = fsale_batch_add_euler(S->batch[0], &R.mesh, R.eos, &R.ep, R.d.data(), R.v.data(), R.p.data(), &prob);
  else
    rc = fsale_batch_add_maire(S->batch[1], &R.mesh, R.cv, R.gamma.data(), &R.mp, R.d.data(), R.v.data(),
                               R.p.data(), &prob);
  if (rc) return rc;
  R.problem = prob;
  R.state = Q_RUNNING;
  R.admitted = now_s();
  S->running[R.solver][prob] = id;
  release_inputs(R);
  return FSALE_OK;
}

// one request alone in a plain context: the per-problem GPU path
int serve_alone(fsale_server *S, int64_t id) {
  Request &R = S->reqs[id];
  R.mesh.coords = R.coords.data();
  R.mesh.cells = R.cells.data();
  R.admitted = now_s();
  fsale_ctx *ctx = nullptr;
  int rc = fsale_ctx_create(&R.mesh, R.solver, &ctx);
  if (rc) return rc;
  int64_t steps = 0;
  double t = 0, dt = 0;
  if (R.solver == FSALE_SOLVER_EULER) {
    rc = fsale_eul_initial_conditions(ctx, R.eos, R.d.data(), R.v.data(), R.p.data());
    if (!rc)
      rc = fsale_eul_advance(ctx, R.eos, R.ep.cfl, R.ep.final_time, (int)R.ep.max_steps, R.ep.flux_kind,
                             R.ep.first_order, R.ep.stop_at_final, S->cfg.eager ? 0 : 1);
  } else {
    rc = fsale_maire_initial_conditions(ctx, R.cv, R.gamma.data(), R.d.data(), R.v.data(), R.p.data());
    if (!rc) rc = fsale_set_time(ctx, 0.0, R.mp.initial_time_step);
    if (!rc)
      rc = fsale_maire_advance(ctx, R.mp.cfl_accoustic, R.mp.cfl_volume, R.mp.cfl_growth, R.mp.final_time,
                               (int)R.mp.max_steps, S->cfg.eager ? 0 : 1);
  }
  int err = rc ? rc : fsale_sync(ctx);
  if (err == FSALE_ERR_NEG_STATE || err == FSALE_ERR_DT) {
    R.status.error = err;
  } else if (err) {
    fsale_ctx_destroy(ctx);
    return err;
  } else {
    R.rd.resize(R.nc);
    R.rv.resize((size_t)R.nc * R.dim);
    R.re.resize(R.nc);
    R.rp.resize(R.nc);
    if ((rc = fsale_get_field(ctx, FSALE_F_DENSITY, R.rd.data())) ||
        (rc = fsale_get_field(ctx, FSALE_F_VELOCITY, R.rv.data())) ||
        (rc = fsale_get_field(ctx, FSALE_F_INTERNAL_ENERGY, R.re.data())) ||
        (rc = fsale_get_field(ctx, FSALE_F_PRESSURE, R.rp.data()))) {
      fsale_ctx_destroy(ctx);
      return rc;
    }
    if (R.solver == FSALE_SOLVER_MAIRE) {
      R.rx.resize((size_t)R.nv * R.dim);
      if ((rc = fsale_get_field(ctx, FSALE_F_COORDS, R.rx.data()))) {
        fsale_ctx_destroy(ctx);
        return rc;
      }
    }
  }
  (void)fsale_get_time(ctx, &t, &dt, &steps);
  fsale_ctx_destroy(ctx);
  R.status.problem = -1;
  R.status.steps = steps;
  R.status.time = t;
  R.status.time_step = dt;
  R.status.done = 1;
  R.state = Q_DONE;
  R.finished = now_s();
  release_inputs(R);
  return FSALE_OK;
}

}  // namespace

extern "C" {

double fsale_server_clock(void) { return now_s(); }

int fsale_server_create(const fsale_server_config *config, double frac_hbm, fsale_server **out) {
  if (!config || !out) return FSALE_ERR_ARG;
  *out = nullptr;
  fsale_server_config c = *config;
  if (c.mode != 0 && c.mode != 1) return fail(FSALE_ERR_ARG, "server mode is 0 (batched) or 1 (one at a time)");
  if (c.chunk_steps <= 0) c.chunk_steps = 10;
  if (c.lookahead < 0) c.lookahead = 0;
  if (c.max_problems <= 0) c.max_problems = 1024;
  fsale_server *S = new fsale_server();
  if (c.mode == 0) {
    if (c.euler_cells == 0 || c.maire_cells == 0) {
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || !(frac_hbm > 0 && frac_hbm < 1)) {
        delete S;
        return fail(FSALE_ERR_ARG, "batch capacities are 0 and the free device memory is unknown");
      }
      // device bytes per cell of a 2D quad batch (state, gradients or corner
      // arrays, geometry, connectivity; ~2 faces and ~1 vertex per cell)
      const double per_cell[2] = {600.0, 750.0};
      const double budget = frac_hbm * (double)free_b / 2;
      if (c.euler_cells == 0) {
        c.euler_cells = (int64_t)(budget / per_cell[0]);
        c.euler_faces = c.euler_cells * 2 + 4096;
        c.euler_vertices = c.euler_cells + c.euler_cells / 4 + 4096;
        c.euler_neighbours = c.euler_cells * 8;
      }
      if (c.maire_cells == 0) {
        c.maire_cells = (int64_t)(budget / per_cell[1]);
        c.maire_faces = c.maire_cells * 2 + 4096;
        c.maire_vertices = c.maire_cells + c.maire_cells / 4 + 4096;
        c.maire_neighbours = c.maire_cells * 8;
      }
    }
    int rc = fsale_batch_create(FSALE_SOLVER_EULER, c.dim, c.verts_per_cell, c.max_problems, c.euler_cells,
                                c.euler_faces, c.euler_vertices, c.euler_neighbours, &S->batch[0]);
    if (!rc && c.dim == 2 && c.verts_per_cell == 4)
      rc = fsale_batch_create(FSALE_SOLVER_MAIRE, c.dim, c.verts_per_cell, c.max_problems, c.maire_cells,
                              c.maire_faces, c.maire_vertices, c.maire_neighbours, &S->batch[1]);
    for (int s = 0; s < 2 && !rc; ++s)
      if (hipEventCreate(&S->ev_start[s]) != hipSuccess || hipEventCreate(&S->ev_end[s]) != hipSuccess)
        rc = fail(FSALE_ERR_HIP, "hipEventCreate failed");
    if (rc) {
      fsale_server_destroy(S);
      return rc;
    }
  }
  S->cfg = c;
  *out = S;
  return FSALE_OK;
}

int fsale_server_destroy(fsale_server *S) {
  if (!S) return FSALE_ERR_ARG;
  for (fsale_ctx *b : S->batch)
    if (b) fsale_ctx_destroy(b);
  for (int s = 0; s < 2; ++s) {
    if (S->ev_start[s]) (void)hipEventDestroy(S->ev_start[s]);
    if (S->ev_end[s]) (void)hipEventDestroy(S->ev_end[s]);
  }
  delete S;
  return FSALE_OK;
}

int fsale_server_submit(fsale_server *S, const fsale_request *q, double arrival, int64_t *id) {
  if (!S || !q || !id) return FSALE_ERR_ARG;
  const fsale_mesh_desc &m = q->mesh;
  if (q->solver != FSALE_SOLVER_EULER && q->solver != FSALE_SOLVER_MAIRE) return FSALE_ERR_ARG;
  if (m.dim != S->cfg.dim || m.verts_per_cell != S->cfg.verts_per_cell || !m.coords || !m.cells ||
      m.num_cells <= 0 || m.num_vertices <= 0 || !q->d || !q->v || !q->p)
    return fail(FSALE_ERR_ARG, "request mesh or state does not match the server");
  if (q->solver == FSALE_SOLVER_MAIRE && (!q->gamma || (S->cfg.mode == 0 && !S->batch[1])))
    return fail(FSALE_ERR_ARG, "Lagrangian requests need per-cell gamma and a 2D quad server");
  Request R;
  R.solver = q->solver;
  R.arrival = arrival;
  R.nc = m.num_cells;
  R.nv = m.num_vertices;
  R.dim = m.dim;
  R.coords.assign(m.coords, m.coords + (size_t)m.num_vertices * m.dim);
  R.cells.assign(m.cells, m.cells + (size_t)m.num_cells * m.verts_per_cell);
  R.mesh = m;
  R.eos = q->eos;
  R.ep = q->euler;
  R.cv = q->cv;
  R.mp = q->maire;
  if (q->solver == FSALE_SOLVER_MAIRE) R.gamma.assign(q->gamma, q->gamma + R.nc);
  R.d.assign(q->d, q->d + R.nc);
  R.v.assign(q->v, q->v + (size_t)R.nc * m.dim);
  R.p.assign(q->p, q->p + R.nc);
  const int64_t k = S->next_id++;
  S->reqs.emplace(k, std::move(R));
  S->queue.push_back(k);
  *id = k;
  return FSALE_OK;
}

int fsale_server_poll(fsale_server *S, int32_t *completed, int32_t *pending) {
  if (!S) return FSALE_ERR_ARG;
  int done = 0, rc;
  if (S->cfg.mode == 1) {
    // one request, alone, start to finish
    if (!S->queue.empty()) {
      const int64_t id = S->queue.front();
      S->queue.pop_front();
      if ((rc = serve_alone(S, id))) return rc;
      done = 1;
    }
  } else {
    // the round launched by the previous poll has run meanwhile: wait for it
    // and collect what finished in it
    std::vector<fsale_problem_status> st;
    std::vector<fsale_problem_status> fin[2];
    for (int s = 0; s < 2; ++s) {
      if (S->running[s].empty() && !S->retired_last[s]) continue;
      st.resize(S->running[s].size() + 1);
      int32_t n = 0;
      double t0 = now_s();
      if ((rc = fsale_batch_status(S->batch[s], (int32_t)st.size(), st.data(), &n))) return rc;
      S->t_wait += now_s() - t0;
      if (S->retired_last[s]) ++S->epoch;   // their slots are free now
      S->retired_last[s] = false;
      for (int i = 0; i < n; ++i)
        if (st[i].done) fin[s].push_back(st[i]);
    }
    // their results are copied out on each batch's upload stream, behind the
    // finished round and ahead of any admission into their ranges, while the
    // next round already runs: finished problems are skipped by its kernels
    const double tr = now_s();
    std::vector<PendingFields> pend[2];
    for (int s = 0; s < 2; ++s) {
      std::vector<int> probs;
      std::vector<std::array<double *, 5>> outs;
      std::vector<double *const *> outp;
      int32_t ids[5];
      int nf = 0;
      for (const auto &f : fin[s]) {
        if (f.error != 0) continue;
        Request &R = S->reqs[S->running[s][f.problem]];
        std::array<double *, 5> o{};
        nf = result_targets(R, s, ids, o.data());
        probs.push_back(f.problem);
        outs.push_back(o);
      }
      for (auto &o : outs) outp.push_back(o.data());
      if (!probs.empty() &&
          (rc = stage_field_downloads(S->batch[s], probs.data(), (int)probs.size(), nf, ids, outp.data(), pend[s])))
        return rc;
    }
    S->t_retire += now_s() - tr;
    // GPU time of the round just finished
    double round_s = 0;
    for (int s = 0; s < 2; ++s) {
      if (!S->timed[s]) continue;
      float ms = 0;
      if (hipEventElapsedTime(&ms, S->ev_start[s], S->ev_end[s]) == hipSuccess) {
        S->gpu_s[s] += ms * 1e-3;
        round_s = std::max(round_s, ms * 1e-3);
      }
      S->timed[s] = false;
    }
    if (round_s > 0) S->gpu_round_s = round_s;
    // launch the next round of the resident problems: both batches advance
    // concurrently, each on its own stream
    const double tl = now_s();
    bool launched = false;
    for (int s = 0; s < 2; ++s) {
      if (S->running[s].size() <= fin[s].size()) continue;
      hipStream_t stream = S->batch[s]->stream;
      FSALE_HIP(hipEventRecord(S->ev_start[s], stream));
      if ((rc = fsale_batch_step(S->batch[s], S->cfg.chunk_steps, S->cfg.eager ? 0 : 1))) return rc;
      FSALE_HIP(hipEventRecord(S->ev_end[s], stream));
      S->timed[s] = true;
      launched = true;
    }
    S->t_launch += now_s() - tl;
    if (S->running[0].size() > fin[0].size() || S->running[1].size() > fin[1].size()) ++S->rounds;
    // retire: results to the host, ranges released for the admissions below
    const double tr2 = now_s();
    for (int s = 0; s < 2; ++s) {
      if ((rc = finish_field_downloads(S->batch[s], pend[s]))) return rc;
      for (const auto &f : fin[s]) {
        if ((rc = retire_deferred(S->batch[s], f.problem))) return rc;
        mark_retired(S, s, f.problem, f);
        S->retired_last[s] = true;
        ++done;
      }
    }
    S->t_retire += now_s() - tr2;
    // admission while the round runs: arrival order; a request that does not
    // fit blocks, but up to `lookahead` later requests may still be tried
    // behind it.  Admitted requests join at the next round.  The host time
    // spent admitting is held to ~70% of the last round's GPU time (20 ms
    // while nothing runs), so the next round is launched as soon as the GPU
    // is done; problems admitted in one burst would also finish in one and
    // leave the device idle while their successors are admitted.
    const double ta = now_s();
    const double budget = (launched && S->gpu_round_s > 0) ? 0.7 * S->gpu_round_s - (ta - tl) : 0.02;
    int blocked = 0;
    for (auto it = S->queue.begin(); it != S->queue.end();) {
      if (now_s() - ta > budget) break;
      Request &Rq = S->reqs[*it];
      if (Rq.blocked_epoch == S->epoch) {   // nothing was freed since it last failed
        if (++blocked > S->cfg.lookahead) break;
        ++it;
        continue;
      }
      rc = admit(S, *it);
      if (rc == FSALE_OK) {
        ++S->admissions;
        it = S->queue.erase(it);
        continue;
      }
      if (rc != FSALE_ERR_OOM) {
        // not admissible at all (bad mesh or state): finishes with its error
        Request &R = S->reqs[*it];
        R.status.problem = -1;
        R.status.error = rc;
        R.status.done = 1;
        R.state = Q_DONE;
        R.admitted = R.finished = now_s();
        release_inputs(R);
        ++done;
        it = S->queue.erase(it);
        continue;
      }
      Rq.blocked_epoch = S->epoch;
      if (++blocked > S->cfg.lookahead) break;
      ++it;
    }
    S->t_admit += now_s() - ta;
  }
  if (completed) *completed = done;
  if (pending) *pending = (int32_t)(S->queue.size() + S->running[0].size() + S->running[1].size());
  return FSALE_OK;
}

int fsale_server_stats(fsale_server *S, double *stats) {
  if (!S || !stats) return FSALE_ERR_ARG;
  stats[0] = S->t_admit;
  stats[1] = S->t_wait;
  stats[2] = S->t_retire;
  stats[3] = S->t_launch;
  stats[4] = (double)S->rounds;
  stats[5] = (double)S->admissions;
  stats[6] = S->gpu_s[0];
  stats[7] = S->gpu_s[1];
  return FSALE_OK;
}

int fsale_server_times(fsale_server *S, int64_t id, double *arrival, double *admitted, double *finished) {
  if (!S) return FSALE_ERR_ARG;
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(FSALE_ERR_ARG, "unknown request");
  if (arrival) *arrival = it->second.arrival;
  if (admitted) *admitted = it->second.admitted;
  if (finished) *finished = it->second.finished;
  return FSALE_OK;
}

int fsale_server_result(fsale_server *S, int64_t id, int keep, fsale_problem_status *status, double *d, double *v,
                        double *e, double *p, double *coords) {
  if (!S) return FSALE_ERR_ARG;
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(FSALE_ERR_ARG, "unknown request");
  Request &R = it->second;
  if (R.state != Q_DONE) return fail(FSALE_ERR_STATE, "request not finished");
  if (status) *status = R.status;
  if (R.status.error == 0) {
    if (d) std::memcpy(d, R.rd.data(), R.rd.size() * sizeof(double));
    if (v) std::memcpy(v, R.rv.data(), R.rv.size() * sizeof(double));
    if (e) std::memcpy(e, R.re.data(), R.re.size() * sizeof(double));
    if (p) std::memcpy(p, R.rp.data(), R.rp.size() * sizeof(double));
    if (coords && !R.rx.empty()) std::memcpy(coords, R.rx.data(), R.rx.size() * sizeof(double));
  }
  if (!keep) S->reqs.erase(it);
  return FSALE_OK;
}

}  // extern "C"
