// Batched ensemble: many independent suspensions, each with R right-hand sides sharing its
// positions, applied together.
//
// Semantics: for every system k and right-hand side r, V_k[:, :, r] equals mobility! on that
// system's FFCMConfig with forces F_k[:, :, r] (src/mobility.jl:23-45); there is no reference
// function for the ensemble itself.
//
// Design: each system is an ffcm_ctx built for R right-hand sides (cell list, tile binning and
// stencil anchors once per system; the R force grids transformed in one batched FFT). The
// ensemble apply runs in one of two forms, each over 1..4 stream lanes (systems assigned to
// lanes longest-processing-time first by a cost model):
//
//   grouped     every pipeline step of a lane's systems is one launch (group.hpp): the systems'
//               pipelines are recorded, invocations of the same step and kernel are merged into
//               one grouped launch whose blocks are split into per-system ranges, a (system,
//               tile) queue. The FFT passes of all systems that share a plan are one launch
//               each, their systems' resident grids dispatched one after another. Each system
//               holds its own force / velocity grids (one buffer: the solve writes the velocity
//               over the consumed forces) and spectrum. The grouped launches (about 140 per lane)
//               are replayed at every apply, lanes forked and joined with events.
//   per-system  each system's launches in turn on its lane, sharing that lane's buffers (sized
//               for the largest system), captured once into a hipGraph and replayed: for
//               ensembles whose per-system grids do not fit.
//
// The first apply times every candidate (both forms, 1..4 lanes) over two replays and keeps the
// fastest (the autotuner); FFCM_BATCH_MODE=grouped|lanes restricts the candidates to one form,
// FFCM_BATCH_LANES to at most that many lanes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "ffcm_internal.hpp"
#include "group.hpp"

using namespace ffcm;

struct ffcm_batch {
  int K = 0, R = 1, dtype = FFCM_F32;
  std::vector<ffcm_ctx*> sys;
  int lanes_max = 1;
  std::vector<hipStream_t> lane;
  std::vector<void*> lane_force, lane_vel, lane_spec;
  std::vector<hipEvent_t> ev_join;
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;
  // batch-owned device arrays of the host-array entry point
  std::vector<void*> Y_in, F_in, V_out;
  // captured graph and the pointers it was captured with
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> key;
  int lanes_used = 0;  // 0: not tuned yet; l > 0: l lanes of per-system launches; -l: grouped, l lanes
  int64_t graph_nodes = 0;
  double tune_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // per-system 1..4 lanes, grouped 1..4 lanes (0: not timed)
  // grouped execution: per-system grids (force = velocity buffer) and spectra; the grouped
  // launches of the current execution, replayed on their lanes at every apply, and their device
  // tables (each prepared candidate owns its tables; they are freed with it, once the stream
  // is idle)
  std::vector<void*> sys_grid, sys_spec;
  void* group_tables = nullptr;
  struct Direct {
    const void* fn;
    unsigned blocks, threads;
    size_t lds;
    const void* table;
    int n;
    int lane;
  };
  std::vector<Direct> direct;
  bool ready = false;  // the current execution matches `key`
  std::string err;
};

namespace {

thread_local std::string g_batch_create_error;

int bfail(ffcm_batch* b, int code, const std::string& msg) {
  if (b)
    b->err = msg;
  else
    g_batch_create_error = msg;
  return code;
}

int bhip(ffcm_batch* b, hipError_t e, const char* what) {
  (void)hipGetLastError();
  return bfail(b, e == hipErrorOutOfMemory ? FFCM_ERR_MEMORY : FFCM_ERR_DEVICE,
               std::string(what) + ": " + hipGetErrorString(e));
}

#define BHIP(b, call)                                  \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return bhip((b), e_, #call); \
  } while (0)

void destroy_all(ffcm_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  for (auto* v : {&b->sys_grid, &b->sys_spec})
    for (void* p : *v)
      if (p) (void)hipFree(p);
  if (b->group_tables) (void)hipFree(b->group_tables);
  for (ffcm_ctx* c : b->sys)
    if (c) ffcm_destroy(c);
  for (void* p : b->lane_force)
    if (p) (void)hipFree(p);
  for (void* p : b->lane_vel)
    if (p) (void)hipFree(p);
  for (void* p : b->lane_spec)
    if (p) (void)hipFree(p);
  for (auto* v : {&b->Y_in, &b->F_in, &b->V_out})
    for (void* p : *v)
      if (p) (void)hipFree(p);
  for (hipEvent_t e : b->ev_join)
    if (e) (void)hipEventDestroy(e);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_t0) (void)hipEventDestroy(b->ev_t0);
  if (b->ev_t1) (void)hipEventDestroy(b->ev_t1);
  for (hipStream_t s : b->lane)
    if (s) (void)hipStreamDestroy(s);
}

// Relative cost of one system's apply: the FFT traffic (grid points times log) dominates for
// large grids, the particle stages for small ones.
double cost_of(const ffcm_ctx* c, int R) {
  const double M3 = (double)c->M3;
  return R * (M3 * (1.0 + std::log2(std::max(2.0, std::cbrt(M3))) / 8.0) + 40.0 * (double)c->N) + 2.0e5;
}

// Longest-processing-time assignment of systems to `lanes` lanes; returns per-lane lists.
std::vector<std::vector<int>> assign(const ffcm_batch* b, int lanes) {
  std::vector<int> order(b->K);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return cost_of(b->sys[x], b->R) > cost_of(b->sys[y], b->R); });
  std::vector<double> load(lanes, 0.0);
  std::vector<std::vector<int>> out(lanes);
  for (int k : order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    out[l].push_back(k);
    load[l] += cost_of(b->sys[k], b->R);
  }
  return out;
}

// Captures the ensemble apply with `lanes` lanes into an executable graph.
int capture(ffcm_batch* b, int lanes, void* const* V, const void* const* Y, const void* const* F,
            hipGraphExec_t* exec_out, int64_t* nodes) {
  const auto plan = assign(b, lanes);
  hipStream_t s0 = b->lane[0];
  BHIP(b, hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
  auto abort_capture = [&](int rc) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s0, &g);
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return rc;
  };
  hipError_t e = hipEventRecord(b->ev_fork, s0);
  for (int l = 1; l < lanes && e == hipSuccess; ++l) e = hipStreamWaitEvent(b->lane[l], b->ev_fork, 0);
  if (e != hipSuccess) return abort_capture(bhip(b, e, "graph fork"));
  for (int l = 0; l < lanes; ++l) {
    for (int k : plan[l]) {
      ffcm_ctx* c = b->sys[k];
      c->stream = b->lane[l];
      c->d.force = b->lane_force[l];
      c->d.velocity = b->lane_vel[l];
      c->d.spectrum = b->lane_spec[l];
      if (int rc = enqueue_mobility(c, V[k], Y[k], F[k], b->R))
        return abort_capture(bfail(b, rc, "system " + std::to_string(k) + ": " + c->err));
    }
  }
  for (int l = 1; l < lanes && e == hipSuccess; ++l) {
    e = hipEventRecord(b->ev_join[l], b->lane[l]);
    if (e == hipSuccess) e = hipStreamWaitEvent(s0, b->ev_join[l], 0);
  }
  if (e != hipSuccess) return abort_capture(bhip(b, e, "graph join"));
  hipGraph_t graph = nullptr;
  BHIP(b, hipStreamEndCapture(s0, &graph));
  size_t n = 0;
  (void)hipGraphGetNodes(graph, nullptr, &n);
  *nodes = (int64_t)n;
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return bhip(b, e, "hipGraphInstantiate");
  *exec_out = exec;
  return FFCM_OK;
}

// One grouped launch: the invocations of one step and kernel, one per system.
struct Group {
  int tag;
  const void* kernel;
  const void* group;
  unsigned threads;
  size_t lds = 0;
  std::vector<const LaunchRecord*> recs;
  unsigned blocks = 0;
  size_t table_offset = 0;
  std::vector<unsigned char> table;
};

// Block ranges of a group's entries: each starts on a multiple of 8 (blocks keep their XCD) and
// keeps its own grid. Entries follow each other in block order, so the blocks of a persistent
// kernel's systems are dispatched system after system, each with its full resident grid.
void layout_group(Group* g) {
  const size_t esz = g->recs[0]->entry_size;
  g->table.assign(esz * g->recs.size(), 0);
  int64_t b0 = 0;
  for (size_t e = 0; e < g->recs.size(); ++e) {
    const LaunchRecord* r = g->recs[e];
    const int b0i = (int)b0, nbi = (int)r->grid;
    unsigned char* ent = g->table.data() + e * esz;
    std::memcpy(ent, &b0i, sizeof(int));
    std::memcpy(ent + sizeof(int), &nbi, sizeof(int));
    std::memcpy(ent + r->args_offset, r->args.data(), r->args.size());
    g->lds = std::max(g->lds, r->lds);
    b0 += ((int64_t)r->grid + 7) / 8 * 8;
  }
  g->blocks = (unsigned)b0;
}

// Records every system's pipeline on its own grids and merges each lane's invocations into
// grouped launches (the systems split over `lanes` lanes as in capture), with their tables.
int prepare_grouped(ffcm_batch* b, int lanes, void* const* V, const void* const* Y, const void* const* F,
                    int64_t* launches, void** tables_out, std::vector<ffcm_batch::Direct>* direct_out) {
  const auto plan = assign(b, lanes);
  std::vector<Recorder> recs(b->K);
  for (int k = 0; k < b->K; ++k) {
    ffcm_ctx* c = b->sys[k];
    c->d.force = b->sys_grid[k];
    c->d.velocity = b->sys_grid[k];
    c->d.spectrum = b->sys_spec[k];
    c->rec = &recs[k];
    const int rc = enqueue_mobility(c, V[k], Y[k], F[k], b->R);
    c->rec = nullptr;
    if (rc) return bfail(b, rc, "system " + std::to_string(k) + ": " + c->err);
  }
  // per lane: merge by (tag, kernel, block size), groups ordered by tag (each system's tags
  // increase along its pipeline, so every system's steps stay in order)
  std::vector<std::vector<Group>> lane_groups(lanes);
  for (int l = 0; l < lanes; ++l) {
    std::vector<Group>& groups = lane_groups[l];
    for (int k : plan[l])
      for (const LaunchRecord& r : recs[k].launches) {
        if (r.grid == 0) continue;
        Group* g = nullptr;
        for (Group& q : groups)
          if (q.tag == r.tag && q.kernel == r.kernel && q.threads == r.threads) {
            g = &q;
            break;
          }
        if (!g) {
          groups.push_back(Group{r.tag, r.kernel, r.group, r.threads});
          g = &groups.back();
        }
        g->recs.push_back(&r);
      }
    std::stable_sort(groups.begin(), groups.end(), [](const Group& x, const Group& y) { return x.tag < y.tag; });
  }
  size_t total = 0;
  for (auto& groups : lane_groups)
    for (Group& g : groups) {
      layout_group(&g);
      g.table_offset = total;
      total += (g.table.size() + 255) / 256 * 256;
    }
  // this graph's own tables (another candidate's graph may still be replayed with its own)
  unsigned char* tables = nullptr;
  BHIP(b, hipMalloc((void**)&tables, total));
  *tables_out = tables;  // the caller frees it, also on failure below
  std::vector<unsigned char> host(total, 0);
  for (const auto& groups : lane_groups)
    for (const Group& g : groups) std::memcpy(host.data() + g.table_offset, g.table.data(), g.table.size());
  BHIP(b, hipMemcpy(tables, host.data(), total, hipMemcpyHostToDevice));
  std::vector<ffcm_batch::Direct>& direct = *direct_out;
  direct.clear();
  for (int l = 0; l < lanes; ++l)
    for (const Group& g : lane_groups[l])
      direct.push_back({g.group, g.blocks, g.threads, g.lds, tables + g.table_offset, (int)g.recs.size(), l});
  *launches = (int64_t)direct.size();
  return FFCM_OK;
}

// One grouped apply: each lane's grouped launches on its stream, forked from and joined into
// lane 0 with events (no graph: captures of grouped launches have ended in host segmentation
// faults inside the runtime, whose cause was not found; the launches are few).
int replay_grouped(ffcm_batch* b, const std::vector<ffcm_batch::Direct>& direct, int lanes) {
  hipStream_t s0 = b->lane[0];
  if (lanes > 1) {
    BHIP(b, hipEventRecord(b->ev_fork, s0));
    for (int l = 1; l < lanes; ++l) BHIP(b, hipStreamWaitEvent(b->lane[l], b->ev_fork, 0));
  }
  for (const ffcm_batch::Direct& d : direct) {
    const void* table = d.table;
    int n = d.n;
    void* args[] = {(void*)&table, (void*)&n};
    BHIP(b, hipLaunchKernel(d.fn, dim3(d.blocks), dim3(d.threads), args, d.lds, b->lane[d.lane]));
  }
  for (int l = 1; l < lanes; ++l) {
    BHIP(b, hipEventRecord(b->ev_join[l], b->lane[l]));
    BHIP(b, hipStreamWaitEvent(s0, b->ev_join[l], 0));
  }
  return FFCM_OK;
}

int time_grouped(ffcm_batch* b, const std::vector<ffcm_batch::Direct>& direct, int lanes, double* ms) {
  hipStream_t s0 = b->lane[0];
  if (int rc = replay_grouped(b, direct, lanes)) return rc;  // warm-up replay
  BHIP(b, hipEventRecord(b->ev_t0, s0));
  const int reps = 2;
  for (int i = 0; i < reps; ++i)
    if (int rc = replay_grouped(b, direct, lanes)) return rc;
  BHIP(b, hipEventRecord(b->ev_t1, s0));
  BHIP(b, hipEventSynchronize(b->ev_t1));
  float t = 0;
  BHIP(b, hipEventElapsedTime(&t, b->ev_t0, b->ev_t1));
  *ms = t / reps;
  return FFCM_OK;
}

int time_graph(ffcm_batch* b, hipGraphExec_t exec, double* ms) {
  hipStream_t s0 = b->lane[0];
  BHIP(b, hipGraphLaunch(exec, s0));  // warm-up replay
  BHIP(b, hipEventRecord(b->ev_t0, s0));
  const int reps = 2;
  for (int i = 0; i < reps; ++i) BHIP(b, hipGraphLaunch(exec, s0));
  BHIP(b, hipEventRecord(b->ev_t1, s0));
  BHIP(b, hipEventSynchronize(b->ev_t1));
  float t = 0;
  BHIP(b, hipEventElapsedTime(&t, b->ev_t0, b->ev_t1));
  *ms = t / reps;
  return FFCM_OK;
}

// First apply: capture every lane-count candidate, time two replays of each and keep the
// fastest. Later applies with other pointers re-capture with the chosen lane count.
int ensure_graph(ffcm_batch* b, void* const* V, const void* const* Y, const void* const* F) {
  std::vector<const void*> key;
  for (int k = 0; k < b->K; ++k) {
    key.push_back(V[k]);
    key.push_back(Y[k]);
    key.push_back(F[k]);
  }
  if (b->ready && key == b->key) return FFCM_OK;
  if (b->ready) {  // the previous execution and its tables, once its last apply has finished
    BHIP(b, hipStreamSynchronize(b->lane[0]));
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    if (b->group_tables) (void)hipFree(b->group_tables);
    b->group_tables = nullptr;
    b->direct.clear();
    b->ready = false;
  }
  if (b->lanes_used == 0) {
    // Every kernel a captured candidate uses is first launched outside any capture: each
    // system's pipeline once on lane 0 with that lane's buffers. Stream captures then only take
    // kernels that have run before (a first apply whose first launches were all inside
    // captures ended in a host segmentation fault inside the runtime).
    for (int k = 0; k < b->K; ++k) {
      ffcm_ctx* c = b->sys[k];
      c->stream = b->lane[0];
      c->d.force = b->lane_force[0];
      c->d.velocity = b->lane_vel[0];
      c->d.spectrum = b->lane_spec[0];
      if (int rc = enqueue_mobility(c, V[k], Y[k], F[k], b->R))
        return bfail(b, rc, "system " + std::to_string(k) + ": " + c->err);
    }
    BHIP(b, hipStreamSynchronize(b->lane[0]));
    double best = 1e300;
    int best_l = 1;
    hipGraphExec_t best_exec = nullptr;
    void* best_tables = nullptr;  // the grouped tables of best_exec
    std::vector<ffcm_batch::Direct> best_direct;
    int64_t best_nodes = 0;
    // (timing leaves the stream idle, so a candidate's graph or tables can go at once)
    auto release = [](hipGraphExec_t ex, void* tables) {
      if (ex) (void)hipGraphExecDestroy(ex);
      if (tables) (void)hipFree(tables);
    };
    const char* mode = std::getenv("FFCM_BATCH_MODE");
    const bool try_grouped = !b->sys_grid.empty() && !(mode && std::strcmp(mode, "lanes") == 0);
    const bool try_lanes = !(try_grouped && mode && std::strcmp(mode, "grouped") == 0);
    // candidates: -l grouped launches on l lanes, +l per-system launches on l lanes
    // FFCM_BATCH_LANES caps the lanes (1 keeps every candidate on one stream: kernel-trace
    // profiling of multi-stream captures has crashed inside the profiler)
    int lmax = std::min(b->lanes_max, 4);
    if (const char* e = std::getenv("FFCM_BATCH_LANES")) lmax = std::max(1, std::min(lmax, std::atoi(e)));
    std::vector<int> cands;
    for (int l = 1; l <= lmax; ++l) {
      if (l > 1 && l > b->K) break;
      if (try_grouped) cands.push_back(-l);
      if (try_lanes) cands.push_back(l);
    }
    for (int cand : cands) {
      const int l = cand < 0 ? -cand : cand;
      hipGraphExec_t ex = nullptr;
      void* tables = nullptr;
      std::vector<ffcm_batch::Direct> direct;
      int64_t nodes = 0;
      int rc = cand < 0 ? prepare_grouped(b, l, V, Y, F, &nodes, &tables, &direct)
                        : capture(b, l, V, Y, F, &ex, &nodes);
      double ms = 0;
      if (rc == FFCM_OK) rc = cand < 0 ? time_grouped(b, direct, l, &ms) : time_graph(b, ex, &ms);
      if (rc != FFCM_OK) {
        (void)hipStreamSynchronize(b->lane[0]);
        release(ex, tables);
        release(best_exec, best_tables);
        return rc;
      }
      b->tune_ms[cand < 0 ? 3 + l : l - 1] = ms;
      if (ms < best) {
        release(best_exec, best_tables);
        best = ms;
        best_l = cand;
        best_exec = ex;
        best_tables = tables;
        best_direct = std::move(direct);
        best_nodes = nodes;
      } else {
        release(ex, tables);
      }
    }
    b->lanes_used = best_l;
    b->exec = best_exec;
    b->group_tables = best_tables;
    b->direct = std::move(best_direct);
    b->graph_nodes = best_nodes;
  } else if (b->lanes_used < 0) {
    const int rc = prepare_grouped(b, -b->lanes_used, V, Y, F, &b->graph_nodes, &b->group_tables, &b->direct);
    if (rc != FFCM_OK) {  // (nothing was launched with these tables)
      if (b->group_tables) (void)hipFree(b->group_tables);
      b->group_tables = nullptr;
      b->direct.clear();
      return rc;
    }
  } else {
    if (int rc = capture(b, b->lanes_used, V, Y, F, &b->exec, &b->graph_nodes)) return rc;
  }
  b->key = key;
  b->ready = true;
  return FFCM_OK;
}

}  // namespace

extern "C" {

int ffcm_batch_create(ffcm_batch** out, int n_systems, const ffcm_params* params, int n_rhs) {
  g_batch_create_error.clear();
  if (!out) return bfail(nullptr, FFCM_ERR_ARGUMENT, "out must not be NULL");
  *out = nullptr;
  if (n_systems < 1 || !params) return bfail(nullptr, FFCM_ERR_ARGUMENT, "n_systems must be >= 1 with params");
  if (n_rhs < 1 || n_rhs > 4096) return bfail(nullptr, FFCM_ERR_ARGUMENT, "n_rhs must be between 1 and 4096");
  for (int k = 1; k < n_systems; ++k)
    if (params[k].dtype != params[0].dtype)
      return bfail(nullptr, FFCM_ERR_ARGUMENT, "every system of a batch must use the same dtype");
  auto* b = new ffcm_batch();
  b->K = n_systems;
  b->R = n_rhs;
  b->dtype = params[0].dtype;
  b->sys.assign(n_systems, nullptr);
  int rc = FFCM_OK;
  size_t gmax = 0, smax = 0;
  for (int k = 0; k < n_systems && rc == FFCM_OK; ++k) {
    rc = create_context(&b->sys[k], &params[k], n_rhs, false);
    if (rc != FFCM_OK) {
      g_batch_create_error = "system " + std::to_string(k) + ": " + ffcm_last_error(nullptr);
      break;
    }
    ffcm_ctx* c = b->sys[k];
    (void)hipStreamDestroy(c->stream);  // the batch's lanes drive every system
    c->stream = nullptr;
    c->borrowed_stream = true;
    gmax = std::max(gmax, grid_bytes(c));
    smax = std::max(smax, spectrum_bytes(c));
  }
  if (rc != FFCM_OK) {
    destroy_all(b);
    delete b;
    return rc;
  }
  // lanes: up to 4, each holding R force + velocity grids and spectra of the largest system,
  // within half of the free device memory
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  (void)hipGetLastError();
  const size_t per_lane = (size_t)n_rhs * (2 * gmax + smax);
  int lanes = 4;
  while (lanes > 1 && (size_t)lanes * per_lane > free_b / 2) --lanes;
  lanes = std::min(lanes, n_systems);
  b->lanes_max = lanes;
  auto cleanup = [&](int code) {
    destroy_all(b);
    delete b;
    return code;
  };
  b->lane.assign(lanes, nullptr);
  b->ev_join.assign(lanes, nullptr);
  b->lane_force.assign(lanes, nullptr);
  b->lane_vel.assign(lanes, nullptr);
  b->lane_spec.assign(lanes, nullptr);
  for (int l = 0; l < lanes; ++l) {
    hipError_t e = hipStreamCreateWithFlags(&b->lane[l], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_join[l], hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&b->lane_force[l], (size_t)n_rhs * gmax);
    if (e == hipSuccess) e = hipMalloc(&b->lane_vel[l], (size_t)n_rhs * gmax);
    if (e == hipSuccess) e = hipMalloc(&b->lane_spec[l], (size_t)n_rhs * smax);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      g_batch_create_error = "batch lane buffers: " + std::string(hipGetErrorString(e));
      return cleanup(e == hipErrorOutOfMemory ? FFCM_ERR_MEMORY : FFCM_ERR_DEVICE);
    }
  }
  // grouped execution: every system's own R grids (force and velocity share one) and spectra,
  // if they fit in what the lanes leave free (with a tenth of it to spare); else lanes only
  {
    size_t need = 0;
    for (ffcm_ctx* c : b->sys) need += (size_t)n_rhs * (grid_bytes(c) + spectrum_bytes(c));
    size_t free_now = 0, total_now = 0;
    if (hipMemGetInfo(&free_now, &total_now) != hipSuccess) free_now = 0;
    (void)hipGetLastError();
    if (need + free_now / 10 < free_now) {
      b->sys_grid.assign(n_systems, nullptr);
      b->sys_spec.assign(n_systems, nullptr);
      bool ok = true;
      for (int k = 0; k < n_systems && ok; ++k) {
        ok = hipMalloc(&b->sys_grid[k], (size_t)n_rhs * grid_bytes(b->sys[k])) == hipSuccess &&
             hipMalloc(&b->sys_spec[k], (size_t)n_rhs * spectrum_bytes(b->sys[k])) == hipSuccess;
      }
      if (!ok) {
        (void)hipGetLastError();
        for (auto* v : {&b->sys_grid, &b->sys_spec}) {
          for (void* p : *v)
            if (p) (void)hipFree(p);
          v->clear();
        }
      }
    }
  }
  hipError_t e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&b->ev_t0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev_t1);
  const size_t ts = b->dtype == FFCM_F32 ? 4 : 8;
  b->Y_in.assign(n_systems, nullptr);
  b->F_in.assign(n_systems, nullptr);
  b->V_out.assign(n_systems, nullptr);
  for (int k = 0; k < n_systems && e == hipSuccess; ++k) {
    const size_t n3 = 3 * (size_t)b->sys[k]->N * ts;
    e = hipMalloc(&b->Y_in[k], n3);
    if (e == hipSuccess) e = hipMalloc(&b->F_in[k], n3 * n_rhs);
    if (e == hipSuccess) e = hipMalloc(&b->V_out[k], n3 * n_rhs);
  }
  // HIP loads the library's device code at its first kernel launch in the process. Nothing above
  // launches a kernel, and the first apply's launches are all inside stream captures; one small
  // kernel here (clearing the first system's tile counts, which every apply rewrites) makes that
  // first launch happen outside any capture.
  if (e == hipSuccess) e = launch_fill_i32(b->lane[0], b->sys[0]->d.tile_count, b->sys[0]->tiling.ntiles, 0);
  if (e == hipSuccess) e = hipStreamSynchronize(b->lane[0]);
  if (e != hipSuccess) {
    g_batch_create_error = "batch buffers: " + std::string(hipGetErrorString(e));
    (void)hipGetLastError();
    return cleanup(e == hipErrorOutOfMemory ? FFCM_ERR_MEMORY : FFCM_ERR_DEVICE);
  }
  *out = b;
  return FFCM_OK;
}

int ffcm_batch_destroy(ffcm_batch* b) {
  if (!b) return FFCM_OK;
  if (!b->lane.empty() && b->lane[0]) (void)hipStreamSynchronize(b->lane[0]);
  destroy_all(b);
  delete b;
  return FFCM_OK;
}

const char* ffcm_batch_last_error(const ffcm_batch* b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int ffcm_batch_mobility_device(ffcm_batch* b, void* const* V, const void* const* Y, const void* const* F) {
  if (!b) return FFCM_ERR_ARGUMENT;
  b->err.clear();
  (void)hipGetLastError();
  if (!V || !Y || !F) return bfail(b, FFCM_ERR_ARGUMENT, "V, Y and F must not be NULL");
  for (int k = 0; k < b->K; ++k)
    if (!V[k] || !Y[k] || !F[k]) return bfail(b, FFCM_ERR_ARGUMENT, "system " + std::to_string(k) + ": NULL array");
  if (int rc = ensure_graph(b, V, Y, F)) return rc;
  if (b->lanes_used < 0) return replay_grouped(b, b->direct, -b->lanes_used);
  BHIP(b, hipGraphLaunch(b->exec, b->lane[0]));
  return FFCM_OK;
}

int ffcm_batch_mobility(ffcm_batch* b, void* const* V, const void* const* Y, const void* const* F) {
  if (!b) return FFCM_ERR_ARGUMENT;
  b->err.clear();
  if (!V || !Y || !F) return bfail(b, FFCM_ERR_ARGUMENT, "V, Y and F must not be NULL");
  const size_t ts = b->dtype == FFCM_F32 ? 4 : 8;
  hipStream_t s0 = b->lane[0];
  for (int k = 0; k < b->K; ++k) {
    if (!V[k] || !Y[k] || !F[k]) return bfail(b, FFCM_ERR_ARGUMENT, "system " + std::to_string(k) + ": NULL array");
    const size_t n3 = 3 * (size_t)b->sys[k]->N * ts;
    BHIP(b, hipMemcpyAsync(b->Y_in[k], Y[k], n3, hipMemcpyHostToDevice, s0));
    BHIP(b, hipMemcpyAsync(b->F_in[k], F[k], n3 * b->R, hipMemcpyHostToDevice, s0));
  }
  BHIP(b, hipStreamSynchronize(s0));
  if (int rc = ffcm_batch_mobility_device(b, b->V_out.data(), (const void* const*)b->Y_in.data(),
                                          (const void* const*)b->F_in.data()))
    return rc;
  for (int k = 0; k < b->K; ++k) {
    const size_t n3 = 3 * (size_t)b->sys[k]->N * ts;
    BHIP(b, hipMemcpyAsync(V[k], b->V_out[k], n3 * b->R, hipMemcpyDeviceToHost, s0));
  }
  BHIP(b, hipStreamSynchronize(s0));
  return FFCM_OK;
}

int ffcm_batch_get_stream(ffcm_batch* b, void** stream) {
  if (!b || !stream) return FFCM_ERR_ARGUMENT;
  *stream = (void*)b->lane[0];
  return FFCM_OK;
}

int ffcm_batch_synchronize(ffcm_batch* b) {
  if (!b) return FFCM_ERR_ARGUMENT;
  BHIP(b, hipStreamSynchronize(b->lane[0]));
  return FFCM_OK;
}

int ffcm_batch_get_info(const ffcm_batch* b, int32_t* lanes, int64_t* graph_nodes, double* tune_ms) {
  if (!b) return FFCM_ERR_ARGUMENT;
  if (lanes) lanes[0] = b->lanes_used, lanes[1] = b->lanes_max;
  if (graph_nodes) *graph_nodes = b->graph_nodes;
  if (tune_ms)
    for (int i = 0; i < 8; ++i) tune_ms[i] = b->tune_ms[i];
  return FFCM_OK;
}

}  // extern "C"
