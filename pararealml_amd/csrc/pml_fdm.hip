// Host driver of the fused stepping kernels: variant lookup, work-unit
// tables, the ping-pong step loop and its hipGraph capture.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <mutex>
#include <string>
#include <vector>

#include "pml_fdm_kernels.h"

namespace pml {

static const std::vector<Variant> *find_variants(int eq, int integ, int prec, int ndim,
                                                 int *n_fields) {
  using Table = const std::vector<Variant> *(*)(int, int, int, int *);
  static const Table tables[3][5] = {
      {fdm_table_f32_a, fdm_table_f32_b, fdm_table_f32_c, fdm_table_f32_d, fdm_table_f32_e},
      {fdm_table_f64_a, fdm_table_f64_b, fdm_table_f64_c, fdm_table_f64_d, fdm_table_f64_e},
      {fdm_table_f32c_a, fdm_table_f32c_b, fdm_table_f32c_c, fdm_table_f32c_d,
       fdm_table_f32c_e},
  };
  int row;
  switch (prec) {
    case PML_F32: row = 0; break;
    case PML_F64: row = 1; break;
    case PML_F32C: row = 2; break;
    default: return nullptr;
  }
  for (Table t : tables[row]) {
    const std::vector<Variant> *v = t(eq, integ, ndim, n_fields);
    if (v) return v;
  }
  return nullptr;
}

// Device copies of work-unit and grid tables.  They are derived from the
// caller's grid array and the variant's unit shape alone, so the cache key
// is that array (plus unit shape, dimensionality and device) and the unit
// list is only built on a miss.  Tables live until pml_graph_cache_clear().
struct TableSet {
  void *units = nullptr;
  void *grids = nullptr;
  int n_units = 0;
};

static std::mutex g_mu;
static std::unordered_map<std::string, TableSet> g_tables;

// Instantiated step-loop graphs.  A graph records its kernels' arguments by
// value (buffer pointers, dt, coefficients, tables), so the key holds all of
// them; a hit replays the graph without capturing anything.  At most
// kMaxGraphs are kept; the oldest is destroyed after a device synchronise
// (a replay may still be in flight on any stream).
struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int device = 0;
  uint64_t stamp = 0;
};
static std::unordered_map<std::string, GraphEntry> g_graphs;
static uint64_t g_stamp = 0;
static constexpr size_t kMaxGraphs = 64;

static std::string grid_key(const pml_grid *grids, int n_grids, const Variant &V, int ndim,
                            int dev) {
  std::string key(reinterpret_cast<const char *>(grids), sizeof(pml_grid) * n_grids);
  const int extra[4] = {V.ox, V.oy, ndim, dev};
  key.append(reinterpret_cast<const char *>(extra), sizeof(extra));
  return key;
}

static void destroy_entry(GraphEntry &e) {
  if (e.exec) hipGraphExecDestroy(e.exec);
  if (e.graph) hipGraphDestroy(e.graph);
  e.exec = nullptr;
  e.graph = nullptr;
}

// Caller holds g_mu.
static int evict_oldest_graph() {
  auto oldest = g_graphs.end();
  for (auto it = g_graphs.begin(); it != g_graphs.end(); ++it)
    if (oldest == g_graphs.end() || it->second.stamp < oldest->second.stamp) oldest = it;
  if (oldest == g_graphs.end()) return PML_OK;
  int cur = 0;
  PML_HIP_CHECK(hipGetDevice(&cur));
  PML_HIP_CHECK(hipSetDevice(oldest->second.device));
  hipError_t e = hipDeviceSynchronize();
  hipSetDevice(cur);
  if (e != hipSuccess)
    return fail(PML_E_HIP, std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
  destroy_entry(oldest->second);
  g_graphs.erase(oldest);
  return PML_OK;
}

}  // namespace pml

using namespace pml;

extern "C" int pml_graph_cache_clear(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  int cur = 0;
  PML_HIP_CHECK(hipGetDevice(&cur));
  int count = 0;
  PML_HIP_CHECK(hipGetDeviceCount(&count));
  // Everything in flight on any device may still read the tables or run a
  // graph: wait for all of it before releasing them.
  for (int d = 0; d < count; ++d) {
    PML_HIP_CHECK(hipSetDevice(d));
    PML_HIP_CHECK(hipDeviceSynchronize());
  }
  PML_HIP_CHECK(hipSetDevice(cur));
  for (auto &kv : g_graphs) destroy_entry(kv.second);
  g_graphs.clear();
  for (auto &kv : g_tables) {
    if (kv.second.units) hipFree(kv.second.units);
    if (kv.second.grids) hipFree(kv.second.grids);
  }
  g_tables.clear();
  return PML_OK;
}

namespace pml {

static int device_tables(const pml_grid *grids, int n_grids, const Variant &V, int ndim,
                         TableSet *out) {
  int dev = 0;
  PML_HIP_CHECK(hipGetDevice(&dev));
  const std::string key = grid_key(grids, n_grids, V, ndim, dev);
  std::lock_guard<std::mutex> lock(g_mu);
  auto it = g_tables.find(key);
  if (it != g_tables.end()) {
    *out = it->second;
    return PML_OK;
  }
  std::vector<GridDev> gd(n_grids);
  std::vector<TileDev> units;
  for (int g = 0; g < n_grids; ++g) {
    const pml_grid &s = grids[g];
    GridDev &d = gd[g];
    d.n0 = s.n0;
    d.n1 = s.n1;
    d.off = s.state_offset;
    d.bc = s.bc_offset;
    d.inv_dx0sq = 1.0 / (s.dx0 * s.dx0);
    d.inv_dx1sq = 1.0 / (s.dx1 * s.dx1);
    d.inv_2dx0 = 1.0 / (2.0 * s.dx0);
    d.inv_2dx1 = 1.0 / (2.0 * s.dx1);
    d.m2dx0 = -2.0 * s.dx0;
    d.m2dx1 = -2.0 * s.dx1;
    d.p2dx0 = 2.0 * s.dx0;
    d.p2dx1 = 2.0 * s.dx1;
    for (int i0 = 0; i0 < s.n0; i0 += V.oy)
      for (int j0 = 0; j0 < s.n1; j0 += V.ox) units.push_back(TileDev{g, i0, j0, 0});
  }
  if (units.size() > (1u << 30)) return fail(PML_E_INVALID, "batch too large");
  TableSet t;
  PML_HIP_CHECK(hipMalloc(&t.units, std::max<size_t>(units.size() * sizeof(TileDev), 16)));
  PML_HIP_CHECK(hipMalloc(&t.grids, std::max<size_t>(gd.size() * sizeof(GridDev), 16)));
  // Synchronous copies: the host vectors are temporaries.
  PML_HIP_CHECK(hipMemcpy(t.units, units.data(), units.size() * sizeof(TileDev),
                          hipMemcpyHostToDevice));
  PML_HIP_CHECK(hipMemcpy(t.grids, gd.data(), gd.size() * sizeof(GridDev),
                          hipMemcpyHostToDevice));
  t.n_units = static_cast<int>(units.size());
  g_tables.emplace(key, t);
  *out = t;
  return PML_OK;
}

}  // namespace pml

using namespace pml;

extern "C" int pml_fdm_configs(int eq_id, int integrator_id, int precision) {
  int nf = 0;
  const auto *v = find_variants(eq_id, integrator_id, precision, 2, &nf);
  return v ? static_cast<int>(v->size()) : 0;
}

extern "C" const char *pml_fdm_config_name(int eq_id, int integrator_id, int precision,
                                           int config) {
  int nf = 0;
  const auto *v = find_variants(eq_id, integrator_id, precision, 2, &nf);
  if (!v || config < 0 || config >= static_cast<int>(v->size())) return "";
  return (*v)[config].name;
}

// The work-unit and grid tables of the allocation-free entry point live in
// this file-scope device array (one copy per GPU, held for the life of the
// process).  Host code never passes the array's name to a kernel: it uses
// the device address from hipGetSymbolAddress, looked up once per device.
static __device__ __align__(256) unsigned char scratch[PML_FDM_SCRATCH_BYTES];

namespace pml {

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64] = {};

// The device address of `scratch` on the current device (looked up on the
// first call made with that device current).
static int scratch_address(void **out) {
  int dev = 0;
  PML_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(PML_E_INVALID, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    PML_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(PML_E_HIP, "hipGetSymbolAddress returned no address");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return PML_OK;
}

// Builds the tables on the host (reusing this thread's buffer) and copies
// them into `scratch` on the stream: grid table first, then the units at
// the next 256-byte boundary.  Nothing is allocated on the device.
static int scratch_tables(const pml_grid *grids, int n_grids, const Variant &V,
                          hipStream_t st, TableSet *out) {
  thread_local std::vector<unsigned char> host;
  const size_t grid_bytes = sizeof(GridDev) * static_cast<size_t>(n_grids);
  const size_t unit_off = (grid_bytes + 255) / 256 * 256;
  size_t n_units = 0;
  for (int g = 0; g < n_grids; ++g)
    n_units += static_cast<size_t>((grids[g].n0 + V.oy - 1) / V.oy) *
               static_cast<size_t>((grids[g].n1 + V.ox - 1) / V.ox);
  const size_t need = unit_off + sizeof(TileDev) * n_units;
  if (need > PML_FDM_SCRATCH_BYTES)
    return fail(PML_E_INVALID, "the batch needs " + std::to_string(need) +
                                   " bytes of work-unit tables; pml_fdm_solve_noalloc holds at "
                                   "most " + std::to_string(PML_FDM_SCRATCH_BYTES));
  void *base = nullptr;
  if (int rc = scratch_address(&base); rc != PML_OK) return rc;
  if (host.size() < need) host.resize(need);
  auto *gd = reinterpret_cast<GridDev *>(host.data());
  auto *units = reinterpret_cast<TileDev *>(host.data() + unit_off);
  size_t u = 0;
  for (int g = 0; g < n_grids; ++g) {
    const pml_grid &s = grids[g];
    GridDev &d = gd[g];
    d.n0 = s.n0;
    d.n1 = s.n1;
    d.off = s.state_offset;
    d.bc = s.bc_offset;
    d.inv_dx0sq = 1.0 / (s.dx0 * s.dx0);
    d.inv_dx1sq = 1.0 / (s.dx1 * s.dx1);
    d.inv_2dx0 = 1.0 / (2.0 * s.dx0);
    d.inv_2dx1 = 1.0 / (2.0 * s.dx1);
    d.m2dx0 = -2.0 * s.dx0;
    d.m2dx1 = -2.0 * s.dx1;
    d.p2dx0 = 2.0 * s.dx0;
    d.p2dx1 = 2.0 * s.dx1;
    for (int i0 = 0; i0 < s.n0; i0 += V.oy)
      for (int j0 = 0; j0 < s.n1; j0 += V.ox) units[u++] = TileDev{g, i0, j0, 0};
  }
  // a pageable-memory copy: the host buffer may be reused once it returns
  PML_HIP_CHECK(hipMemcpyAsync(base, host.data(), need, hipMemcpyHostToDevice, st));
  out->grids = base;
  out->units = static_cast<unsigned char *>(base) + unit_off;
  out->n_units = static_cast<int>(n_units);
  return PML_OK;
}

static int fdm_solve_impl(int eq_id, const double *eq_params, int n_params,
                          int integrator_id, int precision, int ndim, const pml_grid *grids,
                          int n_grids, const double *bc_values, void *state_hi,
                          void *state_lo, void *work_hi, void *work_lo, int64_t total_elems,
                          double dt, int64_t n_steps, void *snapshots_hi,
                          void *snapshots_lo, int64_t snapshot_every, int tile_config,
                          int flags, void *stream, bool no_alloc) {
  int n_fields = 0;
  if (ndim != 1 && ndim != 2) return fail(PML_E_INVALID, "ndim must be 1 or 2");
  const auto *vars = find_variants(eq_id, integrator_id, precision, ndim, &n_fields);
  if (!vars) return fail(PML_E_UNSUPPORTED, "no device kernel for this equation/integrator/precision");
  if (tile_config < 0 || tile_config >= static_cast<int>(vars->size()))
    return fail(PML_E_INVALID, "tile_config out of range");
  if (n_params < 0 || n_params > 6) return fail(PML_E_INVALID, "at most 6 equation parameters");
  if (n_grids <= 0 || !grids) return fail(PML_E_INVALID, "no grids");
  if (!state_hi || !work_hi || !bc_values) return fail(PML_E_INVALID, "null buffer");
  if (precision == PML_F32C && (!state_lo || !work_lo))
    return fail(PML_E_INVALID, "compensated precision needs lo planes");
  if (n_steps < 0) return fail(PML_E_INVALID, "n_steps must be non-negative");
  if (snapshots_hi && snapshot_every <= 0)
    return fail(PML_E_INVALID, "snapshot_every must be positive");
  if (snapshots_hi && precision == PML_F32C && !snapshots_lo)
    return fail(PML_E_INVALID, "compensated snapshots need a lo plane");
  const bool use_graph = flags & PML_FLAG_GRAPH;
  const bool constrain = flags & PML_FLAG_CONSTRAIN_INPUT;
  const bool prepare = flags & PML_FLAG_PREPARE;

  // Two-column stream variants move two values per lane instruction: every
  // grid's row length and state offset must be even and the planes aligned
  // to two values; otherwise the default variant (bitwise-identical
  // results) runs instead.
  const size_t esz_v = precision == PML_F64 ? 8 : 4;
  auto cpl_ok = [&](const Variant &v) {
    if (v.cpl == 1) return true;
    const uintptr_t align = static_cast<uintptr_t>(v.cpl) * esz_v;
    for (const void *ptr : {static_cast<const void *>(state_hi), static_cast<const void *>(state_lo),
                            static_cast<const void *>(work_hi), static_cast<const void *>(work_lo),
                            static_cast<const void *>(snapshots_hi),
                            static_cast<const void *>(snapshots_lo)})
      if (reinterpret_cast<uintptr_t>(ptr) % align) return false;
    if ((total_elems * static_cast<int64_t>(esz_v)) % static_cast<int64_t>(align)) return false;
    for (int g = 0; g < n_grids; ++g)
      if (grids[g].n1 % v.cpl || grids[g].state_offset % v.cpl) return false;
    return true;
  };
  const Variant &V = cpl_ok((*vars)[tile_config]) ? (*vars)[tile_config] : (*vars)[0];
  // Fused steps per launch must divide the snapshot interval.
  const bool can_fuse = V.steps > 1 && (!snapshots_hi || snapshot_every % V.steps == 0);
  const int steps_per_launch = can_fuse ? V.steps : 1;

  for (int g = 0; g < n_grids; ++g) {
    const pml_grid &s = grids[g];
    if (s.n1 < 3 || (ndim == 2 && s.n0 < 3) || (ndim == 1 && s.n0 != 1))
      return fail(PML_E_INVALID, "each active axis needs at least 3 vertices");
    if (!(s.dx0 > 0) || !(s.dx1 > 0)) return fail(PML_E_INVALID, "spacing must be positive");
    const int64_t need = s.state_offset + static_cast<int64_t>(n_fields) * s.n0 * s.n1;
    if (s.state_offset < 0 || need > total_elems)
      return fail(PML_E_INVALID, "grid state exceeds total_elems");
    if (s.bc_offset < 0) return fail(PML_E_INVALID, "negative bc_offset");
    if (static_cast<int64_t>(n_fields) * s.n0 * s.n1 >= (int64_t(1) << 31))
      return fail(PML_E_INVALID, "a single grid must hold fewer than 2^31 values");
  }
  hipStream_t st = as_stream(stream);
  TableSet tab;
  if (no_alloc && (flags & PML_FLAG_PREPARE))
    return fail(PML_E_INVALID, "the allocation-free entry point has nothing to prepare");
  int rc = no_alloc ? scratch_tables(grids, n_grids, V, st, &tab)
                    : device_tables(grids, n_grids, V, ndim, &tab);
  if (rc) return rc;

  FdmArgs a{};
  a.tiles = static_cast<const TileDev *>(tab.units);
  a.grids = static_cast<const GridDev *>(tab.grids);
  a.bc = bc_values;
  a.dt = dt;
  for (int i = 0; i < 6; ++i) a.p[i] = i < n_params ? eq_params[i] : 0.0;
  a.n_tiles = tab.n_units;
  a.ndim = ndim;
  // Stream kernels run four independent wavefronts (units) per workgroup.
  const int blocks = V.kind == kStreamKind ? (tab.n_units + 3) / 4 : tab.n_units;
  const dim3 grid_dim(blocks), block_dim(V.nt);
  const size_t esz = precision == PML_F64 ? 8 : 4;

  void *cur_hi = state_hi, *cur_lo = state_lo, *nxt_hi = work_hi, *nxt_lo = work_lo;
  int64_t done = 0, snap_idx = 0;

  auto launch = [&](KernelFn fn, int steps) -> int {
    a.in_hi = cur_hi;
    a.in_lo = cur_lo;
    a.out_hi = nxt_hi;
    a.out_lo = nxt_lo;
    a.snap_hi = nullptr;
    a.snap_lo = nullptr;
    if (snapshots_hi && (done + steps) % snapshot_every == 0) {
      a.snap_hi = static_cast<char *>(snapshots_hi) + snap_idx * total_elems * esz;
      a.snap_lo = snapshots_lo ? static_cast<char *>(snapshots_lo) + snap_idx * total_elems * esz
                               : nullptr;
      ++snap_idx;
    }
    if (a.snap_hi) fn = fn == V.fn_single ? V.fn_single_snap : V.fn_snap;
    hipLaunchKernelGGL(fn, grid_dim, block_dim, 0, st, a);
    PML_LAUNCH_CHECK();
    std::swap(cur_hi, nxt_hi);
    std::swap(cur_lo, nxt_lo);
    done += steps;
    return PML_OK;
  };

  // (no graphs without allocation: instantiating one allocates)
  const bool graphed = !no_alloc && use_graph && st != nullptr && !snapshots_hi;
  KernelFn main_fn = can_fuse ? V.fn : V.fn_single;
  const int64_t fused_launches = n_steps / steps_per_launch;
  if (prepare && !(graphed && fused_launches >= 2)) return PML_OK;  // tables only

  // Dirichlet values imposed on the input state in place (it is overwritten
  // by the result anyway).
  if (constrain && n_steps > 0 && !prepare) {
    const dim3 cgrid(n_grids, n_fields);
    if (precision == PML_F64)
      hipLaunchKernelGGL(fdm_constrain<PrecF64>, cgrid, dim3(256), 0, st, a.grids, bc_values,
                         state_hi, state_lo, ndim);
    else if (precision == PML_F32C)
      hipLaunchKernelGGL(fdm_constrain<PrecF32C>, cgrid, dim3(256), 0, st, a.grids, bc_values,
                         state_hi, state_lo, ndim);
    else
      hipLaunchKernelGGL(fdm_constrain<PrecF32>, cgrid, dim3(256), 0, st, a.grids, bc_values,
                         state_hi, state_lo, ndim);
    PML_LAUNCH_CHECK();
  }

  int64_t li = 0;

  // Graph replay (a created stream and no snapshots): a run of up to
  // kWhole launches is one graph replayed once; a longer run replays a
  // graph of kBlock launches (even, so the ping-pong parity is the same at
  // every replay) and launches the remainder directly.  Instantiated graphs
  // are cached with every argument they recorded, so repeated calls replay
  // without capturing.
  constexpr int64_t kWhole = 64, kBlock = 32;
  if (graphed && fused_launches >= 2) {
    const int64_t glen = fused_launches <= kWhole ? fused_launches : kBlock;
    const int64_t reps = fused_launches <= kWhole ? 1 : fused_launches / kBlock;
    int dev = 0;
    PML_HIP_CHECK(hipGetDevice(&dev));
    struct {
      const void *fn, *tiles, *grids, *bc, *p0, *p1, *p2, *p3;
      double dt, p[6];
      int64_t glen, n_tiles;
      int ndim, dev, blocks, nt;
    } gk{};
    gk.fn = reinterpret_cast<const void *>(main_fn);
    gk.tiles = a.tiles;
    gk.grids = a.grids;
    gk.bc = bc_values;
    gk.p0 = state_hi;
    gk.p1 = state_lo;
    gk.p2 = work_hi;
    gk.p3 = work_lo;
    gk.dt = dt;
    for (int i = 0; i < 6; ++i) gk.p[i] = a.p[i];
    gk.glen = glen;
    gk.n_tiles = a.n_tiles;
    gk.ndim = ndim;
    gk.dev = dev;
    gk.blocks = blocks;
    gk.nt = V.nt;
    const std::string key(reinterpret_cast<const char *>(&gk), sizeof(gk));
    hipGraphExec_t exec = nullptr;
    {
      std::lock_guard<std::mutex> lock(g_mu);
      auto it = g_graphs.find(key);
      if (it != g_graphs.end()) {
        it->second.stamp = ++g_stamp;
        exec = it->second.exec;
      }
    }
    if (!exec) {
      hipGraph_t graph = nullptr;
      PML_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      for (int64_t i = 0; i < glen; ++i) {
        rc = launch(main_fn, steps_per_launch);
        if (rc) {
          hipStreamEndCapture(st, &graph);
          if (graph) hipGraphDestroy(graph);
          return rc;
        }
      }
      PML_HIP_CHECK(hipStreamEndCapture(st, &graph));
      hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        hipGraphDestroy(graph);
        return fail(PML_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      }
      // Move the graph's launch data to the device now (stream-ordered),
      // so that the first replay costs what the later ones do.
      e = hipGraphUpload(exec, st);
      if (e != hipSuccess) {
        hipGraphExecDestroy(exec);
        hipGraphDestroy(graph);
        return fail(PML_E_HIP, std::string("hipGraphUpload: ") + hipGetErrorString(e));
      }
      std::lock_guard<std::mutex> lock(g_mu);
      if (g_graphs.size() >= kMaxGraphs) {
        rc = evict_oldest_graph();
        if (rc) {
          hipGraphExecDestroy(exec);
          hipGraphDestroy(graph);
          return rc;
        }
      }
      GraphEntry ent;
      ent.exec = exec;
      ent.graph = graph;
      ent.device = dev;
      ent.stamp = ++g_stamp;
      g_graphs.emplace(key, ent);
      // Capture recorded the launches without running them: rewind.
      cur_hi = state_hi;
      cur_lo = state_lo;
      nxt_hi = work_hi;
      nxt_lo = work_lo;
      done = 0;
    }
    if (prepare) return PML_OK;
    for (int64_t r = 0; r < reps; ++r) {
      hipError_t e = hipGraphLaunch(exec, st);
      if (e != hipSuccess)
        return fail(PML_E_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    }
    // After the replays the state is in state_hi (even graph length or an
    // even number of replays) or in work_hi.
    if ((glen * reps) & 1) {
      cur_hi = work_hi;
      cur_lo = work_lo;
      nxt_hi = state_hi;
      nxt_lo = state_lo;
    }
    done = glen * reps * steps_per_launch;
    li = glen * reps;
  }
  for (; li < fused_launches; ++li) {
    rc = launch(main_fn, steps_per_launch);
    if (rc) return rc;
  }
  while (done < n_steps) {
    rc = launch(V.fn_single, 1);
    if (rc) return rc;
  }
  if (cur_hi != state_hi) {
    PML_HIP_CHECK(hipMemcpyAsync(state_hi, cur_hi, total_elems * esz, hipMemcpyDeviceToDevice, st));
    if (precision == PML_F32C)
      PML_HIP_CHECK(
          hipMemcpyAsync(state_lo, cur_lo, total_elems * esz, hipMemcpyDeviceToDevice, st));
  }
  return PML_OK;
}

}  // namespace pml

extern "C" int pml_fdm_solve(int eq_id, const double *eq_params, int n_params,
                             int integrator_id, int precision, int ndim, const pml_grid *grids,
                             int n_grids, const double *bc_values, void *state_hi,
                             void *state_lo, void *work_hi, void *work_lo, int64_t total_elems,
                             double dt, int64_t n_steps, void *snapshots_hi,
                             void *snapshots_lo, int64_t snapshot_every, int tile_config,
                             int flags, void *stream) {
  return fdm_solve_impl(eq_id, eq_params, n_params, integrator_id, precision, ndim, grids,
                        n_grids, bc_values, state_hi, state_lo, work_hi, work_lo, total_elems,
                        dt, n_steps, snapshots_hi, snapshots_lo, snapshot_every, tile_config,
                        flags, stream, false);
}

extern "C" int pml_fdm_solve_noalloc(int eq_id, const double *eq_params, int n_params,
                                     int integrator_id, int precision, int ndim,
                                     const pml_grid *grids, int n_grids,
                                     const double *bc_values, void *state_hi, void *state_lo,
                                     void *work_hi, void *work_lo, int64_t total_elems,
                                     double dt, int64_t n_steps, void *snapshots_hi,
                                     void *snapshots_lo, int64_t snapshot_every,
                                     int tile_config, int flags, void *stream) {
  return fdm_solve_impl(eq_id, eq_params, n_params, integrator_id, precision, ndim, grids,
                        n_grids, bc_values, state_hi, state_lo, work_hi, work_lo, total_elems,
                        dt, n_steps, snapshots_hi, snapshots_lo, snapshot_every, tile_config,
                        flags, stream, true);
}
