// C ABI of libfsl: context management, stage sequencing and host transfers.
// All device work is enqueued on the context's stream; host reads of small
// results (counts, the CFL maximum, solver scalars) go through pinned memory.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fsl_internal.h"

using fsl::Dom;

namespace fsl {

// Reads the device-resident counts back: the exact particle count and
// leaf-list lengths replace the bounds a device-resident step left.
int host_counts(fsl_ctx *ctx) {
  if (!ctx->counts_on_device) return FSL_OK;
  FSL_HIP(hipMemcpyAsync(ctx->host_u32, ctx->dcount, 3 * sizeof(uint32_t),
                         hipMemcpyDeviceToHost, ctx->stream));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  ctx->np = ctx->host_u32[kCountParticles];
  ctx->n_act = ctx->host_u32[kCountActive];
  ctx->n_fluid = ctx->host_u32[kCountFluid];
  ctx->counts_on_device = false;
  return FSL_OK;
}

}  // namespace fsl

namespace {

// Staging array of fsl_get_scalar_noalloc: a file-scope device array (one
// copy per GPU, held until the process ends), so the call allocates
// nothing. 8 Mi floats (32 MiB): boxes of up to 8,388,608 voxels, i.e.
// domains up to 200^3 (the box is (n + 2)^3). Calls that overlap on one GPU
// share the array and race; the kernel writes every element a call reads,
// so nothing needs clearing.
constexpr int64_t kScratchFloats = int64_t(8) << 20;
__device__ __align__(256) float scratch[kScratchFloats];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `dev` (the current device), looked up
// once per device: host code uses this address, never the array's name,
// which is a host-side stand-in.
int scratch_address(int dev, float **out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return FSL_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FSL_ERR_HIP;
    g_scratch[dev] = static_cast<float *>(p);
  }
  *out = g_scratch[dev];
  return FSL_OK;
}

// Every entry point except the device-resident step works with the counts
// on the host: after a device-resident step they are read back first.
#define CHECK_CTX()                                    \
  do {                                                 \
    if (ctx == nullptr) return FSL_ERR_ARG;            \
    if (ctx->counts_on_device) {                       \
      const int rc_ = fsl::host_counts(ctx);           \
      if (rc_) return rc_;                             \
    }                                                  \
  } while (0)

template <typename T>
int dev_alloc(fsl_ctx *ctx, T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  FSL_HIP(hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T)));
  FSL_HIP(hipMemsetAsync(*p, 0, count * sizeof(T), ctx->stream));
  return FSL_OK;
}

int64_t box_count(const fsl_ctx *ctx) {
  return int64_t(ctx->dom.n[0] + 2) * (ctx->dom.n[1] + 2) * (ctx->dom.n[2] + 2);
}

int read_u32(fsl_ctx *ctx, const uint32_t *dev, uint32_t *out) {
  FSL_HIP(hipMemcpyAsync(ctx->host_u32, dev, sizeof(uint32_t), hipMemcpyDeviceToHost,
                         ctx->stream));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  *out = ctx->host_u32[0];
  return FSL_OK;
}

int sync_flags(fsl_ctx *ctx) {
  FSL_HIP(hipMemcpyAsync(ctx->flags_host, ctx->flags, sizeof(fsl::DevFlags),
                         hipMemcpyDeviceToHost, ctx->stream));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  return FSL_OK;
}

// Advect by dt (0 for a plain re-sort), drop deleted particles, and counting
// sort by cell.
int sort_particles(fsl_ctx *ctx, float dt, bool drop_outside, bool keep_gv,
                   const fsl::KillBoxes &kill = fsl::KillBoxes()) {
  FSL_HIP(fsl::launch_advect(ctx, dt, nullptr, drop_outside, kill));
  FSL_HIP(fsl::exclusive_scan(ctx->vcount, ctx->vstart, ctx->nvox, ctx->scan_tmp, ctx->stream));
  uint32_t alive = 0;
  int rc = read_u32(ctx, ctx->vstart + ctx->nvox, &alive);
  if (rc) return rc;
  FSL_HIP(fsl::launch_sort_tail(ctx, alive, keep_gv));
  ctx->outside_deleted = false;
  return FSL_OK;
}

// Switch to "every leaf active" (grids set directly from the host).
int all_leaves(fsl_ctx *ctx) {
  if (ctx->n_act != ctx->nleaves) {
    FSL_HIP(fsl::launch_all_leaves(ctx));
    ctx->n_act = ctx->nleaves;
    FSL_HIP(fsl::launch_clear_leaves(ctx, fsl::act_leaves(ctx)));
  }
  ctx->grid_ready = true;
  return FSL_OK;
}

// The velocity update's consistency flag of the last step (copied to the
// host at the step's end): FSL_ERR_VELOCITY_UPDATE when it is set.
int check_step_flags(fsl_ctx *ctx) {
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->host_u32[16]) {
    ctx->err = "face active in grid_velocity_ but not in grid_velocity_update_";
    return FSL_ERR_VELOCITY_UPDATE;
  }
  return FSL_OK;
}

// FlipDriver::RunOneStep (flip_app_sim.cc:211-254) with the counts on the
// device: particle count, active leaves and FLUID leaves stay on the device
// from the advection to the divergence, and the host reads them (with the
// solver flags and the device dt) once, after the divergence, to size the
// solve. The step is two halves: step_front enqueues everything up to that
// read (it can be captured in a HIP graph: with graph set, every launch is
// sized by an upper bound and reads its count on the device, so the
// captured launches are valid for any later step), step_back waits for the
// counts, solves and enqueues the rest. dt_dev, when set, holds dt on the
// device; with device_dt the front also computes it (CFL and k_dt). The
// velocity update's flag is copied to host_u32[16] and checked by the
// caller (check_step_flags) once the stream has drained.
int step_front(fsl_ctx *ctx, float dt, const float *dt_dev, bool device_dt, bool graph,
               const fsl::DtParams *dtp) {
  int rc;
  auto mark = [&](int ev) -> int {
    if (ctx->timing) FSL_HIP(hipEventRecord(ctx->tev[ev], ctx->stream));
    return FSL_OK;
  };
#define STAGE(s, body)                     \
  if ((rc = mark(2 * (s)))) return rc;     \
  body;                                    \
  if ((rc = mark(2 * (s) + 1))) return rc;
  // The host's counts go to the device (staged in host_u32[8..10], which the
  // caller fills before a graph replay); from here on they are bounds.
  ctx->host_u32[8 + fsl::kCountParticles] = uint32_t(ctx->np);
  ctx->host_u32[8 + fsl::kCountActive] = uint32_t(ctx->n_act);
  ctx->host_u32[8 + fsl::kCountFluid] = uint32_t(ctx->n_fluid);
  FSL_HIP(hipMemcpyAsync(ctx->dcount, ctx->host_u32 + 8, 3 * sizeof(uint32_t),
                         hipMemcpyHostToDevice, ctx->stream));
  ctx->counts_on_device = true;
  if (graph) ctx->n_act = ctx->nleaves;
  if (device_dt) {
    // WhileAndComputeDt: the CFL maximum of the current grid velocity, then
    // dt on the device.
    FSL_HIP(fsl::launch_fill_u32(ctx, &ctx->flags->cfl_bits, 1, 0u));
    if (ctx->grid_ready && ctx->n_act > 0) FSL_HIP(fsl::launch_cfl(ctx));
    FSL_HIP(fsl::launch_dt(ctx, *dtp));
  }
  // MoveParticles + DeleteParticlesInSolids. Within a step nothing reads the
  // particles' grid velocities between the move and G2P, which rewrites
  // them: the sort leaves them behind.
  STAGE(FSL_T_MOVE, {
    FSL_HIP(fsl::launch_advect(ctx, dt, dt_dev, ctx->outside_deleted, fsl::KillBoxes()));
    FSL_HIP(fsl::exclusive_scan(ctx->vcount, ctx->vstart, ctx->nvox, ctx->scan_tmp,
                                ctx->stream));
    FSL_HIP(fsl::launch_sort_tail(ctx, 0, false));
    ctx->outside_deleted = false;
  })
  // ClearGridData: the old list's leaves, then the new active list.
  STAGE(FSL_T_CLEAR, {
    FSL_HIP(fsl::launch_clear_shell(ctx));
    if (ctx->n_act > 0) FSL_HIP(fsl::launch_clear_leaves(ctx, fsl::act_leaves(ctx)));
    FSL_HIP(fsl::launch_leaf_flags(ctx));
    ctx->n_act = ctx->nleaves;
    ctx->n_fluid = ctx->nleaves;
    ctx->grid_ready = true;
    ctx->pcg_ready = false;
  })
  STAGE(FSL_T_P2G, FSL_HIP(fsl::launch_p2g(ctx)))
  // ApplyBoundaryConditions + SaveGridVelocities in one pass.
  STAGE(FSL_T_BC_SAVE, FSL_HIP(fsl::launch_bc_save(ctx)); ctx->bc_applied = true)
  ctx->outside_deleted = true;  // DeleteOutsideParticles: applied by the next sort
  STAGE(FSL_T_GRAVITY_BC, FSL_HIP(fsl::launch_gravity(ctx, dt, dt_dev));
        FSL_HIP(fsl::launch_bc(ctx)))
  STAGE(FSL_T_DIVERGENCE, {
    FSL_HIP(fsl::launch_fill_u32(ctx, ctx->lfluid, ctx->nleaves, 0u));
    FSL_HIP(fsl::launch_fill_u32(ctx, reinterpret_cast<uint32_t *>(ctx->flags),
                                 int64_t(sizeof(fsl::DevFlags) / sizeof(uint32_t)), 0u));
    FSL_HIP(fsl::launch_divergence(ctx));
    FSL_HIP(fsl::launch_compact(ctx, ctx->lfluid, ctx->fluid_list,
                                ctx->dcount + fsl::kCountFluid));
    // The counts, the flags and the step state, read by step_back in one
    // round trip.
    FSL_HIP(hipMemcpyAsync(ctx->host_u32, ctx->dcount, 3 * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, ctx->stream));
    FSL_HIP(hipMemcpyAsync(ctx->flags_host, ctx->flags, sizeof(fsl::DevFlags),
                           hipMemcpyDeviceToHost, ctx->stream));
    FSL_HIP(hipMemcpyAsync(ctx->sstate_host, ctx->sstate, sizeof(fsl::StepState),
                           hipMemcpyDeviceToHost, ctx->stream));
  })
  return FSL_OK;
}

int step_back(fsl_ctx *ctx, const fsl_solve_params *params, fsl_solve_stats *stats) {
  int rc;
  auto mark = [&](int ev) -> int {
    if (ctx->timing) FSL_HIP(hipEventRecord(ctx->tev[ev], ctx->stream));
    return FSL_OK;
  };
  if (!ctx->ev_counts) FSL_HIP(hipEventCreateWithFlags(&ctx->ev_counts, hipEventDisableTiming));
  FSL_HIP(hipEventRecord(ctx->ev_counts, ctx->stream));
  FSL_HIP(hipEventSynchronize(ctx->ev_counts));
  ctx->np = ctx->host_u32[fsl::kCountParticles];
  ctx->n_act = ctx->host_u32[fsl::kCountActive];
  ctx->n_fluid = ctx->host_u32[fsl::kCountFluid];
  ctx->counts_on_device = false;
  if (ctx->flags_host->zero_diag) {
    ctx->err = "FLUID cell with no non-SOLID neighbour (zero Jacobi diagonal)";
    return FSL_ERR_ZERO_DIAGONAL;
  }
  fsl_solve_stats local{};
  fsl_solve_stats *st = stats ? stats : &local;
  *st = fsl_solve_stats{};
  st->n_rows = int(ctx->flags_host->n_rows);
  st->n_cols = st->n_rows;
  STAGE(FSL_T_SOLVE, {
    rc = fsl::pcg_solve_iterations(ctx, params, st);
    if (rc) {
      if (ctx->err.empty()) ctx->err = "solver failed";
      st->iterations = ctx->pcg_host->it;
      return rc;
    }
    // CopySolution, and the true residual for the statistics: read back
    // while the rest of the step runs.
    if ((rc = fsl::pcg_finish_enqueue(ctx))) return rc;
  })
  // The velocity update's consistency flag is read once, at the end.
  STAGE(FSL_T_UPDATE, FSL_HIP(fsl::launch_pressure_update(ctx));
        FSL_HIP(fsl::launch_fill_u32(ctx, &ctx->flags->vel_update, 1, 0u));
        FSL_HIP(fsl::launch_velocity_update(ctx)))
  STAGE(FSL_T_G2P, FSL_HIP(fsl::launch_extrapolate(ctx)); FSL_HIP(fsl::launch_g2p(ctx)))
#undef STAGE
  FSL_HIP(hipMemcpyAsync(ctx->host_u32 + 16, &ctx->flags->vel_update, sizeof(uint32_t),
                         hipMemcpyDeviceToHost, ctx->stream));
  if ((rc = fsl::pcg_finish_collect(ctx, st))) return rc;
  if (ctx->timing) {
    FSL_HIP(hipStreamSynchronize(ctx->stream));
    for (int s = FSL_T_MOVE; s <= FSL_T_G2P; ++s) {
      float ms = 0.0f;
      FSL_HIP(hipEventElapsedTime(&ms, ctx->tev[2 * s], ctx->tev[2 * s + 1]));
      ctx->tacc[s] += ms;
    }
    ctx->tacc[FSL_T_STEPS] += 1;
  }
  return FSL_OK;
}

int run_step(fsl_ctx *ctx, float dt, const float *dt_dev, const fsl_solve_params *params,
             fsl_solve_stats *stats) {
  int rc = step_front(ctx, dt, dt_dev, false, false, nullptr);
  if (rc) return rc;
  return step_back(ctx, params, stats);
}

// The front half of a device-dt step as a HIP graph, one per particle
// buffer (the sort gathers into the other buffer and the buffers swap each
// step), captured on first use and again when the particle count bound
// grows (it sizes the particle launches).
int step_front_graph(fsl_ctx *ctx, const fsl::DtParams &p) {
  const int parity = ctx->P.a[0] == ctx->pbuf0 ? 0 : 1;
  fsl::StepGraph &g = ctx->graphs[parity];
  const bool fresh = !g.exec || ctx->np > g.np || std::memcmp(&g.dtp, &p, sizeof p) != 0 ||
                     ctx->outside_deleted != g.outside_deleted;
  if (fresh) {
    if (g.exec) {
      (void)hipGraphExecDestroy(g.exec);
      g.exec = nullptr;
    }
    g.np = ctx->np;
    g.dtp = p;
    g.outside_deleted = ctx->outside_deleted;
    // Capturing only records the launches: the host state the front
    // changes is restored afterwards (the replay below runs them).
    const fsl::Particles P = ctx->P, Q = ctx->Q;
    const int64_t np = ctx->np, n_act = ctx->n_act, n_fluid = ctx->n_fluid;
    const bool od = ctx->outside_deleted, gr = ctx->grid_ready, bc = ctx->bc_applied;
    FSL_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = step_front(ctx, 0.0f, &ctx->sstate->dt, true, true, &p);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    ctx->P = P;
    ctx->Q = Q;
    ctx->np = np;
    ctx->n_act = n_act;
    ctx->n_fluid = n_fluid;
    ctx->outside_deleted = od;
    ctx->grid_ready = gr;
    ctx->bc_applied = bc;
    ctx->counts_on_device = false;
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    FSL_HIP(e);
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    FSL_HIP(ei);
  }
  ctx->host_u32[8 + fsl::kCountParticles] = uint32_t(ctx->np);
  ctx->host_u32[8 + fsl::kCountActive] = uint32_t(ctx->n_act);
  ctx->host_u32[8 + fsl::kCountFluid] = uint32_t(ctx->n_fluid);
  FSL_HIP(hipGraphLaunch(g.exec, ctx->stream));
  // The host-side effects of the front (as step_front leaves them).
  const fsl::Particles t = ctx->P;
  ctx->P = ctx->Q;
  ctx->Q = t;
  ctx->counts_on_device = true;
  ctx->n_act = ctx->nleaves;
  ctx->n_fluid = ctx->nleaves;
  ctx->grid_ready = true;
  ctx->pcg_ready = false;
  ctx->bc_applied = true;
  ctx->outside_deleted = true;
  return FSL_OK;
}

int require_grid(fsl_ctx *ctx) {
  if (!ctx->grid_ready) {
    ctx->err = "grid not initialised: call fsl_clear after placing particles";
    return FSL_ERR_STATE;
  }
  return FSL_OK;
}

}  // namespace

namespace fsl {
bool on_ctx_device(const fsl_ctx *c, const void *p) {
  if (!p) return false;
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // plain host memory: clear the error the query leaves
    return false;
  }
  return a.type == hipMemoryTypeDevice && a.device == c->device;
}
}  // namespace fsl

extern "C" {

int fsl_create(fsl_ctx **out, const int dims[3], const int tiles[3], int device,
               int64_t particle_capacity) {
  if (!out || !dims) return FSL_ERR_ARG;
  *out = nullptr;
  int T[3] = {1, 1, 1};
  if (tiles) for (int a = 0; a < 3; ++a) T[a] = tiles[a];
  for (int a = 0; a < 3; ++a) {
    if (dims[a] <= 0 || dims[a] % 8 != 0 || T[a] <= 0 || dims[a] % (8 * T[a]) != 0) return FSL_ERR_ARG;
  }
  if (T[0] * T[1] * T[2] > 64 || particle_capacity < 0 || particle_capacity > 0xFFFFFFF0ll)
    return FSL_ERR_ARG;
  fsl_ctx *ctx = new fsl_ctx();
  ctx->device = device;
  for (int a = 0; a < 3; ++a) {
    ctx->dom.n[a] = dims[a];
    ctx->dom.L[a] = dims[a] / 8 + 2;
    ctx->dom.T[a] = T[a];
  }
  ctx->nleaves = fsl::leaf_count(ctx->dom);
  ctx->nvox = ctx->nleaves * fsl::kLeaf;
  if (ctx->nvox >= int64_t(0xFFFFFFF0ll)) { delete ctx; return FSL_ERR_ARG; }
  ctx->cap = particle_capacity;
  auto fail = [&](int rc) { fsl_destroy(ctx); return rc; };
  if (hipSetDevice(device) != hipSuccess) { delete ctx; return FSL_ERR_HIP; }
  if (scratch_address(device, &ctx->scratch) != FSL_OK) { delete ctx; return FSL_ERR_HIP; }
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return FSL_ERR_HIP;
  }
  const size_t nv = size_t(ctx->nvox), nl = size_t(ctx->nleaves);
  int rc = FSL_OK;
  for (int a = 0; a < 3 && !rc; ++a) rc = dev_alloc(ctx, &ctx->v[a], nv);
  for (int a = 0; a < 3 && !rc; ++a) rc = dev_alloc(ctx, &ctx->u[a], nv);
  if (!rc) rc = dev_alloc(ctx, &ctx->vflag, nv);
  if (!rc) rc = dev_alloc(ctx, &ctx->marker, nv);
  if (!rc) rc = dev_alloc(ctx, &ctx->solid, nv);
  if (!rc) rc = dev_alloc(ctx, &ctx->lab, nv);
  if (!rc) rc = dev_alloc(ctx, &ctx->div, nv);
  if (!rc) rc = dev_alloc(ctx, &ctx->pres, nv);
  double *tmp = nullptr;
  if (!rc) { rc = dev_alloc(ctx, &tmp, nv); ctx->x = tmp; }
  if (!rc) { rc = dev_alloc(ctx, &tmp, nv); ctx->r = tmp; }
  if (!rc) { rc = dev_alloc(ctx, &tmp, nv); ctx->z = tmp; }
  if (!rc) { rc = dev_alloc(ctx, &tmp, nv); ctx->p[0] = tmp; }
  if (!rc) { rc = dev_alloc(ctx, &tmp, nv); ctx->p[1] = tmp; }
  if (!rc) { rc = dev_alloc(ctx, &tmp, nv); ctx->Ap = tmp; }
  if (!rc) rc = dev_alloc(ctx, &ctx->x64, nv);
  if (!rc) { rc = dev_alloc(ctx, &tmp, fsl::pcg_partials_bytes() / sizeof(double)); ctx->partials = tmp; }
  if (!rc) rc = dev_alloc(ctx, &ctx->pcg, 1);
  if (!rc) rc = dev_alloc(ctx, &ctx->lhas, nl);
  if (!rc) rc = dev_alloc(ctx, &ctx->lact, nl);
  if (!rc) rc = dev_alloc(ctx, &ctx->lfluid, nl);
  if (!rc) rc = dev_alloc(ctx, &ctx->lscan, nl + 1);
  if (!rc) rc = dev_alloc(ctx, &ctx->act_list, nl);
  if (!rc) rc = dev_alloc(ctx, &ctx->morton, nl);
  if (!rc) rc = dev_alloc(ctx, &ctx->lflag_m, nl);
  if (!rc) rc = dev_alloc(ctx, &ctx->lrank, nl);
  std::vector<int32_t> lrank(nl);
  if (!rc) {
    // Leaves in VDB traversal order of their origins (the particle order).
    std::vector<std::pair<uint64_t, int32_t>> keyed(nl);
    for (int64_t l = 0; l < ctx->nleaves; ++l) {
      int o[3];
      fsl::leaf_origin(ctx->dom, l, o);
      keyed[size_t(l)] = {fsl::vdb_leaf_key(o[0], o[1], o[2]), int32_t(l)};
    }
    std::sort(keyed.begin(), keyed.end());
    for (size_t q = 0; q < nl; ++q) lrank[size_t(keyed[q].second)] = int32_t(q);
    if (hipMemcpyAsync(ctx->lrank, lrank.data(), nl * sizeof(int32_t), hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess)
      rc = FSL_ERR_HIP;
    ctx->dom.lrank = ctx->lrank;
  }
  if (!rc) {
    // Leaves sorted by the Morton code of their coordinates.
    std::vector<std::pair<uint64_t, int32_t>> keyed(nl);
    for (int64_t l = 0; l < ctx->nleaves; ++l) {
      const int64_t l12 = int64_t(ctx->dom.L[1]) * ctx->dom.L[2];
      const uint32_t c[3] = {uint32_t(l / l12), uint32_t((l / ctx->dom.L[2]) % ctx->dom.L[1]),
                             uint32_t(l % ctx->dom.L[2])};
      uint64_t code = 0;
      for (int b = 0; b < 21; ++b)
        for (int a = 0; a < 3; ++a) code |= uint64_t((c[a] >> b) & 1u) << (3 * b + (2 - a));
      keyed[size_t(l)] = {code, int32_t(l)};
    }
    std::sort(keyed.begin(), keyed.end());
    std::vector<int32_t> order(nl);
    for (size_t q = 0; q < nl; ++q) order[q] = keyed[q].second;
    // On the context's stream, behind the zero fill dev_alloc queued there
    // (a null-stream copy is not ordered with it and could be overwritten).
    if (hipMemcpyAsync(ctx->morton, order.data(), nl * sizeof(int32_t), hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = FSL_ERR_HIP;
  }
  if (!rc) rc = dev_alloc(ctx, &ctx->fluid_list, nl);
  for (int a = 0; a < 9 && !rc; ++a)
    rc = dev_alloc(ctx, &ctx->P.a[a], size_t(ctx->cap + fsl::kParticlePad));
  for (int a = 0; a < 9 && !rc; ++a)
    rc = dev_alloc(ctx, &ctx->Q.a[a], size_t(ctx->cap + fsl::kParticlePad));
  ctx->pbuf0 = ctx->P.a[0];
  if (!rc) rc = dev_alloc(ctx, &ctx->key, size_t(ctx->cap));
  if (!rc) rc = dev_alloc(ctx, &ctx->perm, size_t(ctx->cap + fsl::kParticlePad));
  if (!rc) rc = dev_alloc(ctx, &ctx->p2g_w, 3 * nv);
  if (!rc) {
    uint8_t *ring = nullptr;
    rc = dev_alloc(ctx, &ring, fsl::p2g_ring_bytes(ctx->nleaves));
    ctx->p2g_ring = ring;
  }
  if (!rc) rc = dev_alloc(ctx, &ctx->vcount, nv);
  if (!rc) rc = dev_alloc(ctx, &ctx->vstart, nv + 1);
  const int64_t ncell = int64_t(dims[0]) * dims[1] * dims[2];
  ctx->scan_tmp_len = fsl::scan_tmp_elems(std::max<int64_t>(ctx->nvox, ncell));
  if (!rc) rc = dev_alloc(ctx, &ctx->scan_tmp, size_t(ctx->scan_tmp_len));
  if (!rc) rc = dev_alloc(ctx, &ctx->flags, 1);
  if (!rc) rc = dev_alloc(ctx, &ctx->dcount, 4);
  if (!rc) rc = dev_alloc(ctx, &ctx->sstate, 1);
  if (!rc && hipHostMalloc(reinterpret_cast<void **>(&ctx->sstate_host), sizeof(fsl::StepState)) != hipSuccess) rc = FSL_ERR_HIP;
  if (!rc && hipHostMalloc(reinterpret_cast<void **>(&ctx->flags_host), sizeof(fsl::DevFlags)) != hipSuccess) rc = FSL_ERR_HIP;
  if (!rc && hipHostMalloc(reinterpret_cast<void **>(&ctx->pcg_host), 3 * sizeof(fsl::PcgState)) != hipSuccess) rc = FSL_ERR_HIP;
  if (!rc && hipHostMalloc(reinterpret_cast<void **>(&ctx->host_u32), 64 * sizeof(uint32_t)) != hipSuccess) rc = FSL_ERR_HIP;
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = FSL_ERR_HIP;
  if (rc) return fail(rc);
  std::memset(ctx->flags_host, 0, sizeof(fsl::DevFlags));
  *out = ctx;
  return FSL_OK;
}

int fsl_destroy(fsl_ctx *ctx) {
  CHECK_CTX();
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  fsl::adjoint_free(ctx);
  fsl::ic_free(ctx);
  void *bufs[] = {ctx->v[0], ctx->v[1], ctx->v[2], ctx->u[0], ctx->u[1], ctx->u[2], ctx->vflag,
                  ctx->marker, ctx->solid, ctx->lab, ctx->div, ctx->pres, ctx->x, ctx->r, ctx->z,
                  ctx->p[0], ctx->p[1], ctx->Ap, ctx->x64, ctx->partials, ctx->pcg, ctx->lhas,
                  ctx->lact, ctx->lfluid, ctx->lscan, ctx->act_list, ctx->fluid_list, ctx->morton,
                  ctx->lrank,
                  ctx->lflag_m, ctx->key, ctx->p2g_w, ctx->p2g_ring,
                  ctx->perm, ctx->vcount, ctx->vstart, ctx->scan_tmp, ctx->flags, ctx->dcount,
                  ctx->sstate};
  for (void *b : bufs) if (b) (void)hipFree(b);
  for (int a = 0; a < 9; ++a) {
    if (ctx->P.a[a]) (void)hipFree(ctx->P.a[a]);
    if (ctx->Q.a[a]) (void)hipFree(ctx->Q.a[a]);
  }
  if (ctx->src_buf) (void)hipFree(ctx->src_buf);
  for (hipEvent_t e : ctx->tev) if (e) (void)hipEventDestroy(e);
  if (ctx->ev_finish) (void)hipEventDestroy(ctx->ev_finish);
  for (fsl::StepGraph &g : ctx->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (ctx->ev_counts) (void)hipEventDestroy(ctx->ev_counts);
  if (ctx->flags_host) (void)hipHostFree(ctx->flags_host);
  if (ctx->pcg_host) (void)hipHostFree(ctx->pcg_host);
  if (ctx->sstate_host) (void)hipHostFree(ctx->sstate_host);
  if (ctx->host_u32) (void)hipHostFree(ctx->host_u32);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FSL_OK;
}

const char *fsl_last_error(const fsl_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fsl_synchronize(fsl_ctx *ctx) {
  CHECK_CTX();
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  return FSL_OK;
}

int fsl_set_params(fsl_ctx *ctx, float gravity_scaled, const float force[3], float flip) {
  CHECK_CTX();
  if (!force) return FSL_ERR_ARG;
  ctx->gravity = gravity_scaled;
  for (int a = 0; a < 3; ++a) ctx->force[a] = force[a];
  ctx->flip = flip;
  return FSL_OK;
}

int fsl_set_box_boundary(fsl_ctx *ctx, int tank) {
  CHECK_CTX();
  // hipMemcpy runs on the null stream, which does not wait for the
  // context's (non-blocking) stream: drain it first.
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  const Dom &d = ctx->dom;
  std::vector<uint8_t> s(size_t(ctx->nvox), 0);
  FSL_HIP(hipMemcpy(s.data(), ctx->solid, s.size(), hipMemcpyDeviceToHost));
  auto set = [&](int i, int j, int k) { s[size_t(fsl::vox_of(d, i, j, k))] = 1; };
  for (int j = -1; j <= d.n[1]; ++j)
    for (int k = -1; k <= d.n[2]; ++k) { set(-1, j, k); set(d.n[0], j, k); }
  for (int i = -1; i <= d.n[0]; ++i)
    for (int k = -1; k <= d.n[2]; ++k) { set(i, -1, k); if (!tank) set(i, d.n[1], k); }
  for (int i = -1; i <= d.n[0]; ++i)
    for (int j = -1; j <= d.n[1]; ++j) { set(i, j, -1); set(i, j, d.n[2]); }
  FSL_HIP(hipMemcpy(ctx->solid, s.data(), s.size(), hipMemcpyHostToDevice));
  return FSL_OK;
}

int fsl_add_solid_box(fsl_ctx *ctx, const int lo[3], const int hi[3]) {
  CHECK_CTX();
  if (!lo || !hi) return FSL_ERR_ARG;
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  const Dom &d = ctx->dom;
  std::vector<uint8_t> s(size_t(ctx->nvox), 0);
  FSL_HIP(hipMemcpy(s.data(), ctx->solid, s.size(), hipMemcpyDeviceToHost));
  int a[3], b[3];
  for (int q = 0; q < 3; ++q) { a[q] = std::max(lo[q], -1); b[q] = std::min(hi[q], d.n[q]); }
  for (int i = a[0]; i <= b[0]; ++i)
    for (int j = a[1]; j <= b[1]; ++j)
      for (int k = a[2]; k <= b[2]; ++k) s[size_t(fsl::vox_of(d, i, j, k))] = 1;
  FSL_HIP(hipMemcpy(ctx->solid, s.data(), s.size(), hipMemcpyHostToDevice));
  return FSL_OK;
}

int fsl_set_particles(fsl_ctx *ctx, int64_t n, const float *pos, const float *vel,
                      const float *gv) {
  CHECK_CTX();
  if (n < 0 || (n > 0 && (!pos || !vel))) return FSL_ERR_ARG;
  if (n > ctx->cap) { ctx->err = "particle capacity exceeded"; return FSL_ERR_CAPACITY; }
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<float> col(static_cast<size_t>(n));
  for (int a = 0; a < 9; ++a) {
    const float *src = a < 3 ? pos : a < 6 ? vel : (gv ? gv : vel);
    const int comp = a % 3;
    for (int64_t q = 0; q < n; ++q) col[size_t(q)] = src[3 * q + comp];
    if (n) FSL_HIP(hipMemcpy(ctx->P.a[a], col.data(), size_t(n) * sizeof(float), hipMemcpyHostToDevice));
  }
  ctx->np = n;
  return sort_particles(ctx, 0.0f, false, true);
}

int fsl_seed_particles(fsl_ctx *ctx, const uint8_t *mask, int ppc, uint32_t seed,
                       const float velocity[3]) {
  CHECK_CTX();
  if (!mask || ppc <= 0 || !velocity) return FSL_ERR_ARG;
  const int64_t ncell = int64_t(ctx->dom.n[0]) * ctx->dom.n[1] * ctx->dom.n[2];
  uint8_t *mdev = nullptr;
  uint32_t *offs = nullptr;
  FSL_HIP(hipMalloc(reinterpret_cast<void **>(&mdev), size_t(ncell)));
  if (hipMalloc(reinterpret_cast<void **>(&offs), size_t(ncell + 1) * sizeof(uint32_t)) != hipSuccess) {
    (void)hipFree(mdev);
    ctx->err = "out of device memory";
    return FSL_ERR_HIP;
  }
  int rc = FSL_OK;
  uint32_t cells = 0;
  auto done = [&](int code) { (void)hipStreamSynchronize(ctx->stream); (void)hipFree(mdev); (void)hipFree(offs); return code; };
  if (hipMemcpy(mdev, mask, size_t(ncell), hipMemcpyHostToDevice) != hipSuccess) return done(FSL_ERR_HIP);
  if (fsl::launch_mask_count(ctx, mdev, ncell, offs) != hipSuccess) return done(FSL_ERR_HIP);
  if (fsl::exclusive_scan(offs, offs, ncell, ctx->scan_tmp, ctx->stream) != hipSuccess) return done(FSL_ERR_HIP);
  rc = read_u32(ctx, offs + ncell, &cells);
  if (rc) return done(rc);
  const int64_t n = int64_t(cells) * ppc;
  if (n > ctx->cap) { ctx->err = "particle capacity exceeded"; return done(FSL_ERR_CAPACITY); }
  if (fsl::launch_seed(ctx, mdev, offs, ppc, seed, velocity) != hipSuccess) return done(FSL_ERR_HIP);
  ctx->np = n;
  rc = sort_particles(ctx, 0.0f, false, true);
  return done(rc);
}

int fsl_add_source(fsl_ctx *ctx, const float lo[3], const float hi[3], const float velocity[3],
                   int ppc, uint32_t seed, uint32_t stream) {
  CHECK_CTX();
  if (!lo || !hi || !velocity || ppc <= 0) return FSL_ERR_ARG;
  fsl::SourceBox s{};
  int64_t ncell = 1;
  for (int e = 0; e < 3; ++e) {
    // Coord(lo), Coord(hi): conversion to int truncates; then the
    // intersection with the domain box.
    s.c0[e] = std::max(int(lo[e]), 0);
    s.c1[e] = std::min(int(hi[e]), ctx->dom.n[e] - 1);
    if (s.c1[e] < s.c0[e]) return FSL_OK;
    ncell *= s.c1[e] - s.c0[e] + 1;
    s.lo[e] = lo[e];
    s.hi[e] = hi[e];
    s.vel[e] = velocity[e];
  }
  s.ppc = ppc;
  s.k0 = seed;
  s.k1 = stream;
  if (ctx->src_cap < ncell + 1) {
    if (ctx->src_buf) FSL_HIP(hipFree(ctx->src_buf));
    ctx->src_buf = nullptr;
    ctx->src_cap = 0;
    FSL_HIP(hipMalloc(reinterpret_cast<void **>(&ctx->src_buf), size_t(ncell + 1) * sizeof(uint32_t)));
    ctx->src_cap = ncell + 1;
  }
  FSL_HIP(fsl::launch_source_count(ctx, s, ncell, ctx->src_buf));
  FSL_HIP(fsl::exclusive_scan(ctx->src_buf, ctx->src_buf, ncell, ctx->scan_tmp, ctx->stream));
  uint32_t total = 0;
  int rc = read_u32(ctx, ctx->src_buf + ncell, &total);
  if (rc) return rc;
  if (ctx->np + int64_t(total) > ctx->cap) {
    ctx->err = "particle capacity exceeded";
    return FSL_ERR_CAPACITY;
  }
  if (total) FSL_HIP(fsl::launch_source_emit(ctx, s, ncell, ctx->src_buf, ctx->np));
  ctx->np += total;
  return FSL_OK;
}

int fsl_delete_particles_in_boxes(fsl_ctx *ctx, int n, const float *boxes) {
  CHECK_CTX();
  if (n < 0 || (n > 0 && !boxes)) return FSL_ERR_ARG;
  // A re-sort without advection (dt = 0) whose advection pass also drops
  // the particles in the boxes; up to kMaxKillBoxes boxes per pass.
  for (int b0 = 0; b0 < n; b0 += fsl::kMaxKillBoxes) {
    fsl::KillBoxes kill;
    kill.n = std::min(fsl::kMaxKillBoxes, n - b0);
    for (int b = 0; b < kill.n; ++b)
      for (int e = 0; e < 3; ++e) {
        kill.lo[b][e] = boxes[6 * (b0 + b) + e];
        kill.hi[b][e] = boxes[6 * (b0 + b) + 3 + e];
      }
    int rc = sort_particles(ctx, 0.0f, ctx->outside_deleted, true, kill);
    if (rc) return rc;
  }
  return FSL_OK;
}

int fsl_num_particles(fsl_ctx *ctx, int64_t *n) {
  CHECK_CTX();
  if (!n) return FSL_ERR_ARG;
  if (!ctx->outside_deleted) { *n = ctx->np; return FSL_OK; }
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  // Particles outside the partition box are deleted lazily (dropped by the
  // next sort); count the ones that remain.
  int64_t m = 0;
  for (int a = 0; a < 3; ++a) {
    std::vector<float> col(static_cast<size_t>(ctx->np));
    if (ctx->np) FSL_HIP(hipMemcpy(col.data(), ctx->P.a[a], col.size() * sizeof(float), hipMemcpyDeviceToHost));
    if (a == 0) m = 0;
    static thread_local std::vector<uint8_t> in;
    if (a == 0) in.assign(col.size(), 1);
    for (size_t q = 0; q < col.size(); ++q) {
      const int c = int(std::floor(col[q]));
      if (c < 0 || c >= ctx->dom.n[a]) in[q] = 0;
    }
    if (a == 2) for (uint8_t x : in) m += x;
  }
  *n = m;
  return FSL_OK;
}

int fsl_get_particles(fsl_ctx *ctx, float *pos, float *vel, float *gv) {
  CHECK_CTX();
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  const int64_t n = ctx->np;
  std::vector<float> col[9];
  for (int a = 0; a < 9; ++a) {
    col[a].resize(size_t(n));
    if (n) FSL_HIP(hipMemcpy(col[a].data(), ctx->P.a[a], size_t(n) * sizeof(float), hipMemcpyDeviceToHost));
  }
  int64_t m = 0;
  for (int64_t q = 0; q < n; ++q) {
    if (ctx->outside_deleted) {
      bool in = true;
      for (int a = 0; a < 3; ++a) {
        const int c = int(std::floor(col[a][size_t(q)]));
        if (c < 0 || c >= ctx->dom.n[a]) in = false;
      }
      if (!in) continue;
    }
    for (int a = 0; a < 3; ++a) {
      if (pos) pos[3 * m + a] = col[a][size_t(q)];
      if (vel) vel[3 * m + a] = col[3 + a][size_t(q)];
      if (gv) gv[3 * m + a] = col[6 + a][size_t(q)];
    }
    ++m;
  }
  return FSL_OK;
}

int fsl_cfl(fsl_ctx *ctx, float *inv_speed) {
  CHECK_CTX();
  if (!inv_speed) return FSL_ERR_ARG;
  float v_inf = 0.0f;
  if (ctx->grid_ready && ctx->n_act > 0) {
    FSL_HIP(hipMemsetAsync(&ctx->flags->cfl_bits, 0, sizeof(unsigned), ctx->stream));
    FSL_HIP(fsl::launch_cfl(ctx));
    uint32_t bits = 0;
    int rc = read_u32(ctx, &ctx->flags->cfl_bits, &bits);
    if (rc) return rc;
    std::memcpy(&v_inf, &bits, sizeof(float));
  }
  // flip_simulation.cc:485-487, with the reference's float/double mix.
  float v = float(std::fmax(std::sqrt(0.5 * double(ctx->gravity)), double(v_inf)));
  if (v < 1e-2) v = 1e-2f;
  *inv_speed = float(1.0 / double(v));
  return FSL_OK;
}

int fsl_move_particles(fsl_ctx *ctx, float dt) {
  CHECK_CTX();
  return sort_particles(ctx, dt, ctx->outside_deleted, true);
}

int fsl_clear(fsl_ctx *ctx) {
  CHECK_CTX();
  FSL_HIP(fsl::launch_clear_shell(ctx));
  if (ctx->n_act > 0) FSL_HIP(fsl::launch_clear_leaves(ctx, fsl::act_leaves(ctx)));
  FSL_HIP(fsl::launch_leaf_flags(ctx));
  uint32_t n_act = 0;
  int rc = read_u32(ctx, ctx->lscan + ctx->nleaves, &n_act);
  if (rc) return rc;
  ctx->n_act = n_act;
  ctx->n_fluid = 0;
  ctx->grid_ready = true;
  ctx->bc_applied = false;
  ctx->pcg_ready = false;
  return FSL_OK;
}

int fsl_p2g(fsl_ctx *ctx) {
  CHECK_CTX();
  int rc = require_grid(ctx);
  if (rc) return rc;
  FSL_HIP(fsl::launch_p2g(ctx));
  return FSL_OK;
}

int fsl_apply_bc(fsl_ctx *ctx) {
  CHECK_CTX();
  int rc = require_grid(ctx);
  if (rc) return rc;
  FSL_HIP(fsl::launch_bc(ctx));
  ctx->bc_applied = true;
  return FSL_OK;
}

int fsl_save_velocities(fsl_ctx *ctx) {
  CHECK_CTX();
  int rc = require_grid(ctx);
  if (rc) return rc;
  FSL_HIP(fsl::launch_save(ctx));
  return FSL_OK;
}

int fsl_delete_outside(fsl_ctx *ctx) {
  CHECK_CTX();
  ctx->outside_deleted = true;
  return FSL_OK;
}

int fsl_add_gravity(fsl_ctx *ctx, float dt) {
  CHECK_CTX();
  int rc = require_grid(ctx);
  if (rc) return rc;
  FSL_HIP(fsl::launch_gravity(ctx, dt));
  return FSL_OK;
}

int fsl_divergence(fsl_ctx *ctx) {
  CHECK_CTX();
  int rc = require_grid(ctx);
  if (rc) return rc;
  FSL_HIP(hipMemsetAsync(ctx->lfluid, 0, size_t(ctx->nleaves) * sizeof(uint32_t), ctx->stream));
  FSL_HIP(hipMemsetAsync(ctx->flags, 0, sizeof(fsl::DevFlags), ctx->stream));
  FSL_HIP(fsl::launch_divergence(ctx));
  FSL_HIP(fsl::launch_compact(ctx, ctx->lfluid, ctx->fluid_list, ctx->dcount + fsl::kCountFluid));
  // The fluid-leaf count and the flags in one round trip.
  FSL_HIP(hipMemcpyAsync(ctx->host_u32, ctx->lscan + ctx->nleaves, sizeof(uint32_t),
                         hipMemcpyDeviceToHost, ctx->stream));
  FSL_HIP(hipMemcpyAsync(ctx->flags_host, ctx->flags, sizeof(fsl::DevFlags),
                         hipMemcpyDeviceToHost, ctx->stream));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  ctx->n_fluid = ctx->host_u32[0];
  ctx->pcg_ready = false;
  return FSL_OK;
}

int fsl_pcg_solve(fsl_ctx *ctx, const fsl_solve_params *params, fsl_solve_stats *stats) {
  CHECK_CTX();
  if (!params) return FSL_ERR_ARG;
  if (ctx->flags_host->zero_diag) {
    ctx->err = "FLUID cell with no non-SOLID neighbour (zero Jacobi diagonal)";
    return FSL_ERR_ZERO_DIAGONAL;
  }
  fsl_solve_stats local{};
  fsl_solve_stats *st = stats ? stats : &local;
  *st = fsl_solve_stats{};
  st->n_rows = int(ctx->flags_host->n_rows);
  st->n_cols = st->n_rows;
  int rc = fsl::pcg_run(ctx, params, st);
  if (rc) {
    if (ctx->err.empty()) ctx->err = "solver failed";
    st->iterations = ctx->pcg_host->it;
  }
  return rc;
}

int fsl_pressure_update(fsl_ctx *ctx) {
  CHECK_CTX();
  int rc = require_grid(ctx);
  if (rc) return rc;
  FSL_HIP(fsl::launch_pressure_update(ctx));
  return FSL_OK;
}

int fsl_velocity_update(fsl_ctx *ctx) {
  CHECK_CTX();
  int rc = require_grid(ctx);
  if (rc) return rc;
  FSL_HIP(hipMemsetAsync(&ctx->flags->vel_update, 0, sizeof(unsigned), ctx->stream));
  FSL_HIP(fsl::launch_velocity_update(ctx));
  uint32_t bad = 0;
  rc = read_u32(ctx, &ctx->flags->vel_update, &bad);
  if (rc) return rc;
  if (bad) {
    ctx->err = "face active in grid_velocity_ but not in grid_velocity_update_";
    return FSL_ERR_VELOCITY_UPDATE;
  }
  return FSL_OK;
}

int fsl_extrapolate(fsl_ctx *ctx) {
  CHECK_CTX();
  int rc = require_grid(ctx);
  if (rc) return rc;
  FSL_HIP(fsl::launch_extrapolate(ctx));
  return FSL_OK;
}

int fsl_g2p(fsl_ctx *ctx) {
  CHECK_CTX();
  int rc = require_grid(ctx);
  if (rc) return rc;
  FSL_HIP(fsl::launch_extrapolate(ctx));
  FSL_HIP(fsl::launch_g2p(ctx));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  return FSL_OK;
}

int fsl_step(fsl_ctx *ctx, float dt, const fsl_solve_params *params, fsl_solve_stats *stats) {
  CHECK_CTX();
  if (!params) return FSL_ERR_ARG;
  int rc = run_step(ctx, dt, nullptr, params, stats);
  if (rc) return rc;
  return check_step_flags(ctx);
}

int fsl_advance_frame(fsl_ctx *ctx, int cfl, float frame_step, const fsl_solve_params *params,
                      fsl_frame_stats *stats) {
  CHECK_CTX();
  if (!params || !(frame_step > 0.0f) || cfl <= 0) return FSL_ERR_ARG;
  fsl_frame_stats local{};
  fsl_frame_stats *fs = stats ? stats : &local;
  *fs = fsl_frame_stats{};
  // WhileAndComputeDt's constants, in the host's float/double arithmetic
  // (fsl_compute_dt).
  fsl::DtParams p;
  p.v_gravity = std::sqrt(0.5 * double(ctx->gravity));
  p.snap = 0.9 * double(frame_step);
  p.cfl = float(cfl);
  p.frame_step = frame_step;
  p.dt_max = float(1.0 / double(frame_step));
  p.dt_min = frame_step / 10;
  int rc;
  // The frame's elapsed time starts at 0.
  FSL_HIP(fsl::launch_fill_u32(ctx, reinterpret_cast<uint32_t *>(&ctx->sstate->step), 1, 0u));
  for (int s = 0;; ++s) {
    if (s >= 100000) {
      ctx->err = "frame did not end";
      return FSL_ERR_STATE;
    }
    // The front half (CFL and dt on the device, then the stages up to the
    // divergence), replayed from a graph once the grids exist and the
    // timing hook is off.
    if (ctx->use_graph && ctx->grid_ready && !ctx->timing && ctx->outside_deleted)
      rc = step_front_graph(ctx, p);
    else
      rc = step_front(ctx, 0.0f, &ctx->sstate->dt, true, false, &p);
    if (rc) return rc;
    fsl_solve_stats st{};
    if ((rc = step_back(ctx, params, &st))) return rc;
    fs->steps += 1;
    fs->pcg_iterations += st.iterations;
    fs->dt_last = ctx->sstate_host->dt;
    fs->last_solve = st;
    if (!ctx->sstate_host->more) break;
  }
  return check_step_flags(ctx);
}

int fsl_set_step_graph(fsl_ctx *ctx, int enable) {
  CHECK_CTX();
  ctx->use_graph = enable != 0;
  return FSL_OK;
}

int fsl_set_timing(fsl_ctx *ctx, int enable) {
  CHECK_CTX();
  if (enable && !ctx->tev[0]) {
    for (hipEvent_t &e : ctx->tev) FSL_HIP(hipEventCreate(&e));
  }
  ctx->timing = enable != 0;
  for (double &t : ctx->tacc) t = 0.0;
  return FSL_OK;
}

int fsl_get_timing(fsl_ctx *ctx, double *out, int n) {
  CHECK_CTX();
  if (!out || n < 0) return FSL_ERR_ARG;
  for (int q = 0; q < n && q < FSL_T_COUNT; ++q) out[q] = ctx->tacc[q];
  return FSL_OK;
}

int fsl_compute_dt(fsl_ctx *ctx, int cfl, float frame_step, float *step, float *dt) {
  CHECK_CTX();
  if (!step || !dt) return FSL_ERR_ARG;
  float inv = 0.0f;
  int rc = fsl_cfl(ctx, &inv);
  if (rc) return rc;
  const float local_dt = float(cfl) * inv;
  float g = float(1.0 / double(frame_step));
  g = std::min(g, local_dt);
  if (g < frame_step / 10) g = frame_step / 10;
  if (*step + g > frame_step) g = frame_step - *step;
  if (double(*step + g) > 0.9 * double(frame_step)) g = frame_step - *step;
  *step += g;
  *dt = g;
  return FSL_OK;
}

int fsl_get_vec(fsl_ctx *ctx, int which, float *out, uint8_t *active) {
  CHECK_CTX();
  if (!out || (which != FSL_GRID_VELOCITY && which != FSL_GRID_VELOCITY_UPDATE)) return FSL_ERR_ARG;
  if (fsl::on_ctx_device(ctx, out) && (!active || fsl::on_ctx_device(ctx, active))) {
    // Device arrays: written in place on the context's stream.
    FSL_HIP(fsl::launch_box_vec(ctx, which, out, active, true));
    FSL_HIP(hipStreamSynchronize(ctx->stream));
    return FSL_OK;
  }
  const int64_t nb = box_count(ctx);
  float *dbox = nullptr;
  uint8_t *dact = nullptr;
  FSL_HIP(hipMalloc(reinterpret_cast<void **>(&dbox), size_t(nb) * 3 * sizeof(float)));
  FSL_HIP(hipMalloc(reinterpret_cast<void **>(&dact), size_t(nb)));
  FSL_HIP(fsl::launch_box_vec(ctx, which, dbox, dact, true));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  FSL_HIP(hipMemcpy(out, dbox, size_t(nb) * 3 * sizeof(float), hipMemcpyDeviceToHost));
  if (active) FSL_HIP(hipMemcpy(active, dact, size_t(nb), hipMemcpyDeviceToHost));
  (void)hipFree(dbox);
  (void)hipFree(dact);
  return FSL_OK;
}

int fsl_set_vec(fsl_ctx *ctx, int which, const float *in, const uint8_t *active) {
  CHECK_CTX();
  if (!in || !active || (which != FSL_GRID_VELOCITY && which != FSL_GRID_VELOCITY_UPDATE)) return FSL_ERR_ARG;
  int rc = all_leaves(ctx);
  if (rc) return rc;
  if (fsl::on_ctx_device(ctx, in) && fsl::on_ctx_device(ctx, active)) {
    // Device arrays: read in place on the context's stream.
    FSL_HIP(fsl::launch_box_vec(ctx, which, const_cast<float *>(in), const_cast<uint8_t *>(active),
                                false));
    FSL_HIP(hipStreamSynchronize(ctx->stream));
    return FSL_OK;
  }
  const int64_t nb = box_count(ctx);
  float *dbox = nullptr;
  uint8_t *dact = nullptr;
  FSL_HIP(hipMalloc(reinterpret_cast<void **>(&dbox), size_t(nb) * 3 * sizeof(float)));
  FSL_HIP(hipMalloc(reinterpret_cast<void **>(&dact), size_t(nb)));
  FSL_HIP(hipMemcpy(dbox, in, size_t(nb) * 3 * sizeof(float), hipMemcpyHostToDevice));
  FSL_HIP(hipMemcpy(dact, active, size_t(nb), hipMemcpyHostToDevice));
  FSL_HIP(fsl::launch_box_vec(ctx, which, dbox, dact, false));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  (void)hipFree(dbox);
  (void)hipFree(dact);
  return FSL_OK;
}

int fsl_get_scalar(fsl_ctx *ctx, int which, float *out) {
  CHECK_CTX();
  if (!out || (which != FSL_GRID_DIVERGENCE && which != FSL_GRID_PRESSURE)) return FSL_ERR_ARG;
  const int64_t nb = box_count(ctx);
  float *dbox = nullptr;
  FSL_HIP(hipMalloc(reinterpret_cast<void **>(&dbox), size_t(nb) * sizeof(float)));
  FSL_HIP(fsl::launch_box_scalar(ctx, which == FSL_GRID_DIVERGENCE ? ctx->div : ctx->pres, dbox, true));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  FSL_HIP(hipMemcpy(out, dbox, size_t(nb) * sizeof(float), hipMemcpyDeviceToHost));
  (void)hipFree(dbox);
  return FSL_OK;
}

int fsl_get_scalar_noalloc(fsl_ctx *ctx, int which, float *out) {
  CHECK_CTX();
  if (!out || (which != FSL_GRID_DIVERGENCE && which != FSL_GRID_PRESSURE)) return FSL_ERR_ARG;
  const int64_t nb = box_count(ctx);
  if (nb > kScratchFloats) {
    ctx->err = "fsl_get_scalar_noalloc: the box has " + std::to_string(nb) +
               " voxels, more than its staging array's " + std::to_string(kScratchFloats) +
               " (domains up to 200^3); use fsl_get_scalar";
    return FSL_ERR_ARG;
  }
  if (!ctx->scratch) {
    ctx->err = "fsl_get_scalar_noalloc: no staging array on this device";
    return FSL_ERR_STATE;
  }
  FSL_HIP(fsl::launch_box_scalar(ctx, which == FSL_GRID_DIVERGENCE ? ctx->div : ctx->pres,
                                 ctx->scratch, true));
  FSL_HIP(hipMemcpyAsync(out, ctx->scratch, size_t(nb) * sizeof(float), hipMemcpyDeviceToHost,
                         ctx->stream));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  return FSL_OK;
}

int fsl_set_scalar(fsl_ctx *ctx, int which, const float *in) {
  CHECK_CTX();
  if (!in || (which != FSL_GRID_DIVERGENCE && which != FSL_GRID_PRESSURE)) return FSL_ERR_ARG;
  int rc = all_leaves(ctx);
  if (rc) return rc;
  const int64_t nb = box_count(ctx);
  float *dbox = nullptr;
  FSL_HIP(hipMalloc(reinterpret_cast<void **>(&dbox), size_t(nb) * sizeof(float)));
  FSL_HIP(hipMemcpy(dbox, in, size_t(nb) * sizeof(float), hipMemcpyHostToDevice));
  FSL_HIP(fsl::launch_box_scalar(ctx, which == FSL_GRID_DIVERGENCE ? ctx->div : ctx->pres, dbox, false));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  (void)hipFree(dbox);
  return FSL_OK;
}

int fsl_get_int(fsl_ctx *ctx, int which, int32_t *out) {
  CHECK_CTX();
  if (!out || which < 0 || which > 3) return FSL_ERR_ARG;
  const int64_t nb = box_count(ctx);
  int32_t *dbox = nullptr;
  FSL_HIP(hipMalloc(reinterpret_cast<void **>(&dbox), size_t(nb) * sizeof(int32_t)));
  if (which == FSL_GRID_CELL_COUNT) {
    FSL_HIP(fsl::launch_cell_counts_box(ctx, dbox));
  } else {
    FSL_HIP(fsl::launch_box_u8(ctx, which == FSL_GRID_SOLID ? ctx->solid : ctx->marker, dbox, true));
  }
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  FSL_HIP(hipMemcpy(out, dbox, size_t(nb) * sizeof(int32_t), hipMemcpyDeviceToHost));
  (void)hipFree(dbox);
  if (which == FSL_GRID_MARKER || which == FSL_GRID_ROW_INDEX) {
    // Leaves off the active list are never read, but the reference marks
    // every solid voxel SOLID once boundary conditions ran.
    if (ctx->bc_applied) {
      std::vector<int32_t> s(static_cast<size_t>(nb));
      int rc = fsl_get_int(ctx, FSL_GRID_SOLID, s.data());
      if (rc) return rc;
      for (int64_t q = 0; q < nb; ++q) if (s[size_t(q)]) out[q] = fsl::kSolid;
    }
  }
  if (which == FSL_GRID_ROW_INDEX) {
    // CreateIndexAndMaskTree (solver.cc:114-151): lexicographic order over
    // bbox + 1, interior FLUID cells first (single partition: all of them).
    int32_t row = 0;
    for (int64_t q = 0; q < nb; ++q) out[q] = out[q] == fsl::kFluid ? row++ : -1;
  }
  return FSL_OK;
}

int fsl_set_marker(fsl_ctx *ctx, const int32_t *in) {
  CHECK_CTX();
  if (!in) return FSL_ERR_ARG;
  int rc = all_leaves(ctx);
  if (rc) return rc;
  const int64_t nb = box_count(ctx);
  int32_t *dbox = nullptr;
  FSL_HIP(hipMalloc(reinterpret_cast<void **>(&dbox), size_t(nb) * sizeof(int32_t)));
  FSL_HIP(hipMemcpy(dbox, in, size_t(nb) * sizeof(int32_t), hipMemcpyHostToDevice));
  FSL_HIP(fsl::launch_box_u8(ctx, ctx->marker, dbox, false));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  (void)hipFree(dbox);
  return FSL_OK;
}

int fsl_get_active_leaves(fsl_ctx *ctx, int which, int32_t *origins, int64_t cap, int64_t *count) {
  CHECK_CTX();
  if (!count) return FSL_ERR_ARG;
  const int64_t n = which == 0 ? ctx->n_act : ctx->n_fluid;
  *count = n;
  if (!origins) return FSL_OK;
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<int32_t> list(static_cast<size_t>(n));
  if (n) FSL_HIP(hipMemcpy(list.data(), which == 0 ? ctx->act_list : ctx->fluid_list,
                           size_t(n) * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int64_t q = 0; q < n && q < cap; ++q) {
    int o[3];
    fsl::leaf_origin(ctx->dom, list[size_t(q)], o);
    origins[3 * q] = o[0]; origins[3 * q + 1] = o[1]; origins[3 * q + 2] = o[2];
  }
  return FSL_OK;
}

int fsl_tile_rows(fsl_ctx *ctx, int64_t *rows_per_tile) {
  CHECK_CTX();
  if (!rows_per_tile) return FSL_ERR_ARG;
  const int nt = ctx->dom.T[0] * ctx->dom.T[1] * ctx->dom.T[2];
  for (int t = 0; t < nt; ++t) rows_per_tile[t] = int64_t(ctx->flags_host->tile_rows[t]);
  return FSL_OK;
}

int fsl_pcg_prepare(fsl_ctx *ctx, const fsl_solve_params *params, double *rr, double *rz) {
  CHECK_CTX();
  if (!params) return FSL_ERR_ARG;
  if (ctx->flags_host->zero_diag) return FSL_ERR_ZERO_DIAGONAL;
  int rc = fsl::pcg_prepare(ctx, params);
  if (rr) *rr = ctx->pcg_host->rr;
  if (rz) *rz = ctx->pcg_host->rz;
  return rc;
}

int fsl_pcg_local1(fsl_ctx *ctx, double *pAp) {
  CHECK_CTX();
  return fsl::pcg_local1(ctx, pAp);
}

int fsl_pcg_local2(fsl_ctx *ctx, double *alpha, double *rr, double *rz) {
  CHECK_CTX();
  return fsl::pcg_local2(ctx, alpha, rr, rz);
}

int fsl_pcg_local3(fsl_ctx *ctx, double *beta) {
  CHECK_CTX();
  return fsl::pcg_local3(ctx, beta);
}

int fsl_pcg_iterate(fsl_ctx *ctx, double *pAp, double *alpha, double *rr, double *rz,
                    double *beta) {
  CHECK_CTX();
  int rc = fsl::pcg_iterate_once(ctx);
  const fsl::PcgState *h = ctx->pcg_host;
  if (pAp) *pAp = h->pAp;
  if (alpha) *alpha = h->alpha;
  if (rr) *rr = h->rr;
  if (rz) *rz = h->rz;
  if (beta) *beta = h->beta;
  return rc;
}

int fsl_pcg_finish(fsl_ctx *ctx, fsl_solve_stats *stats) {
  CHECK_CTX();
  if (stats) {
    *stats = fsl_solve_stats{};
    stats->n_rows = int(ctx->flags_host->n_rows);
    stats->n_cols = stats->n_rows;
  }
  return fsl::pcg_finish(ctx, stats);
}

int fsl_project_forward(fsl_ctx *ctx, const fsl_solve_params *params, fsl_solve_stats *stats) {
  CHECK_CTX();
  int rc;
  if ((rc = fsl_divergence(ctx))) return rc;
  if ((rc = fsl_pcg_solve(ctx, params, stats))) return rc;
  return fsl_pressure_update(ctx);
}

int fsl_project_backward(fsl_ctx *ctx, const fsl_solve_params *params, const float *grad_out,
                         float *grad_in, fsl_solve_stats *stats) {
  CHECK_CTX();
  if (!params || !grad_out || !grad_in) return FSL_ERR_ARG;
  return fsl::project_backward(ctx, params, grad_out, grad_in, stats);
}

}  // extern "C"
