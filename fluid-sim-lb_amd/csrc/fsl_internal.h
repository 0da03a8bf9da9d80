// Internal declarations shared by the HIP translation units of libfsl.
//
// Storage model. The domain [0,n)^3 plus every voxel the reference can touch
// ([-1, n] on each axis) is tiled by 8^3 leaves; leaf (a,b,c) holds voxels
// [8a-8, 8a) x [8b-8, 8b) x [8c-8, 8c), so there are L = n/8 + 2 leaves per
// axis. Every field is a dense array over all leaves in leaf-major order
// (512 voxels per leaf, voxel offset (x<<6)|(y<<3)|z, the OpenVDB LeafNode
// order). Work is proportional to the active-leaf list, not to the box:
// kernels run one workgroup per listed leaf, and leaves off the list hold
// zeros, which is the OpenVDB background value, so a read anywhere returns
// what the reference's tree would return.
#ifndef FSL_INTERNAL_H_
#define FSL_INTERNAL_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "fsl.h"

namespace fsl {

struct AdjState;  // adjoint buffers and rollout checkpoints (adjoint.hip)
constexpr int kLeaf = 512;           // voxels per leaf
constexpr uint32_t kDead = 0xFFFFFFFFu;  // particle key of a deleted particle
constexpr int kAir = 0, kFluid = 1, kSolid = 2;

struct Dom {
  int n[3];   // cells per axis
  int L[3];   // leaves per axis, n/8 + 2
  int T[3];   // tiles (reference partitions) per axis
  // Rank of each leaf in VDB traversal order (root tiles of 64^3 by origin,
  // then the leaf's offset in its tile): the particle sort orders cells by
  // (rank, voxel offset), the reference's particle traversal order, and the
  // per-cell particle ranges vstart are indexed that way (sort_index).
  const int32_t *lrank;
};

__host__ __device__ inline int64_t vox_of(const Dom &d, int i, int j, int k) {
  const int a = i + 8, b = j + 8, c = k + 8;
  const int64_t leaf = (int64_t(a >> 3) * d.L[1] + (b >> 3)) * d.L[2] + (c >> 3);
  return (leaf << 9) | ((a & 7) << 6) | ((b & 7) << 3) | (c & 7);
}
// Index of voxel v (storage order) in the particle sort order.
__device__ inline int64_t sort_index(const Dom &d, int64_t v) {
  return (int64_t(d.lrank[v >> 9]) << 9) | (v & 511);
}
// First sort index of a leaf's voxels (its range ends at sort_begin + 512).
__device__ inline int64_t sort_begin(const Dom &d, int64_t leaf) {
  return int64_t(d.lrank[leaf]) << 9;
}
// VDB traversal key of a leaf origin (cell coordinates, multiples of 8):
// root tile of 64^3 in lexicographic order of its origin, then the leaf's
// offset in the tile, x-major.
inline uint64_t vdb_leaf_key(int i, int j, int k) {
  const uint64_t ri = uint64_t((i >> 6) + 1024), rj = uint64_t((j >> 6) + 1024),
                 rk = uint64_t((k >> 6) + 1024);
  const uint64_t off = uint64_t(((i >> 3) & 7) << 6 | ((j >> 3) & 7) << 3 | ((k >> 3) & 7));
  return ((ri * 2048 + rj) * 2048 + rk) << 9 | off;
}

__host__ __device__ inline bool in_store(const Dom &d, int i, int j, int k) {
  return i >= -8 && j >= -8 && k >= -8 && i < d.n[0] + 8 && j < d.n[1] + 8 && k < d.n[2] + 8;
}
__host__ __device__ inline bool in_bbox(const Dom &d, int i, int j, int k) {
  return i >= 0 && j >= 0 && k >= 0 && i < d.n[0] && j < d.n[1] && k < d.n[2];
}
__host__ __device__ inline void leaf_origin(const Dom &d, int64_t leaf, int o[3]) {
  const int64_t l12 = int64_t(d.L[1]) * d.L[2];
  o[0] = int(leaf / l12) * 8 - 8;
  o[1] = int((leaf / d.L[2]) % d.L[1]) * 8 - 8;
  o[2] = int(leaf % d.L[2]) * 8 - 8;
}
__host__ __device__ inline int64_t leaf_count(const Dom &d) {
  return int64_t(d.L[0]) * d.L[1] * d.L[2];
}
// Reference partition rank of a cell (MetaGrid::ComputeRank, metagrid.cc:84-87).
__host__ __device__ inline int tile_of(const Dom &d, int i, int j, int k) {
  int c[3] = {i, j, k};
  int t[3];
  for (int a = 0; a < 3; ++a) {
    int w = d.n[a] / d.T[a];
    int x = c[a] < 0 ? 0 : (c[a] >= d.n[a] ? d.n[a] - 1 : c[a]);
    t[a] = x / w;
  }
  return t[0] + d.T[0] * (t[1] + d.T[1] * t[2]);
}

// Scalars of one PCG solve, resident on the device.
struct PcgState {
  double rz, rz_old, rr, pAp, alpha, beta;
  double bnorm2, thr, rr_true;
  double rzh[2];  // rz of iteration k in slot k & 1
  double alph[2]; // alpha of iteration k in slot k & 1 (single-reduction CG)
  int it, done, status, max_iter;
  int stop_mode, precision, refresh_due;
  int pc_fail;    // the IC(0) build met a pivot below 1e-5
  // Single-reduction path: the iteration at which a k_cg launch stopped
  // (INT_MAX while running). Launches of later iterations exit on it; the
  // launch that stops still runs in every workgroup, since a workgroup may
  // start after the stopping decision was stored.
  int stop_k;
  unsigned cnt_a, cnt_b, cnt_r, cnt_x;
  // The iteration a convergence check enqueued without a host read
  // refreshed (the stopping iteration, or the next one when still running).
  int refresh_kn, pad;
};

// Device-side error flags of the grid stages.
// Particle deletion boxes applied by the advection kernel (positions after
// the move; bounds inclusive, as Cube::IsInside, primitives.cc:57-64).
constexpr int kMaxKillBoxes = 16;
struct KillBoxes {
  int n = 0;
  float lo[kMaxKillBoxes][3], hi[kMaxKillBoxes][3];
};
// One evaluated water source (AddContributionFromSources,
// flip_simulation.cc:1206-1266): the cells c0..c1 (inclusive, inside the
// domain), the box used for the inside test, the velocity, the particles per
// cell and the Philox key.
struct SourceBox {
  int c0[3], c1[3];
  float lo[3], hi[3], vel[3];
  int ppc;
  uint32_t k0, k1;
};

// Time step on the device (fsl_advance_frame): the frame's elapsed time and
// the dt of the current step, from the CFL maximum of the grid velocity.
struct StepState {
  float dt;       // dt of the current step
  float step;     // time elapsed in the frame after the current step
  int more;       // another step follows in this frame
  int pad;
};
// Constants of WhileAndComputeDt (flip_app_sim.cc:357-410) for k_dt, formed
// on the host in the reference's float/double arithmetic.
struct DtParams {
  double v_gravity;  // sqrt(0.5 * gravity) (flip_simulation.cc:485)
  double snap;       // 0.9 * frame_step
  float cfl, frame_step, dt_max, dt_min;  // dt_max = 1 / frame_step, dt_min = frame_step / 10
};

// A captured front half of a device-resident step (api.hip) and what it
// was captured with.
struct StepGraph {
  hipGraphExec_t exec = nullptr;
  int64_t np = 0;
  DtParams dtp{};
  bool outside_deleted = false;
};

struct DevFlags {
  unsigned zero_diag;
  unsigned vel_update;
  unsigned cfl_bits;
  unsigned pad;
  unsigned long long n_rows;
  unsigned long long tile_rows[64];
};

struct Particles {
  float *a[9];  // px py pz vx vy vz gx gy gz
};
// Particle-sized arrays that are read four elements at a time (the particle
// arrays and the adjoint's particle values in P2G, the sort permutation in
// the in-cell sort) hold this many elements beyond the capacity: the loads
// reach up to three elements past a leaf's last particle.
constexpr int64_t kParticlePad = 4;

// A leaf list on the device, list[0, n). When `count` is set the length is
// the value there (written on the device by the compaction that built the
// list) and n is only an upper bound of it, used to size launches, so work
// can be enqueued before the host knows the length.
struct LeafList {
  const int32_t *list;
  const uint32_t *count;
  int64_t n;
};

// Device-resident counts (fsl_ctx::dcount).
constexpr int kCountParticles = 0;  // particles alive after the last sort
constexpr int kCountActive = 1;     // entries of act_list
constexpr int kCountFluid = 2;      // entries of fluid_list

}  // namespace fsl

struct fsl_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  fsl::Dom dom{};
  int64_t nleaves = 0, nvox = 0;
  float gravity = 9.8f, force[3] = {0, 9.8f, 0}, flip = 0.95f;
  // Grids.
  float *v[3] = {}, *u[3] = {};
  uint8_t *vflag = nullptr;  // bit0: active in grid_velocity_, bit1: in grid_velocity_update_
  uint8_t *marker = nullptr, *solid = nullptr, *lab = nullptr;
  float *div = nullptr, *pres = nullptr;
  // PCG vectors (sized for doubles; FP32 modes use the first half).
  void *x = nullptr, *r = nullptr, *z = nullptr, *p[2] = {}, *Ap = nullptr;
  double *x64 = nullptr;
  void *partials = nullptr;
  fsl::PcgState *pcg = nullptr;
  fsl::PcgState *pcg_host = nullptr;  // pinned mirror; entries 1 and 2: state after pipelined batches
  int pcg_precision = FSL_PCG_FP32, pcg_stop = FSL_STOP_ABS_RR, pcg_refresh = 32;
  int pcg_pbuf = 0;            // which p buffer holds the current direction
  bool pcg_ready = false;
  bool pcg_sr = false;         // the last solve ran the single-reduction iteration
  bool pcg_roworder = false;   // dot products in the oracle's row order (two-pass)
  int pcg_stage = 0;           // stage API: next stage of the iteration (0, 1, 2 = Local1..3)
  // Block IC(0) preconditioner (allocated by the first IC solve): L's
  // diagonal d and -1/d per voxel (sized for doubles), the FLUID leaves
  // sorted by tile-local level, and the level offsets into that list.
  bool pcg_ic = false;
  void *ic_d = nullptr, *ic_o = nullptr;
  int32_t *ic_leaves = nullptr;
  std::vector<int> ic_off;
  // Leaves.
  uint32_t *lhas = nullptr;   // leaf holds particles
  uint32_t *lact = nullptr;   // leaf active (26-dilation of lhas)
  uint32_t *lfluid = nullptr; // FLUID rows in the leaf (nonzero: a fluid leaf)
  uint32_t *lscan = nullptr;  // scan scratch over leaves (nleaves + 1)
  int32_t *act_list = nullptr, *fluid_list = nullptr;
  int32_t *morton = nullptr;   // leaf indices in Morton (Z-curve) order of leaf coordinates
  int32_t *lrank = nullptr;    // leaf -> rank in VDB traversal order (dom.lrank)
  uint32_t *lflag_m = nullptr; // flags gathered in Morton order (scan input)
  int64_t n_act = 0, n_fluid = 0;
  // Device-resident step (fsl_advance_frame): the counts are kept on the
  // device (dcount, kCount*), and n_act, n_fluid and np are then upper
  // bounds until the host reads the counts back (sync_counts).
  uint32_t *dcount = nullptr;
  bool counts_on_device = false;
  fsl::StepState *sstate = nullptr, *sstate_host = nullptr;  // device and pinned mirror
  bool use_graph = false;          // fsl_advance_frame replays its steps' front halves
  fsl::StepGraph graphs[2];        // by particle buffer (which one P is)
  float *pbuf0 = nullptr;          // the first particle buffer's px array
  float *scratch = nullptr;        // device address of the staging array (api.hip)
  // The previous single-reduction solve's iteration count: with use_graph,
  // the next solve enqueues that many launches at once (pcg.hip).
  int pcg_last_iters = 0;
  hipEvent_t ev_finish = nullptr;  // the solver state's copy after the solve's finish
  hipEvent_t ev_counts = nullptr;  // the counts' copy within a device-resident step
  bool bc_applied = false, grid_ready = false, outside_deleted = false;
  // Particles.
  int64_t cap = 0, np = 0;
  fsl::Particles P{}, Q{};
  uint32_t *key = nullptr, *perm = nullptr;
  // P2G scratch: weight sums of the own cells (3 * nvox) and the ring blocks
  // that carry contributions to neighbouring leaves.
  float *p2g_w = nullptr;
  void *p2g_ring = nullptr;
  uint32_t *vcount = nullptr, *vstart = nullptr;  // per voxel; vstart has nvox + 1
  uint32_t *scan_tmp = nullptr;
  int64_t scan_tmp_len = 0;
  // Per-cell counts and offsets of a water source (grown on demand).
  uint32_t *src_buf = nullptr;
  int64_t src_cap = 0;
  // Flags and staging.
  fsl::DevFlags *flags = nullptr, *flags_host = nullptr;
  uint32_t *host_u32 = nullptr;  // pinned scratch for small readbacks
  // Measurement hook.
  bool timing = false;
  double tacc[FSL_T_COUNT] = {};
  hipEvent_t tev[240] = {};  // [0,32): stage pairs, [32,232): 5 per PCG iteration
  std::string err;
  fsl::AdjState *adj = nullptr;  // allocated by the first adjoint call
};

namespace fsl {

// api.hip: reads the device-resident counts back (after a device-resident
// step), so the host's n_act, n_fluid and np are exact again.
int host_counts(fsl_ctx *ctx);

// The active and FLUID leaf lists as the kernels take them.
inline LeafList act_leaves(const fsl_ctx *c) {
  return {c->act_list, c->counts_on_device ? c->dcount + kCountActive : nullptr, c->n_act};
}
inline LeafList fluid_leaves(const fsl_ctx *c) {
  return {c->fluid_list, c->counts_on_device ? c->dcount + kCountFluid : nullptr, c->n_fluid};
}

// scan.hip
int64_t scan_tmp_elems(int64_t n);
hipError_t exclusive_scan(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *tmp,
                          hipStream_t s);

// particles.hip
hipError_t launch_seed(fsl_ctx *c, const uint8_t *mask_dev, const uint32_t *offs, int ppc,
                       uint32_t seed, const float vel[3]);
hipError_t launch_advect(fsl_ctx *c, float dt, const float *dt_dev, bool drop_outside,
                         const KillBoxes &kill);
hipError_t launch_source_count(fsl_ctx *c, const SourceBox &s, int64_t ncell, uint32_t *cnt);
hipError_t launch_source_emit(fsl_ctx *c, const SourceBox &s, int64_t ncell, const uint32_t *off,
                              int64_t base);
hipError_t launch_sort_tail(fsl_ctx *c, uint32_t n_alive, bool keep_gv = true);
hipError_t launch_mask_count(fsl_ctx *c, const uint8_t *mask_dev, int64_t n, uint32_t *cnt);
hipError_t launch_leaf_flags(fsl_ctx *c);
hipError_t launch_compact(fsl_ctx *c, const uint32_t *flag, int32_t *list, uint32_t *count);
size_t p2g_ring_bytes(int64_t nleaves);
hipError_t launch_p2g(fsl_ctx *c, float *const wsum[3] = nullptr);
hipError_t launch_g2p_transpose(fsl_ctx *c, float *const pval[3], float *const out[3]);
hipError_t launch_p2g_transpose(fsl_ctx *c, float *const g[3], float *const wsum[3], float keep,
                                float *const io[3]);
hipError_t launch_g2p(fsl_ctx *c);
hipError_t launch_cell_counts_box(fsl_ctx *c, int32_t *out_dev);

// grid.hip
hipError_t launch_clear_leaves(fsl_ctx *c, const LeafList &L);
hipError_t launch_bc(fsl_ctx *c);
hipError_t launch_save(fsl_ctx *c);
hipError_t launch_bc_save(fsl_ctx *c);
hipError_t launch_gravity(fsl_ctx *c, float dt, const float *dt_dev = nullptr);
hipError_t launch_divergence(fsl_ctx *c);
hipError_t launch_pressure_update(fsl_ctx *c);
hipError_t launch_velocity_update(fsl_ctx *c);
hipError_t launch_extrapolate(fsl_ctx *c);
hipError_t launch_clear_shell(fsl_ctx *c);
hipError_t launch_cfl(fsl_ctx *c);
// dt of the next step on the device, from the CFL maximum in flags->cfl_bits.
hipError_t launch_dt(fsl_ctx *c, const DtParams &p);
hipError_t launch_fill_u32(fsl_ctx *c, uint32_t *p, int64_t n, uint32_t value);
hipError_t launch_box_vec(fsl_ctx *c, int which, float *out3, uint8_t *act, bool to_box);
// True when p points into device memory of the context's device (a caller's
// array that kernels can read or write directly); false for host memory.
bool on_ctx_device(const fsl_ctx *c, const void *p);
hipError_t launch_box_scalar(fsl_ctx *c, float *grid, float *box, bool to_box);
hipError_t launch_box_u8(fsl_ctx *c, uint8_t *grid, int32_t *box, bool to_box);
hipError_t launch_all_leaves(fsl_ctx *c);

// pcg.hip
size_t pcg_partials_bytes();
int pcg_prepare(fsl_ctx *c, const fsl_solve_params *prm);
int pcg_run(fsl_ctx *c, const fsl_solve_params *prm, fsl_solve_stats *st);
int pcg_iterate_once(fsl_ctx *c);
int pcg_local1(fsl_ctx *c, double *pAp);
int pcg_local2(fsl_ctx *c, double *alpha, double *rr, double *rz);
int pcg_local3(fsl_ctx *c, double *beta);
int pcg_finish(fsl_ctx *c, fsl_solve_stats *st);
// pcg_finish in two parts: the solution copy and the true-residual kernels
// with the state copied to the host (enqueued), then the wait for that copy
// and the statistics; work enqueued in between overlaps the wait.
int pcg_finish_enqueue(fsl_ctx *c);
int pcg_finish_collect(fsl_ctx *c, fsl_solve_stats *st);
// The solve without its finish (pcg_run = pcg_solve_iterations + pcg_finish).
int pcg_solve_iterations(fsl_ctx *c, const fsl_solve_params *prm, fsl_solve_stats *st);
int pcg_true_residual(fsl_ctx *c, double *rr_true);

// ic.hip: block IC(0). ic_factor sorts the FLUID leaves by level and factors
// (once per solve, after the labels); ic_sweeps enqueues z = M^-1 r (y in Ap);
// force: run even when the solve is done (a forced refresh).
int ic_factor(fsl_ctx *ctx);
void ic_sweeps(fsl_ctx *c, int force);
void ic_free(fsl_ctx *c);

// adjoint.hip
int project_backward(fsl_ctx *c, const fsl_solve_params *prm, const float *grad_out,
                     float *grad_in, fsl_solve_stats *st);
void adjoint_free(fsl_ctx *c);

}  // namespace fsl

#define FSL_HIP(expr)                                                        \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);          \
      return FSL_ERR_HIP;                                                    \
    }                                                                        \
  } while (0)

#endif  // FSL_INTERNAL_H_
