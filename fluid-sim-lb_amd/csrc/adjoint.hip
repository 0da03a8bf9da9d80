// Adjoint of the pressure projection, and of FLIP substeps with the particle
// positions, labels and active faces of a recorded forward trajectory.
//
// Projection. With D the divergence on FLUID rows (flip_simulation.cc:715-736),
// G = M G~ the pressure-gradient update (flip_simulation.cc:738-778; G~ the
// plain difference p(c) - p(c - e_d) with p = 0 off the rows, M the faces the
// update writes) and A = -D M G~ the assembled Laplacian
// (solver.cc:197-233), the projection is u' = u + G A^-1 D u. A is symmetric
// and D^T = -G~, so
//     P^T g = g + G~ A^-1 D (M g):
// a masked divergence, one PCG solve on the same operator, and an unmasked
// gradient. For a row cell c the faces M keeps are those whose other cell is
// not SOLID.
//
// Substep. Once the particle positions are fixed, one FLIP substep
// (flip_app_sim.cc:211-254) is linear in the particle velocities, the grid
// velocity and the body force:
//   v1 = P2G(vel); v2 = BC(v1); s = v2; v3 = v2 + dt F f; v4 = BC(v3);
//   u' = P v4; dv = mask ? u' - s : s; (ue, dve) = Extrapolate(u', dv);
//   vel' = (1 - flip) I(ue) + flip I(dve) + flip vel.
// The backward pass applies the transposes in reverse order. Each substep's
// masks and P2G weight sums are recomputed from a checkpoint of its
// (sorted) particles, taken by the forward pass right after advection.
#include <algorithm>
#include <cstring>
#include <vector>

#include "device_util.h"
#include "fsl_internal.h"

namespace fsl {

struct Checkpoint {
  float *pv[6] = {};        // positions and velocities after advection + sort
  uint32_t *perm = nullptr; // new index -> index before the advection
  int64_t n = 0, n_prev = 0;
  float dt = 0.0f;
};

struct AdjState {
  float *g[3] = {}, *s[3] = {}, *wsum[3] = {};
  float *rhs = nullptr, *y = nullptr;
  float *pa[2][3] = {};  // particle adjoints (ping-pong)
  float *box = nullptr;  // box-format staging, 3 floats per box voxel
  double *fpart = nullptr, *fsum = nullptr;
  int64_t nfpart = 0;
  std::vector<Checkpoint> ck;
  int steps = 0;
};

namespace {

constexpr int kThreads = 256;
inline unsigned blocks_for(int64_t n, int t) { return unsigned((n + t - 1) / t); }

__device__ inline bool leaf_voxel(const Dom &d, const int32_t *list, int64_t nlist, int64_t &v,
                                  int c[3]) {
  const int64_t w = xcd_work(nlist);
  if (w >= nlist) return false;
  const int64_t leaf = list[w];
  int o[3];
  leaf_origin(d, leaf, o);
  const int t = threadIdx.x;
  c[0] = o[0] + (t >> 6); c[1] = o[1] + ((t >> 3) & 7); c[2] = o[2] + (t & 7);
  v = (leaf << 9) | t;
  return true;
}

__device__ inline bool in_faces(const Dom &d, const int c[3]) {
  return c[0] >= 0 && c[1] >= 0 && c[2] >= 0 && c[0] <= d.n[0] && c[1] <= d.n[1] &&
         c[2] <= d.n[2];
}

__device__ inline bool is_row(uint8_t l) { return (l & 3) == kFluid; }

// Box format (3 floats per voxel of [-1, n]^3, x-major) -> leaf storage.
__global__ void k_box_in(Dom d, const float *box, float *g0, float *g1, float *g2, int add) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int bx = d.n[0] + 2, by = d.n[1] + 2, bz = d.n[2] + 2;
  if (q >= int64_t(bx) * by * bz) return;
  const int i = int(q / (int64_t(by) * bz)) - 1, j = int((q / bz) % by) - 1, k = int(q % bz) - 1;
  const int64_t v = vox_of(d, i, j, k);
  if (add) {
    g0[v] = g0[v] + box[3 * q]; g1[v] = g1[v] + box[3 * q + 1]; g2[v] = g2[v] + box[3 * q + 2];
  } else {
    g0[v] = box[3 * q]; g1[v] = box[3 * q + 1]; g2[v] = box[3 * q + 2];
  }
}

// Leaf storage -> box format; with bit >= 0 only where vflag has that bit.
__global__ void k_box_out(Dom d, const float *g0, const float *g1, const float *g2,
                          const uint8_t *vflag, int bit, float *box) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int bx = d.n[0] + 2, by = d.n[1] + 2, bz = d.n[2] + 2;
  if (q >= int64_t(bx) * by * bz) return;
  const int i = int(q / (int64_t(by) * bz)) - 1, j = int((q / bz) % by) - 1, k = int(q % bz) - 1;
  const int64_t v = vox_of(d, i, j, k);
  const bool on = bit < 0 || ((vflag[v] >> bit) & 1);
  box[3 * q] = on ? g0[v] : 0.0f;
  box[3 * q + 1] = on ? g1[v] : 0.0f;
  box[3 * q + 2] = on ? g2[v] : 0.0f;
}

// Sets the grid velocity on faces active after P2G (vflag bit 0) from a box.
__global__ void k_override(Dom d, const float *box, const uint8_t *vflag, float *v0, float *v1,
                           float *v2) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int bx = d.n[0] + 2, by = d.n[1] + 2, bz = d.n[2] + 2;
  if (q >= int64_t(bx) * by * bz) return;
  const int i = int(q / (int64_t(by) * bz)) - 1, j = int((q / bz) % by) - 1, k = int(q % bz) - 1;
  const int64_t v = vox_of(d, i, j, k);
  if (!(vflag[v] & 1)) return;
  v0[v] = box[3 * q]; v1[v] = box[3 * q + 1]; v2[v] = box[3 * q + 2];
}

// rhs = D (M g) on rows, 0 elsewhere in the fluid leaves.
__global__ __launch_bounds__(512) void k_adj_div(Dom d, const int32_t *list, int64_t nlist,
                                                 const uint8_t *lab, const uint8_t *marker,
                                                 const float *gx, const float *gy,
                                                 const float *gz, float *rhs) {
#pragma clang fp contract(off)
  int64_t v;
  int c[3];
  if (!leaf_voxel(d, list, nlist, v, c)) return;
  float r = 0.0f;
  if (is_row(lab[v])) {
    const float *g[3] = {gx, gy, gz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      int up[3] = {c[0], c[1], c[2]}, dn[3] = {c[0], c[1], c[2]};
      up[a] += 1;
      dn[a] -= 1;
      const int64_t vu = vox_of(d, up[0], up[1], up[2]);
      const int64_t vd = vox_of(d, dn[0], dn[1], dn[2]);
      const float gu = marker[vu] != kSolid ? g[a][vu] : 0.0f;
      const float gd = marker[vd] != kSolid ? g[a][v] : 0.0f;
      r = r + (gu - gd);
    }
  }
  rhs[v] = r;
}

// g += G~ y on faces [0, n]: g_d(c) += y(c) - y(c - e_d), y = 0 off the rows.
__global__ __launch_bounds__(512) void k_adj_grad(Dom d, const int32_t *list, int64_t nlist,
                                                  const uint8_t *lab, const float *y, float *gx,
                                                  float *gy, float *gz) {
#pragma clang fp contract(off)
  int64_t v;
  int c[3];
  if (!leaf_voxel(d, list, nlist, v, c)) return;
  if (!in_faces(d, c)) return;
  const float yc = is_row(lab[v]) ? y[v] : 0.0f;
  float *g[3] = {gx, gy, gz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int dn[3] = {c[0], c[1], c[2]};
    dn[a] -= 1;
    const int64_t vd = vox_of(d, dn[0], dn[1], dn[2]);
    const float yd = is_row(lab[vd]) ? y[vd] : 0.0f;
    g[a][v] = g[a][v] + (yc - yd);
  }
}

// G2P transpose output T (in g) split into the adjoints of the extrapolated
// velocity and velocity-update grids: g = (1 - flip) T, s = flip T.
__global__ __launch_bounds__(512) void k_adj_split(const int32_t *list, int64_t nlist,
                                                   float flip, float *gx, float *gy, float *gz,
                                                   float *sx, float *sy, float *sz) {
  const int64_t w = xcd_work(nlist);
  if (w >= nlist) return;
  const int64_t v = (int64_t(list[w]) << 9) | threadIdx.x;
  const float keep = 1.0f - flip;
  float *g[3] = {gx, gy, gz}, *s[3] = {sx, sy, sz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float t = g[a][v];
    g[a][v] = keep * t;
    s[a][v] = flip * t;
  }
}

// Transpose of one extrapolation pass (k_extrapolate's copy of the
// tangential components from the layer inside to the boundary layer).
__global__ void k_adj_extrap(Dom d, int axis, float *gx, float *gy, float *gz, float *sx,
                             float *sy, float *sz) {
  const int a = axis == 0 ? 1 : 0, b = axis == 2 ? 1 : 2;
  const int na = d.n[a] + 1, nb = d.n[b] + 1;
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= 2 * int64_t(na) * nb) return;
  const int side = int(q / (int64_t(na) * nb));
  const int ia = int((q / nb) % na), ib = int(q % nb);
  int out[3], in[3];
  out[axis] = side ? d.n[axis] : -1;
  in[axis] = side ? d.n[axis] - 1 : 0;
  out[a] = in[a] = ia;
  out[b] = in[b] = ib;
  const int64_t vo = vox_of(d, out[0], out[1], out[2]);
  const int64_t vi = vox_of(d, in[0], in[1], in[2]);
  float *g[3] = {gx, gy, gz}, *s[3] = {sx, sy, sz};
  const int comp[2] = {a, b};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int cc = comp[t];
    g[cc][vi] = g[cc][vi] + g[cc][vo];
    g[cc][vo] = 0.0f;
    s[cc][vi] = s[cc][vi] + s[cc][vo];
    s[cc][vo] = 0.0f;
  }
}

// Transpose of GetVelocityUpdate: on faces [0, n] active in both grids the
// update grid holds u' - s, elsewhere it keeps s.
__global__ __launch_bounds__(512) void k_adj_velupd(Dom d, const int32_t *list, int64_t nlist,
                                                    const uint8_t *vflag, float *gx, float *gy,
                                                    float *gz, float *sx, float *sy,
                                                    float *sz) {
  int64_t v;
  int c[3];
  if (!leaf_voxel(d, list, nlist, v, c)) return;
  if (!in_faces(d, c) || (vflag[v] & 3) != 3) return;
  float *g[3] = {gx, gy, gz}, *s[3] = {sx, sy, sz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float t = s[a][v];
    g[a][v] = g[a][v] + t;
    s[a][v] = -t;
  }
}

// Transpose of ApplyBoundaryConditions: zero the component on faces between
// a solid and a non-solid cell. (The forward zeroes active faces only; an
// inactive face holds the constant 0, so its adjoint is never used.)
__global__ __launch_bounds__(512) void k_adj_bc(Dom d, const int32_t *list, int64_t nlist,
                                                const uint8_t *solid, float *gx, float *gy,
                                                float *gz) {
  int64_t v;
  int c[3];
  if (!leaf_voxel(d, list, nlist, v, c)) return;
  const uint8_t sc = solid[v];
  float *g[3] = {gx, gy, gz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int dn[3] = {c[0], c[1], c[2]};
    dn[a] -= 1;
    const uint8_t sd = in_store(d, dn[0], dn[1], dn[2]) ? solid[vox_of(d, dn[0], dn[1], dn[2])]
                                                        : uint8_t(0);
    if (sc != sd) g[a][v] = 0.0f;
  }
}

// Per-workgroup sums of g over the faces that were active when gravity was
// added (vflag bit 1, set by SaveGridVelocities before AddGravity).
__global__ __launch_bounds__(512) void k_adj_force(Dom d, const int32_t *list, int64_t nlist,
                                                   const uint8_t *vflag, const float *gx,
                                                   const float *gy, const float *gz,
                                                   double *part) {
  __shared__ double s_r[16];
  const int64_t w = xcd_work(nlist);
  double a[3] = {0.0, 0.0, 0.0};
  if (w < nlist) {
    const int64_t v = (int64_t(list[w]) << 9) | threadIdx.x;
    if (vflag[v] & 2) { a[0] = gx[v]; a[1] = gy[v]; a[2] = gz[v]; }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double t = block_sum<double>(a[c], s_r);
    if (threadIdx.x == 0) part[3 * int64_t(blockIdx.x) + c] = t;
  }
}

// fsum += scale * (sum of the partials), in a fixed order.
__global__ void k_adj_force_total(const double *part, int64_t nb, double sx, double sy, double sz,
                                  double *fsum) {
  __shared__ double s_r[16];
  const double scale[3] = {sx, sy, sz};
  for (int c = 0; c < 3; ++c) {
    double a = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += blockDim.x) a += part[3 * b + c];
    const double t = block_sum<double>(a, s_r);
    if (threadIdx.x == 0) fsum[c] += scale[c] * t;
  }
}

// g += s on the active leaves (transpose of SaveGridVelocities' copy).
__global__ __launch_bounds__(512) void k_adj_add(const int32_t *list, int64_t nlist, float *gx,
                                                 float *gy, float *gz, const float *sx,
                                                 const float *sy, const float *sz) {
  const int64_t w = xcd_work(nlist);
  if (w >= nlist) return;
  const int64_t v = (int64_t(list[w]) << 9) | threadIdx.x;
  gx[v] = gx[v] + sx[v];
  gy[v] = gy[v] + sy[v];
  gz[v] = gz[v] + sz[v];
}

// out[perm[i]] = in[i] (particle adjoints back through a sort).
__global__ void k_perm_scatter(int64_t n, const uint32_t *perm, const float *i0, const float *i1,
                               const float *i2, float *o0, float *o1, float *o2) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint32_t p = perm[q];
  o0[p] = i0[q]; o1[p] = i1[q]; o2[p] = i2[q];
}

// out[i] = in[perm[i]] (particle velocities forward through a sort).
__global__ void k_perm_gather(int64_t n, const uint32_t *perm, const float *i0, const float *i1,
                              const float *i2, float *o0, float *o1, float *o2) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint32_t p = perm[q];
  o0[q] = i0[p]; o1[q] = i1[p]; o2[q] = i2[p];
}

int64_t box_count(const Dom &d) { return int64_t(d.n[0] + 2) * (d.n[1] + 2) * (d.n[2] + 2); }

template <typename T>
int alloc(fsl_ctx *ctx, T **p, int64_t count) {
  FSL_HIP(hipMalloc(reinterpret_cast<void **>(p), size_t(std::max<int64_t>(count, 1)) * sizeof(T)));
  FSL_HIP(hipMemsetAsync(*p, 0, size_t(std::max<int64_t>(count, 1)) * sizeof(T), ctx->stream));
  return FSL_OK;
}

int ensure_state(fsl_ctx *ctx) {
  if (ctx->adj) return FSL_OK;
  ctx->adj = new AdjState();
  AdjState &A = *ctx->adj;
  int rc;
  for (int a = 0; a < 3; ++a) {
    if ((rc = alloc(ctx, &A.g[a], ctx->nvox))) return rc;
    if ((rc = alloc(ctx, &A.s[a], ctx->nvox))) return rc;
    if ((rc = alloc(ctx, &A.wsum[a], ctx->nvox))) return rc;
    for (int b = 0; b < 2; ++b)
      if ((rc = alloc(ctx, &A.pa[b][a], ctx->cap + fsl::kParticlePad))) return rc;
  }
  if ((rc = alloc(ctx, &A.rhs, ctx->nvox))) return rc;
  if ((rc = alloc(ctx, &A.y, ctx->nvox))) return rc;
  if ((rc = alloc(ctx, &A.box, 3 * box_count(ctx->dom)))) return rc;
  A.nfpart = ctx->nleaves + kXcd;
  if ((rc = alloc(ctx, &A.fpart, 3 * A.nfpart))) return rc;
  if ((rc = alloc(ctx, &A.fsum, 3))) return rc;
  return FSL_OK;
}

int free_checkpoints(fsl_ctx *ctx) {
  AdjState &A = *ctx->adj;
  for (Checkpoint &c : A.ck) {
    for (float *&p : c.pv) { if (p) FSL_HIP(hipFree(p)); p = nullptr; }
    if (c.perm) FSL_HIP(hipFree(c.perm));
    c.perm = nullptr;
  }
  A.ck.clear();
  A.steps = 0;
  return FSL_OK;
}

// `host` may also be a device array of the context's device: it is then read
// in place.
int upload_box(fsl_ctx *ctx, const float *host, float *const g[3], bool add) {
  AdjState &A = *ctx->adj;
  const int64_t nb = box_count(ctx->dom);
  const float *src = A.box;
  if (on_ctx_device(ctx, host)) {
    src = host;
  } else {
    FSL_HIP(hipMemcpyAsync(A.box, host, size_t(3 * nb) * sizeof(float), hipMemcpyHostToDevice,
                           ctx->stream));
  }
  k_box_in<<<blocks_for(nb, kThreads), kThreads, 0, ctx->stream>>>(ctx->dom, src, g[0], g[1],
                                                                    g[2], add ? 1 : 0);
  FSL_HIP(hipGetLastError());
  return FSL_OK;
}

// `host` may also be a device array of the context's device: it is then
// written in place.
int download_box(fsl_ctx *ctx, float *const g[3], int bit, float *host) {
  AdjState &A = *ctx->adj;
  const int64_t nb = box_count(ctx->dom);
  const bool dev = on_ctx_device(ctx, host);
  k_box_out<<<blocks_for(nb, kThreads), kThreads, 0, ctx->stream>>>(ctx->dom, g[0], g[1], g[2],
                                                                     ctx->vflag, bit,
                                                                     dev ? host : A.box);
  FSL_HIP(hipGetLastError());
  if (!dev)
    FSL_HIP(hipMemcpyAsync(host, A.box, size_t(3 * nb) * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  return FSL_OK;
}

// g <- P^T g on the labels of the last divergence pass.
int adj_project(fsl_ctx *ctx, const fsl_solve_params *prm, fsl_solve_stats *st) {
  AdjState &A = *ctx->adj;
  if (ctx->flags_host->zero_diag) {
    ctx->err = "FLUID cell with no non-SOLID neighbour (zero Jacobi diagonal)";
    return FSL_ERR_ZERO_DIAGONAL;
  }
  if (ctx->n_fluid > 0) {
    k_adj_div<<<xcd_grid(ctx->n_fluid), 512, 0, ctx->stream>>>(
        ctx->dom, ctx->fluid_list, ctx->n_fluid, ctx->lab, ctx->marker, A.g[0], A.g[1], A.g[2],
        A.rhs);
    FSL_HIP(hipGetLastError());
  }
  // The solver reads b from the divergence grid and writes x to the pressure
  // grid; point both at the adjoint's buffers for this solve.
  fsl_solve_stats local{};
  fsl_solve_stats *s = st ? st : &local;
  *s = fsl_solve_stats{};
  s->n_rows = int(ctx->flags_host->n_rows);
  s->n_cols = s->n_rows;
  fsl_solve_params p = *prm;
  p.warm_start = 0;
  std::swap(ctx->div, A.rhs);
  std::swap(ctx->pres, A.y);
  int rc = ctx->flags_host->n_rows > 0 ? pcg_run(ctx, &p, s) : FSL_OK;
  std::swap(ctx->div, A.rhs);
  std::swap(ctx->pres, A.y);
  if (rc) return rc;
  if (ctx->n_act > 0) {
    k_adj_grad<<<xcd_grid(ctx->n_act), 512, 0, ctx->stream>>>(
        ctx->dom, ctx->act_list, ctx->n_act, ctx->lab, A.y, A.g[0], A.g[1], A.g[2]);
    FSL_HIP(hipGetLastError());
  }
  return FSL_OK;
}

int adj_bc(fsl_ctx *ctx) {
  AdjState &A = *ctx->adj;
  if (ctx->n_act == 0) return FSL_OK;
  k_adj_bc<<<xcd_grid(ctx->n_act), 512, 0, ctx->stream>>>(ctx->dom, ctx->act_list, ctx->n_act,
                                                           ctx->solid, A.g[0], A.g[1], A.g[2]);
  FSL_HIP(hipGetLastError());
  return FSL_OK;
}

// Restores checkpoint k into the particle arrays (grid velocity zeroed) and
// rebuilds the cell ranges with a zero-length advection (the order is kept:
// the particles are already sorted and the sort is stable).
int restore(fsl_ctx *ctx, int k, bool velocities) {
  AdjState &A = *ctx->adj;
  const Checkpoint &c = A.ck[k];
  const size_t bytes = size_t(c.n) * sizeof(float);
  for (int a = 0; a < 6; ++a) {
    if (a >= 3 && !velocities) continue;
    FSL_HIP(hipMemcpyAsync(ctx->P.a[a], c.pv[a], bytes, hipMemcpyDeviceToDevice, ctx->stream));
  }
  for (int a = 6; a < 9; ++a) FSL_HIP(hipMemsetAsync(ctx->P.a[a], 0, bytes, ctx->stream));
  ctx->np = c.n;
  ctx->outside_deleted = false;
  return fsl_move_particles(ctx, 0.0f);
}

// The grid part of a substep up to the pressure solve's right-hand side:
// clear, P2G (optionally with weight sums), optional override of the P2G
// velocity from a box, boundary conditions + save, gravity, boundary
// conditions, divergence and labels.
int grid_front(fsl_ctx *ctx, float dt, float *const wsum[3], const float *u0_box) {
  int rc;
  if ((rc = fsl_clear(ctx))) return rc;
  FSL_HIP(launch_p2g(ctx, wsum));
  if (u0_box) {
    AdjState &A = *ctx->adj;
    const int64_t nb = box_count(ctx->dom);
    FSL_HIP(hipMemcpyAsync(A.box, u0_box, size_t(3 * nb) * sizeof(float), hipMemcpyHostToDevice,
                           ctx->stream));
    k_override<<<blocks_for(nb, kThreads), kThreads, 0, ctx->stream>>>(
        ctx->dom, A.box, ctx->vflag, ctx->v[0], ctx->v[1], ctx->v[2]);
    FSL_HIP(hipGetLastError());
  }
  FSL_HIP(launch_bc_save(ctx));
  ctx->bc_applied = true;
  if ((rc = fsl_delete_outside(ctx))) return rc;
  if ((rc = fsl_add_gravity(ctx, dt))) return rc;
  if ((rc = fsl_apply_bc(ctx))) return rc;
  return fsl_divergence(ctx);
}

// The rest of a forward substep: solve, pressure update, optional download of
// the post-projection velocity, velocity update, extrapolation and G2P.
int grid_back(fsl_ctx *ctx, const fsl_solve_params *prm, float *u_post_box,
              fsl_solve_stats *st) {
  int rc;
  if ((rc = fsl_pcg_solve(ctx, prm, st))) return rc;
  if ((rc = fsl_pressure_update(ctx))) return rc;
  if (u_post_box) {
    const int64_t nb = box_count(ctx->dom);
    FSL_HIP(launch_box_vec(ctx, 0, ctx->adj->box, nullptr, true));
    FSL_HIP(hipMemcpyAsync(u_post_box, ctx->adj->box, size_t(3 * nb) * sizeof(float),
                           hipMemcpyDeviceToHost, ctx->stream));
    FSL_HIP(hipStreamSynchronize(ctx->stream));
  }
  if ((rc = fsl_velocity_update(ctx))) return rc;
  return fsl_g2p(ctx);
}

}  // namespace

int project_backward(fsl_ctx *ctx, const fsl_solve_params *prm, const float *grad_out,
                     float *grad_in, fsl_solve_stats *st) {
  int rc;
  if ((rc = ensure_state(ctx))) return rc;
  if (!ctx->grid_ready) {
    ctx->err = "no labels: run the forward projection first";
    return FSL_ERR_STATE;
  }
  AdjState &A = *ctx->adj;
  if ((rc = upload_box(ctx, grad_out, A.g, false))) return rc;
  if ((rc = adj_project(ctx, prm, st))) return rc;
  return download_box(ctx, A.g, -1, grad_in);
}

void adjoint_free(fsl_ctx *ctx) {
  if (!ctx->adj) return;
  AdjState &A = *ctx->adj;
  free_checkpoints(ctx);
  for (int a = 0; a < 3; ++a) {
    (void)hipFree(A.g[a]); (void)hipFree(A.s[a]); (void)hipFree(A.wsum[a]);
    (void)hipFree(A.pa[0][a]); (void)hipFree(A.pa[1][a]);
  }
  (void)hipFree(A.rhs); (void)hipFree(A.y); (void)hipFree(A.box);
  (void)hipFree(A.fpart); (void)hipFree(A.fsum);
  delete ctx->adj;
  ctx->adj = nullptr;
}

}  // namespace fsl

using fsl::AdjState;
using fsl::Checkpoint;

extern "C" {

int fsl_rollout_begin(fsl_ctx *ctx, int max_substeps) {
  if (!ctx || max_substeps < 1) return FSL_ERR_ARG;
  if (ctx->counts_on_device && fsl::host_counts(ctx)) return FSL_ERR_HIP;
  int rc;
  if ((rc = fsl::ensure_state(ctx))) return rc;
  if ((rc = fsl::free_checkpoints(ctx))) return rc;
  AdjState &A = *ctx->adj;
  A.ck.resize(size_t(max_substeps));
  for (Checkpoint &c : A.ck) {
    for (float *&p : c.pv) if ((rc = fsl::alloc(ctx, &p, ctx->cap))) return rc;
    if ((rc = fsl::alloc(ctx, &c.perm, ctx->cap))) return rc;
  }
  return FSL_OK;
}

int fsl_rollout_step(fsl_ctx *ctx, float dt, const fsl_solve_params *params, const float *u0_box,
                     float *u_post_box, fsl_solve_stats *stats) {
  if (!ctx || !params) return FSL_ERR_ARG;
  if (ctx->counts_on_device && fsl::host_counts(ctx)) return FSL_ERR_HIP;
  if (!ctx->adj || ctx->adj->ck.empty()) {
    ctx->err = "fsl_rollout_begin first";
    return FSL_ERR_STATE;
  }
  AdjState &A = *ctx->adj;
  if (A.steps >= int(A.ck.size())) {
    ctx->err = "rollout longer than fsl_rollout_begin's max_substeps";
    return FSL_ERR_CAPACITY;
  }
  if (u0_box && A.steps != 0) return FSL_ERR_ARG;
  Checkpoint &c = A.ck[A.steps];
  c.n_prev = ctx->np;
  int rc;
  if ((rc = fsl_move_particles(ctx, dt))) return rc;
  c.n = ctx->np;
  c.dt = dt;
  const size_t bytes = size_t(c.n) * sizeof(float);
  for (int a = 0; a < 6; ++a)
    FSL_HIP(hipMemcpyAsync(c.pv[a], ctx->P.a[a], bytes, hipMemcpyDeviceToDevice, ctx->stream));
  FSL_HIP(hipMemcpyAsync(c.perm, ctx->perm, size_t(c.n) * sizeof(uint32_t),
                         hipMemcpyDeviceToDevice, ctx->stream));
  if ((rc = fsl::grid_front(ctx, dt, nullptr, u0_box))) return rc;
  if ((rc = fsl::grid_back(ctx, params, u_post_box, stats))) return rc;
  ++A.steps;
  return FSL_OK;
}

int fsl_rollout_replay(fsl_ctx *ctx, const fsl_solve_params *params, const float *u0_box,
                       const float force[3], float *u_post_box, fsl_solve_stats *stats) {
  if (!ctx || !params || !force) return FSL_ERR_ARG;
  if (ctx->counts_on_device && fsl::host_counts(ctx)) return FSL_ERR_HIP;
  if (!ctx->adj || ctx->adj->steps == 0) {
    ctx->err = "no recorded rollout";
    return FSL_ERR_STATE;
  }
  AdjState &A = *ctx->adj;
  float saved[3];
  std::memcpy(saved, ctx->force, sizeof(saved));
  std::memcpy(ctx->force, force, sizeof(saved));
  int rc = FSL_OK;
  for (int k = 0; k < A.steps && rc == FSL_OK; ++k) {
    const Checkpoint &c = A.ck[k];
    if (k == 0) {
      rc = fsl::restore(ctx, 0, true);
    } else {
      // Velocities of the previous substep's G2P, through this substep's sort.
      float *tmp[3] = {A.pa[0][0], A.pa[0][1], A.pa[0][2]};
      if (c.n > 0) {
        fsl::k_perm_gather<<<fsl::blocks_for(c.n, fsl::kThreads), fsl::kThreads, 0,
                             ctx->stream>>>(c.n, c.perm, ctx->P.a[3], ctx->P.a[4], ctx->P.a[5],
                                            tmp[0], tmp[1], tmp[2]);
        FSL_HIP(hipGetLastError());
      }
      if ((rc = fsl::restore(ctx, k, false))) break;
      for (int a = 0; a < 3; ++a)
        FSL_HIP(hipMemcpyAsync(ctx->P.a[3 + a], tmp[a], size_t(c.n) * sizeof(float),
                               hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (rc) break;
    if ((rc = fsl::grid_front(ctx, c.dt, nullptr, k == 0 ? u0_box : nullptr))) break;
    rc = fsl::grid_back(ctx, params, k == A.steps - 1 ? u_post_box : nullptr, stats);
  }
  std::memcpy(ctx->force, saved, sizeof(saved));
  return rc;
}

int fsl_rollout_backward(fsl_ctx *ctx, const fsl_solve_params *params, const float *loss_grad,
                         float *grad_u0, double grad_force[3], fsl_solve_stats *stats) {
  if (!ctx || !params || !loss_grad) return FSL_ERR_ARG;
  if (ctx->counts_on_device && fsl::host_counts(ctx)) return FSL_ERR_HIP;
  if (!ctx->adj || ctx->adj->steps == 0) {
    ctx->err = "no recorded rollout";
    return FSL_ERR_STATE;
  }
  AdjState &A = *ctx->adj;
  int rc;
  const int K = A.steps;
  FSL_HIP(hipMemsetAsync(A.fsum, 0, 3 * sizeof(double), ctx->stream));
  // Adjoint of the last substep's G2P output: zero.
  int cur = 0;
  for (int a = 0; a < 3; ++a)
    FSL_HIP(hipMemsetAsync(A.pa[cur][a], 0, size_t(A.ck[K - 1].n) * sizeof(float), ctx->stream));
  for (int k = K - 1; k >= 0; --k) {
    const Checkpoint &c = A.ck[k];
    // Recompute the substep's masks, labels and P2G weight sums.
    if ((rc = fsl::restore(ctx, k, true))) return rc;
    if ((rc = fsl::grid_front(ctx, c.dt, A.wsum, nullptr))) return rc;
    FSL_HIP(fsl::launch_pressure_update(ctx));  // activity only: pressure is 0
    const int32_t *L = ctx->act_list;
    const int64_t nl = ctx->n_act;
    const int gl = fsl::xcd_grid(nl);
    if (nl == 0) continue;
    for (int a = 0; a < 3; ++a) {
      FSL_HIP(hipMemsetAsync(A.g[a], 0, size_t(ctx->nvox) * sizeof(float), ctx->stream));
      FSL_HIP(hipMemsetAsync(A.s[a], 0, size_t(ctx->nvox) * sizeof(float), ctx->stream));
    }
    // G2P transpose, then the split into the two interpolated grids.
    float *pin[3] = {A.pa[cur][0], A.pa[cur][1], A.pa[cur][2]};
    FSL_HIP(fsl::launch_g2p_transpose(ctx, pin, A.g));
    fsl::k_adj_split<<<gl, 512, 0, ctx->stream>>>(L, nl, ctx->flip, A.g[0], A.g[1], A.g[2],
                                                  A.s[0], A.s[1], A.s[2]);
    FSL_HIP(hipGetLastError());
    for (int axis = 2; axis >= 0; --axis) {
      const int a = axis == 0 ? 1 : 0, b = axis == 2 ? 1 : 2;
      const int64_t n = 2 * int64_t(ctx->dom.n[a] + 1) * (ctx->dom.n[b] + 1);
      fsl::k_adj_extrap<<<fsl::blocks_for(n, fsl::kThreads), fsl::kThreads, 0, ctx->stream>>>(
          ctx->dom, axis, A.g[0], A.g[1], A.g[2], A.s[0], A.s[1], A.s[2]);
    }
    fsl::k_adj_velupd<<<gl, 512, 0, ctx->stream>>>(ctx->dom, L, nl, ctx->vflag, A.g[0], A.g[1],
                                                   A.g[2], A.s[0], A.s[1], A.s[2]);
    FSL_HIP(hipGetLastError());
    if (k == K - 1 && (rc = fsl::upload_box(ctx, loss_grad, A.g, true))) return rc;
    if ((rc = fsl::adj_project(ctx, params, stats))) return rc;
    if ((rc = fsl::adj_bc(ctx))) return rc;  // second ApplyBoundaryConditions
    fsl::k_adj_force<<<gl, 512, 0, ctx->stream>>>(ctx->dom, L, nl, ctx->vflag, A.g[0], A.g[1],
                                                  A.g[2], A.fpart);
    const double dt = double(c.dt);
    fsl::k_adj_force_total<<<1, 1024, 0, ctx->stream>>>(A.fpart, gl, dt, -dt, dt, A.fsum);
    fsl::k_adj_add<<<gl, 512, 0, ctx->stream>>>(L, nl, A.g[0], A.g[1], A.g[2], A.s[0], A.s[1],
                                                A.s[2]);
    FSL_HIP(hipGetLastError());
    if ((rc = fsl::adj_bc(ctx))) return rc;  // first ApplyBoundaryConditions
    // Particle adjoint of this substep's input velocities: flip * pin from
    // the FLIP term, plus the P2G transpose (not for substep 0 when its P2G
    // output was replaced by u0).
    float *const *pout = A.pa[cur];
    if (k == 0 && grad_u0) {
      if ((rc = fsl::download_box(ctx, A.g, 1, grad_u0))) return rc;
    }
    FSL_HIP(fsl::launch_p2g_transpose(ctx, A.g, A.wsum, ctx->flip, pout));
    if (k > 0) {
      const int nxt = cur ^ 1;
      for (int a = 0; a < 3; ++a)
        FSL_HIP(hipMemsetAsync(A.pa[nxt][a], 0, size_t(c.n_prev) * sizeof(float), ctx->stream));
      if (c.n > 0) {
        fsl::k_perm_scatter<<<fsl::blocks_for(c.n, fsl::kThreads), fsl::kThreads, 0,
                              ctx->stream>>>(c.n, c.perm, pout[0], pout[1], pout[2],
                                             A.pa[nxt][0], A.pa[nxt][1], A.pa[nxt][2]);
        FSL_HIP(hipGetLastError());
      }
      cur = nxt;
    }
  }
  if (grad_force) {
    FSL_HIP(hipMemcpyAsync(grad_force, A.fsum, 3 * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
  }
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  return FSL_OK;
}

int fsl_rollout_checkpoint(fsl_ctx *ctx, int k, float *pos, float *vel, uint32_t *perm,
                           int64_t *n, int64_t *n_prev, float *dt) {
  if (!ctx || !ctx->adj || k < 0 || k >= ctx->adj->steps) return FSL_ERR_ARG;
  if (ctx->counts_on_device && fsl::host_counts(ctx)) return FSL_ERR_HIP;
  const Checkpoint &c = ctx->adj->ck[k];
  if (n) *n = c.n;
  if (n_prev) *n_prev = c.n_prev;
  if (dt) *dt = c.dt;
  FSL_HIP(hipStreamSynchronize(ctx->stream));  // hipMemcpy does not wait for it
  std::vector<float> tmp(size_t(c.n));
  for (int a = 0; a < 6; ++a) {
    float *dst = a < 3 ? pos : vel;
    if (!dst) continue;
    FSL_HIP(hipMemcpy(tmp.data(), c.pv[a], size_t(c.n) * sizeof(float), hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < c.n; ++q) dst[3 * q + (a % 3)] = tmp[size_t(q)];
  }
  if (perm)
    FSL_HIP(hipMemcpy(perm, c.perm, size_t(c.n) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return FSL_OK;
}

}  // extern "C"
