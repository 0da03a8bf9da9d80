// Grid-stage kernels: clear, boundary conditions, save, gravity, divergence
// and solver labels, pressure-gradient update, velocity update, boundary
// extrapolation, CFL, and box-format conversion. One workgroup of 512
// threads per active leaf, one thread per voxel.
#include "device_util.h"
#include "fsl_internal.h"

namespace fsl {

namespace {

constexpr int kThreads = 256;
inline unsigned blocks_for(int64_t n, int t) { return unsigned((n + t - 1) / t); }

// Runs f(v, c) for each leaf of L with v this thread's voxel (one thread
// per voxel, 512 per workgroup) and c its cell coordinates.
template <typename F>
__device__ inline void for_each_voxel(const Dom &d, const LeafList &L, F &&f) {
  for_each_leaf(L.list, L.count, L.n, [&](int64_t leaf) {
    int o[3];
    leaf_origin(d, leaf, o);
    const int t = threadIdx.x;
    const int c[3] = {o[0] + (t >> 6), o[1] + ((t >> 3) & 7), o[2] + (t & 7)};
    f((leaf << 9) | t, c);
  });
}

__global__ __launch_bounds__(512) void k_clear(LeafList L, float *v0, float *v1, float *v2,
                                               float *u0, float *u1, float *u2, uint8_t *vflag,
                                               uint8_t *marker, uint8_t *lab, float *div,
                                               float *pres) {
  for_each_leaf(L.list, L.count, L.n, [&](int64_t leaf) {
    const int64_t v = (leaf << 9) | threadIdx.x;
    v0[v] = 0; v1[v] = 0; v2[v] = 0; u0[v] = 0; u1[v] = 0; u2[v] = 0;
    vflag[v] = 0; marker[v] = 0; lab[v] = 0; div[v] = 0; pres[v] = 0;
  });
}

__device__ inline uint8_t solid_at(const Dom &d, const uint8_t *solid, int i, int j, int k) {
  return in_store(d, i, j, k) ? solid[vox_of(d, i, j, k)] : uint8_t(0);
}

// ApplyBoundaryConditions (flip_simulation.cc:656-713): a solid voxel is
// marked SOLID; the normal component on a face between a solid and a
// non-solid cell is zeroed when the face is active. Face d of voxel c lies
// between c - e_d and c, so it is zeroed iff solid(c) != solid(c - e_d).
__global__ __launch_bounds__(512) void k_bc(Dom d, LeafList L,
                                            const uint8_t *solid, uint8_t *marker,
                                            const uint8_t *vflag, float *vx, float *vy,
                                            float *vz, int save, float *ux, float *uy, float *uz,
                                            uint8_t *vflag_out) {
  for_each_voxel(d, L, [&](int64_t v, const int *c) {
    const uint8_t s = solid[v];
    if (s) marker[v] = kSolid;
    const uint8_t f = vflag[v];
    float a = vx[v], b = vy[v], e = vz[v];
    if (f & 1) {
      if (s != solid_at(d, solid, c[0] - 1, c[1], c[2])) a = 0.0f;
      if (s != solid_at(d, solid, c[0], c[1] - 1, c[2])) b = 0.0f;
      if (s != solid_at(d, solid, c[0], c[1], c[2] - 1)) e = 0.0f;
      vx[v] = a; vy[v] = b; vz[v] = e;
    }
    if (save) {  // SaveGridVelocities (flip_simulation.cc:490-496)
      ux[v] = a; uy[v] = b; uz[v] = e;
      vflag_out[v] = (f & 1) ? uint8_t(3) : uint8_t(0);
    }
  });
}

__global__ __launch_bounds__(512) void k_save(Dom d, LeafList L,
                                              const float *vx, const float *vy, const float *vz,
                                              float *ux, float *uy, float *uz, uint8_t *vflag) {
  for_each_voxel(d, L, [&](int64_t v, const int *c) {
    ux[v] = vx[v]; uy[v] = vy[v]; uz[v] = vz[v];
    vflag[v] = (vflag[v] & 1) ? uint8_t(3) : uint8_t(0);
  });
}

// AddGravity (flip_simulation.cc:505-515) on active faces.
// dt_dev, when set, holds dt on the device (dt is then unused).
__global__ __launch_bounds__(512) void k_gravity(Dom d, LeafList L,
                                                 const uint8_t *vflag, float *vx, float *vy,
                                                 float *vz, float dt, const float *dt_dev,
                                                 float fx, float fy, float fz) {
#pragma clang fp contract(off)
  if (dt_dev) dt = *dt_dev;
  const float ax = dt * fx, ay = dt * fy, az = dt * fz;
  for_each_voxel(d, L, [&](int64_t v, const int *c) {
#pragma clang fp contract(off)
    if (!(vflag[v] & 1)) return;
    vx[v] = vx[v] + ax;
    vy[v] = vy[v] - ay;
    vz[v] = vz[v] + az;
  });
}

// ComputeDivergence (flip_simulation.cc:715-736) fused with the solver
// labels: lab = marker | diag << 2, diag = count of FLUID or AIR neighbours
// (CreateLaplacianMatrix, solver.cc:214-231). Also stores each leaf's count
// of FLUID rows in lfluid (nonzero marks a fluid leaf); k_row_totals sums
// them.
__global__ __launch_bounds__(512) void k_divergence(Dom d, LeafList L,
                                                    const uint8_t *marker, const float *vx,
                                                    const float *vy, const float *vz,
                                                    float *div, uint8_t *lab, uint32_t *lfluid,
                                                    DevFlags *flags) {
#pragma clang fp contract(off)
  __shared__ unsigned s_cnt[16];
  for_each_voxel(d, L, [&](int64_t v, const int *c) {
#pragma clang fp contract(off)
    // Every load is issued whatever the marker turns out to be, so none waits
    // for it: cells outside the box read their own voxel in place of the
    // neighbours (in the box, the neighbours lie in [-1, n], inside storage).
    const bool inb = in_bbox(d, c[0], c[1], c[2]);
    const int off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    int64_t nv[6];
#pragma unroll
    for (int q = 0; q < 6; ++q)
      nv[q] = inb ? vox_of(d, c[0] + off[q][0], c[1] + off[q][1], c[2] + off[q][2]) : v;
    const uint8_t m = marker[v];
    uint8_t nm[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) nm[q] = marker[nv[q]];
    const float ux = vx[v], uxp = vx[nv[1]], uy = vy[v], uyp = vy[nv[3]], uz = vz[v],
                uzp = vz[nv[5]];
    const bool fluid = m == kFluid && inb;
    float dv = 0.0f;
    uint8_t l = m;
    if (fluid) {
      dv = uxp - ux;
      dv = dv + (uyp - uy);
      dv = dv + (uzp - uz);
      int diag = 0;
#pragma unroll
      for (int q = 0; q < 6; ++q) diag += (nm[q] == kFluid || nm[q] == kAir) ? 1 : 0;
      if (diag == 0) atomicOr(&flags->zero_diag, 1u);
      l = uint8_t(kFluid | (diag << 2));
    }
    // A FLUID marker outside the partition box is not a row of this system
    // (in the reference it is a ghost column); label 3 keeps it out of the
    // solve while still counting as FLUID in its neighbours' diagonals.
    if (m == kFluid && !fluid) l = 3;
    div[v] = dv;
    lab[v] = l;
    const unsigned cnt = unsigned(block_sum<unsigned>(fluid ? 1u : 0u, s_cnt));
    if (threadIdx.x == 0) lfluid[v >> 9] = cnt;
  });
}

// Row totals (solver.cc:80) and rows per tile (metagrid.cc:84-87) from the
// per-leaf FLUID counts: each workgroup sums a contiguous range of leaves
// per tile in LDS, then adds its nonzero tiles to flags (zeroed before the
// divergence) with one atomic each. A leaf lies in one tile (tiles are
// multiples of 8 cells wide); leaves outside the domain hold no rows.
constexpr int kRowTotalGroups = 64;
__global__ __launch_bounds__(256) void k_row_totals(Dom d, int64_t nleaves,
                                                   const uint32_t *lfluid, DevFlags *flags) {
  __shared__ unsigned long long s_tile[64];
  if (threadIdx.x < 64) s_tile[threadIdx.x] = 0;
  __syncthreads();
  const int64_t chunk = (nleaves + gridDim.x - 1) / gridDim.x;
  const int64_t beg = int64_t(blockIdx.x) * chunk;
  const int64_t end = beg + chunk < nleaves ? beg + chunk : nleaves;
  for (int64_t l = beg + threadIdx.x; l < end; l += blockDim.x) {
    const uint32_t cnt = lfluid[l];
    if (!cnt) continue;
    int o[3];
    leaf_origin(d, l, o);
    atomicAdd(&s_tile[tile_of(d, o[0], o[1], o[2]) & 63], (unsigned long long)cnt);
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const unsigned long long t = s_tile[threadIdx.x];
    if (t) {
      atomicAdd(&flags->tile_rows[threadIdx.x], t);
      atomicAdd(&flags->n_rows, t);
    }
  }
}

// UpdateVelocityFromPressure (flip_simulation.cc:738-778) over faces
// [0, n] of active leaves; AIR pressure is 0 (the cleared grid).
__global__ __launch_bounds__(512) void k_pressure_update(Dom d, LeafList L, const uint8_t *marker,
                                                         const float *pres, float *vx, float *vy,
                                                         float *vz, uint8_t *vflag) {
  for_each_voxel(d, L, [&](int64_t v, const int *c) {
    if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] > d.n[0] || c[1] > d.n[1] || c[2] > d.n[2]) return;
    const uint8_t m = marker[v];
    if (m == kSolid) return;
    const float pc = pres[v];
    bool upd = false;
    float *g[3] = {vx, vy, vz};
    const int off[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int64_t w = vox_of(d, c[0] - off[a][0], c[1] - off[a][1], c[2] - off[a][2]);
      const uint8_t mw = marker[w];
      if ((mw == kFluid || m == kFluid) && mw != kSolid) {
        g[a][v] = g[a][v] + (pc - pres[w]);
        upd = true;
      }
    }
    if (upd) vflag[v] = vflag[v] | 1;
  });
}

// GetVelocityUpdate (flip_simulation.cc:780-801): dv = v - saved on faces
// [0, n] active in both; elsewhere the saved value stays.
__global__ __launch_bounds__(512) void k_velocity_update(Dom d, LeafList L, const uint8_t *vflag,
                                                         const float *vx, const float *vy,
                                                         const float *vz, float *ux, float *uy,
                                                         float *uz, DevFlags *flags) {
  for_each_voxel(d, L, [&](int64_t v, const int *c) {
    if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] > d.n[0] || c[1] > d.n[1] || c[2] > d.n[2]) return;
    const uint8_t f = vflag[v];
    // The flag only goes from 0 to 1: one lane per wavefront sets it, and
    // only while a read still shows it unset (same-address atomics from
    // every voxel serialise).
    const unsigned long long need = __ballot((f & 1) && !(f & 2));
    if (need && (threadIdx.x & 63) == unsigned(__ffsll((long long)need) - 1) &&
        !__hip_atomic_load(&flags->vel_update, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicOr(&flags->vel_update, 1u);
    if ((f & 3) == 3) {
      ux[v] = vx[v] - ux[v];
      uy[v] = vy[v] - uy[v];
      uz[v] = vz[v] - uz[v];
    }
  });
}

// VectorGrid::ExtrapolateToBoundary (vector_grid.cc:302-381) for one axis
// (lower and upper layers), on v and dv. The reference runs x, y, z in
// sequence and later axes read values earlier axes wrote, so each axis is
// its own launch. Every face in the copy ranges becomes active, as set()
// activates in the reference; outside active leaves the copied values are the
// background zeros, and the activity flags of the boundary layers are cleared
// again by ClearGridData (k_clear_shell).
__global__ void k_extrapolate(Dom d, int axis, float *vx, float *vy, float *vz, float *ux,
                              float *uy, float *uz, uint8_t *vflag) {
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
  float *gv[3] = {vx, vy, vz};
  float *gu[3] = {ux, uy, uz};
  gv[a][vo] = gv[a][vi]; gv[b][vo] = gv[b][vi];
  gu[a][vo] = gu[a][vi]; gu[b][vo] = gu[b][vi];
  vflag[vo] = 3;
}

// Activity flags of the six boundary layers (index -1 and n on each axis),
// which extrapolation sets; part of ClearGridData.
__global__ void k_clear_shell(Dom d, uint8_t *vflag) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t per = int64_t(d.n[0] + 2) * (d.n[1] + 2) * (d.n[2] + 2);
  if (q >= per) return;
  const int by = d.n[1] + 2, bz = d.n[2] + 2;
  const int i = int(q / (int64_t(by) * bz)) - 1, j = int((q / bz) % by) - 1, k = int(q % bz) - 1;
  if (i == -1 || j == -1 || k == -1 || i == d.n[0] || j == d.n[1] || k == d.n[2])
    vflag[vox_of(d, i, j, k)] = 0;
}

// CFL (flip_simulation.cc:459-488): max |v| over active faces inside the
// partition box, skipping samples past the last face of each component.
// v_inf of CFL (flip_simulation.cc:459-488): a grid-stride walk over the
// active leaves (at most kCflBlocks workgroups, so one atomic each); the
// activity byte and the three components are loaded together.
constexpr int kCflBlocks = 1024;
__global__ __launch_bounds__(512) void k_cfl(Dom d, LeafList L, const uint8_t *vflag,
                                             const float *vx, const float *vy, const float *vz,
                                             unsigned *bits) {
  __shared__ float s_m[8];
  const int t = threadIdx.x;
  const int e0 = d.n[0] - 1, e1 = d.n[1] - 1, e2 = d.n[2] - 1;
  const int64_t nlist = L.count ? int64_t(*L.count) : L.n;
  float m = 0.0f;
  for (int64_t w = blockIdx.x; w < nlist; w += gridDim.x) {
    const int64_t leaf = L.list[w];
    int o[3];
    leaf_origin(d, leaf, o);
    const int c0 = o[0] + (t >> 6), c1 = o[1] + ((t >> 3) & 7), c2 = o[2] + (t & 7);
    const int64_t v = (leaf << 9) | t;
    const uint8_t f = vflag[v];
    const float ax = fabsf(vx[v]), ay = fabsf(vy[v]), az = fabsf(vz[v]);
    if (in_bbox(d, c0, c1, c2) && (f & 1)) {
      if (c1 != e1 && c2 != e2) m = fmaxf(m, ax);
      if (c0 != e0 && c2 != e2) m = fmaxf(m, ay);
      if (c0 != e0 && c1 != e1) m = fmaxf(m, az);
    }
  }
  m = wave_max(m);
  if ((t & 63) == 0) s_m[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    float x = 0.0f;
    for (int q = 0; q < 8; ++q) x = fmaxf(x, s_m[q]);
    atomicMax(bits, __float_as_uint(x));  // non-negative floats order as uints
  }
}

// WhileAndComputeDt's dt (flip_app_sim.cc:357-410) on the device: the CFL
// speed max(sqrt(g / 2), v_inf, 1e-2) (flip_simulation.cc:485-487, v_inf
// from k_cfl), dt = min(1 / frame_step, cfl / speed), floored at frame_step
// / 10, cut at the frame's end and snapped to it past 90% of the frame; the
// elapsed time advances by dt. The float/double mix is the host's
// (fsl_compute_dt). The frame's elapsed time is reset to 0 before its first
// step.
__global__ void k_dt(const unsigned *cfl_bits, DtParams p, StepState *ss) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const float v_inf = __uint_as_float(*cfl_bits);
  float v = float(fmax(p.v_gravity, double(v_inf)));
  if (v < 1e-2f) v = 1e-2f;
  const float inv = float(1.0 / double(v));
  const float local_dt = p.cfl * inv;
  const float step = ss->step;
  float g = p.dt_max < local_dt ? p.dt_max : local_dt;
  if (g < p.dt_min) g = p.dt_min;
  if (step + g > p.frame_step) g = p.frame_step - step;
  if (double(step + g) > p.snap) g = p.frame_step - step;
  const float after = step + g;
  ss->dt = g;
  ss->step = after;
  // The While test, step + THRESH < frame_step (flip_app_sim.cc:363, THRESH
  // 1e-5 in double).
  ss->more = double(after) + 1e-5 < double(p.frame_step) ? 1 : 0;
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t value) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q < n) p[q] = value;
}

// Box format <-> leaf storage.
__device__ inline bool box_coord(const Dom &d, int64_t q, int &i, int &j, int &k) {
  const int bx = d.n[0] + 2, by = d.n[1] + 2, bz = d.n[2] + 2;
  if (q >= int64_t(bx) * by * bz) return false;
  i = int(q / (int64_t(by) * bz)) - 1; j = int((q / bz) % by) - 1; k = int(q % bz) - 1;
  return true;
}
__global__ void k_box_vec(Dom d, float *g0, float *g1, float *g2, uint8_t *vflag, int bit,
                          float *box, uint8_t *act, int to_box) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!box_coord(d, q, i, j, k)) return;
  const int64_t v = vox_of(d, i, j, k);
  if (to_box) {
    box[3 * q] = g0[v]; box[3 * q + 1] = g1[v]; box[3 * q + 2] = g2[v];
    if (act) act[q] = (vflag[v] >> bit) & 1;
  } else {
    g0[v] = box[3 * q]; g1[v] = box[3 * q + 1]; g2[v] = box[3 * q + 2];
    const uint8_t on = act[q] ? uint8_t(1u << bit) : uint8_t(0);
    vflag[v] = uint8_t((vflag[v] & ~(1u << bit)) | on);
  }
}
__global__ void k_box_scalar(Dom d, float *g, float *box, int to_box) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!box_coord(d, q, i, j, k)) return;
  const int64_t v = vox_of(d, i, j, k);
  if (to_box) box[q] = g[v];
  else g[v] = box[q];
}
__global__ void k_box_u8(Dom d, uint8_t *g, int32_t *box, int to_box) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!box_coord(d, q, i, j, k)) return;
  const int64_t v = vox_of(d, i, j, k);
  if (to_box) box[q] = g[v];
  else g[v] = uint8_t(box[q]);
}

__global__ void k_all_leaves(int64_t n, const int32_t *morton, uint32_t *lact, uint32_t *lhas,
                             int32_t *list) {
  const int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l >= n) return;
  lact[l] = 1u;
  lhas[l] = 0u;
  list[l] = morton[l];
}

}  // namespace

hipError_t launch_clear_leaves(fsl_ctx *c, const LeafList &L) {
  if (L.n == 0) return hipSuccess;
  k_clear<<<leaf_grid(L.n, L.count), 512, 0, c->stream>>>(L, c->v[0], c->v[1], c->v[2], c->u[0],
                                                           c->u[1], c->u[2], c->vflag, c->marker,
                                                           c->lab, c->div, c->pres);
  return hipGetLastError();
}
// Workgroups of a per-leaf kernel over the active leaves.
inline int act_grid(const fsl_ctx *c) { return leaf_grid(c->n_act, c->counts_on_device); }
hipError_t launch_bc(fsl_ctx *c) {
  if (c->n_act == 0) return hipSuccess;
  k_bc<<<act_grid(c), 512, 0, c->stream>>>(c->dom, act_leaves(c), c->solid, c->marker, c->vflag,
                                            c->v[0], c->v[1], c->v[2], 0, nullptr, nullptr,
                                            nullptr, nullptr);
  return hipGetLastError();
}
hipError_t launch_bc_save(fsl_ctx *c) {
  if (c->n_act == 0) return hipSuccess;
  k_bc<<<act_grid(c), 512, 0, c->stream>>>(c->dom, act_leaves(c), c->solid, c->marker, c->vflag,
                                            c->v[0], c->v[1], c->v[2], 1, c->u[0], c->u[1],
                                            c->u[2], c->vflag);
  return hipGetLastError();
}
hipError_t launch_save(fsl_ctx *c) {
  if (c->n_act == 0) return hipSuccess;
  k_save<<<act_grid(c), 512, 0, c->stream>>>(c->dom, act_leaves(c), c->v[0], c->v[1], c->v[2],
                                              c->u[0], c->u[1], c->u[2], c->vflag);
  return hipGetLastError();
}
hipError_t launch_gravity(fsl_ctx *c, float dt, const float *dt_dev) {
  if (c->n_act == 0) return hipSuccess;
  k_gravity<<<act_grid(c), 512, 0, c->stream>>>(c->dom, act_leaves(c), c->vflag, c->v[0],
                                                 c->v[1], c->v[2], dt, dt_dev, c->force[0],
                                                 c->force[1], c->force[2]);
  return hipGetLastError();
}
hipError_t launch_divergence(fsl_ctx *c) {
  if (c->n_act == 0) return hipSuccess;
  k_divergence<<<act_grid(c), 512, 0, c->stream>>>(c->dom, act_leaves(c), c->marker, c->v[0],
                                                    c->v[1], c->v[2], c->div, c->lab, c->lfluid,
                                                    c->flags);
  k_row_totals<<<kRowTotalGroups, 256, 0, c->stream>>>(c->dom, c->nleaves, c->lfluid, c->flags);
  return hipGetLastError();
}
hipError_t launch_pressure_update(fsl_ctx *c) {
  if (c->n_act == 0) return hipSuccess;
  k_pressure_update<<<act_grid(c), 512, 0, c->stream>>>(c->dom, act_leaves(c), c->marker,
                                                         c->pres, c->v[0], c->v[1], c->v[2],
                                                         c->vflag);
  return hipGetLastError();
}
hipError_t launch_velocity_update(fsl_ctx *c) {
  if (c->n_act == 0) return hipSuccess;
  k_velocity_update<<<act_grid(c), 512, 0, c->stream>>>(c->dom, act_leaves(c), c->vflag,
                                                         c->v[0], c->v[1], c->v[2], c->u[0],
                                                         c->u[1], c->u[2], c->flags);
  return hipGetLastError();
}
hipError_t launch_extrapolate(fsl_ctx *c) {
  for (int axis = 0; axis < 3; ++axis) {
    const int a = axis == 0 ? 1 : 0, b = axis == 2 ? 1 : 2;
    const int64_t n = 2 * int64_t(c->dom.n[a] + 1) * (c->dom.n[b] + 1);
    k_extrapolate<<<blocks_for(n, kThreads), kThreads, 0, c->stream>>>(
        c->dom, axis, c->v[0], c->v[1], c->v[2], c->u[0], c->u[1], c->u[2], c->vflag);
  }
  return hipGetLastError();
}
hipError_t launch_clear_shell(fsl_ctx *c) {
  const int64_t nb = int64_t(c->dom.n[0] + 2) * (c->dom.n[1] + 2) * (c->dom.n[2] + 2);
  k_clear_shell<<<blocks_for(nb, kThreads), kThreads, 0, c->stream>>>(c->dom, c->vflag);
  return hipGetLastError();
}
hipError_t launch_cfl(fsl_ctx *c) {
  if (c->n_act == 0) return hipSuccess;
  const int g = int(std::min<int64_t>(c->n_act, kCflBlocks));
  k_cfl<<<g, 512, 0, c->stream>>>(c->dom, act_leaves(c), c->vflag, c->v[0], c->v[1], c->v[2],
                                  &c->flags->cfl_bits);
  return hipGetLastError();
}
hipError_t launch_dt(fsl_ctx *c, const DtParams &p) {
  k_dt<<<1, 64, 0, c->stream>>>(&c->flags->cfl_bits, p, c->sstate);
  return hipGetLastError();
}
// Fills n words (in place of a memset, which must not be captured in a graph).
hipError_t launch_fill_u32(fsl_ctx *c, uint32_t *p, int64_t n, uint32_t value) {
  if (n <= 0) return hipSuccess;
  k_fill_u32<<<blocks_for(n, kThreads), kThreads, 0, c->stream>>>(p, n, value);
  return hipGetLastError();
}
hipError_t launch_box_vec(fsl_ctx *c, int which, float *out3, uint8_t *act, bool to_box) {
  float **g = which == FSL_GRID_VELOCITY ? c->v : c->u;
  const int64_t nb = int64_t(c->dom.n[0] + 2) * (c->dom.n[1] + 2) * (c->dom.n[2] + 2);
  k_box_vec<<<blocks_for(nb, kThreads), kThreads, 0, c->stream>>>(
      c->dom, g[0], g[1], g[2], c->vflag, which == FSL_GRID_VELOCITY ? 0 : 1, out3, act,
      to_box ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_box_scalar(fsl_ctx *c, float *grid, float *box, bool to_box) {
  const int64_t nb = int64_t(c->dom.n[0] + 2) * (c->dom.n[1] + 2) * (c->dom.n[2] + 2);
  k_box_scalar<<<blocks_for(nb, kThreads), kThreads, 0, c->stream>>>(c->dom, grid, box,
                                                                      to_box ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_box_u8(fsl_ctx *c, uint8_t *grid, int32_t *box, bool to_box) {
  const int64_t nb = int64_t(c->dom.n[0] + 2) * (c->dom.n[1] + 2) * (c->dom.n[2] + 2);
  k_box_u8<<<blocks_for(nb, kThreads), kThreads, 0, c->stream>>>(c->dom, grid, box,
                                                                  to_box ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_all_leaves(fsl_ctx *c) {
  k_all_leaves<<<blocks_for(c->nleaves, kThreads), kThreads, 0, c->stream>>>(
      c->nleaves, c->morton, c->lact, c->lhas, c->act_list);
  return hipGetLastError();
}

}  // namespace fsl
