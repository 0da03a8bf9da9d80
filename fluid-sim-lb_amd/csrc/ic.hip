// Block incomplete-Cholesky preconditioner, IC(0), for the two-pass PCG.
//
// The reference's IC branch (BuildPreconditioner, solver.cc:103-106) builds
// OpenVDB 4.0.1's pcg::IncompleteCholeskyPreconditioner on each partition's
// interior matrix; openvdb.patch:209-224 skips the ghost columns, so the
// factor is block-diagonal over partitions (the tiles here). Rows are in the
// lexicographic order of CreateIndexAndMaskTree (solver.cc:114-151): x
// outermost, z innermost.
//
// For the 7-point Laplacian (diagonal = count of FLUID or AIR neighbours, -1
// per FLUID neighbour) two lower neighbours of a cell are never neighbours of
// each other, so IC(0) produces no update off the diagonal and the factor is
//   L_mm = sqrt(A_mm - sum over lower coupled neighbours n of L_mn^2)
//   L_im = A_mi / L_mm = -1 / L_mm   for every upper coupled neighbour i of m,
// where "coupled" means FLUID and in the same tile. Each cell stores
// d = L_mm and o = -1 / L_mm (the value of every entry in column m of L, and
// of every entry in row m of L^T). The solves are
//   forward:  y_i = (r_i - sum over x-1, y-1, z-1 of o_m y_m) / d_i
//   backward: z_i = (y_i - sum over z+1, y+1, x+1 of o_i z_m) / d_i
// with the sums in ascending column order, as the triangular row dot runs.
// A pivot below 1e-5 fails the build (the preconditioner's compatibility
// test), reported through PcgState::pc_fail.
//
// Schedule. Cell i depends only on its three lower (upper) neighbours, so any
// order that finishes them first gives the same bits as the sequential loop.
// Leaves go by tile-local hyperplanes: a leaf's level is the sum of its
// leaf coordinates within its tile, and each level is one launch (levels
// ascending for the factor and the forward solve, descending for the
// backward solve). A workgroup of one wavefront takes one leaf: it stages the
// leaf and the three face halos it depends on in LDS, then walks the 22 cell
// hyperplanes x + y + z = t of the leaf (at most 48 cells each, one per lane).
#include <algorithm>
#include <vector>

#include "device_util.h"
#include "fsl_internal.h"

namespace fsl {

namespace {

constexpr int kW = 64;  // one wavefront per leaf

__device__ inline bool fluid_lab(uint8_t l) { return (l & 3) == kFluid; }

// LDS box [-1, 9)^3 of a leaf: index (x+1)*100 + (y+1)*10 + (z+1).
__device__ inline int bix(int x, int y, int z) { return (x + 1) * 100 + (y + 1) * 10 + (z + 1); }
__device__ inline int bix_local(int c) { return bix(c >> 6, (c >> 3) & 7, c & 7); }

// Cell number l (x ascending, then y) of the hyperplane x + y + z = t of a
// leaf, as its local offset x<<6 | y<<3 | z; -1 past the plane's end.
__device__ inline int plane_cell(int t, int l) {
  int c = -1;
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    const int s = t - x;  // y + z
    const int lo = s > 7 ? s - 7 : 0, hi = s < 7 ? s : 7;
    const int cnt = (s >= 0 && hi >= lo) ? hi - lo + 1 : 0;
    if (c < 0 && l >= 0 && l < cnt) {
      const int y = lo + l;
      c = x << 6 | y << 3 | (s - y);
    }
    l -= cnt;
  }
  return c;
}

// Face halo cell (a, b) of face f (0: x, 1: y, 2: z) on the lower (up = false)
// or upper side: its voxel in the neighbouring leaf and its LDS index.
__device__ inline void face_cell(const Dom &d, int64_t leaf, int f, bool up, int a, int b,
                                 int64_t &v, int &bi) {
  const int64_t l12 = int64_t(d.L[1]) * d.L[2];
  const int64_t step = f == 0 ? l12 : (f == 1 ? int64_t(d.L[2]) : 1);
  const int64_t nb = up ? leaf + step : leaf - step;
  const int layer = up ? 0 : 7, side = up ? 8 : -1;
  int off;
  if (f == 0) { off = layer << 6 | a << 3 | b; bi = bix(side, a, b); }
  else if (f == 1) { off = a << 6 | layer << 3 | b; bi = bix(a, side, b); }
  else { off = a << 6 | b << 3 | layer; bi = bix(a, b, side); }
  v = (nb << 9) | off;
}

// Whether the leaf couples to its neighbour across face f: both sides in one
// tile (tiles are whole leaves; a leaf holding FLUID cells lies in [0, n)).
__device__ inline bool face_coupled(const Dom &d, const int o[3], int f, bool up) {
  const int w = d.n[f] / d.T[f];
  const int edge = up ? o[f] + 8 : o[f];
  return edge % w != 0;
}

template <typename V>
__device__ inline V ic_sqrt(V s) { return sqrt(s); }

// Factor one level: d and o of every cell of the level's leaves (0 on
// non-FLUID cells).
template <typename V>
__global__ __launch_bounds__(kW) void k_ic_factor(Dom d, const int32_t *leaves,
                                                  const uint8_t *lab, V *ld, V *lo,
                                                  PcgState *st) {
#pragma clang fp contract(off)
  __shared__ V s_d[1000];
  __shared__ uint8_t s_f[1000];
  const int64_t leaf = leaves[blockIdx.x];
  const int lane = threadIdx.x;
  int o[3];
  leaf_origin(d, leaf, o);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = lane + kW * j;
    const uint8_t l = lab[(leaf << 9) | c];
    const bool f = fluid_lab(l);
    s_f[bix_local(c)] = f;
    s_d[bix_local(c)] = f ? V(l >> 2) : V(0);  // A_mm before the updates
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    int64_t hv;
    int bi;
    face_cell(d, leaf, f, false, lane >> 3, lane & 7, hv, bi);
    const bool fl = face_coupled(d, o, f, false) && fluid_lab(lab[hv]);
    s_f[bi] = fl;
    s_d[bi] = fl ? ld[hv] : V(0);
  }
  __syncthreads();
  bool fail = false;
#pragma unroll
  for (int t = 0; t < 22; ++t) {
    const int c = plane_cell(t, lane);
    if (c >= 0) {
      const int x = c >> 6, y = (c >> 3) & 7, z = c & 7;
      const int bi = bix(x, y, z);
      if (s_f[bi]) {
        const int nb[3] = {bix(x - 1, y, z), bix(x, y - 1, z), bix(x, y, z - 1)};
        V s = s_d[bi];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          if (!s_f[nb[q]]) continue;
          const V om = V(-1) / s_d[nb[q]];
          const V m = om * om;
          s = s - m;
        }
        if (s < V(1.e-5f)) {
          fail = true;
          s = V(1);
        }
        s_d[bi] = ic_sqrt(s);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = lane + kW * j;
    const int64_t v = (leaf << 9) | c;
    const int bi = bix_local(c);
    const bool f = s_f[bi];
    const V dd = f ? s_d[bi] : V(0);
    ld[v] = dd;
    lo[v] = f ? V(-1) / dd : V(0);
  }
  if (fail) st->pc_fail = 1;
}

// Forward solve L y = r over one level (y is 0 on non-FLUID cells).
template <typename V>
__global__ __launch_bounds__(kW) void k_ic_forward(Dom d, const int32_t *leaves,
                                                   const uint8_t *lab, const V *ld, const V *lo,
                                                   const V *r, V *y, const PcgState *st,
                                                   int force) {
#pragma clang fp contract(off)
  if ((!force && st->done) || st->status) return;
  __shared__ V s_y[1000], s_o[1000], s_r[512], s_d[512];
  __shared__ uint8_t s_f[1000];
  const int64_t leaf = leaves[blockIdx.x];
  const int lane = threadIdx.x;
  int o[3];
  leaf_origin(d, leaf, o);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = lane + kW * j;
    const int64_t v = (leaf << 9) | c;
    const int bi = bix_local(c);
    const bool f = fluid_lab(lab[v]);
    s_f[bi] = f;
    s_r[c] = r[v];
    s_d[c] = ld[v];
    s_o[bi] = lo[v];
    s_y[bi] = V(0);
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    int64_t hv;
    int bi;
    face_cell(d, leaf, f, false, lane >> 3, lane & 7, hv, bi);
    const bool fl = face_coupled(d, o, f, false) && fluid_lab(lab[hv]);
    s_f[bi] = fl;
    s_o[bi] = fl ? lo[hv] : V(0);
    s_y[bi] = fl ? y[hv] : V(0);
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 22; ++t) {
    const int c = plane_cell(t, lane);
    if (c >= 0) {
      const int x = c >> 6, yy = (c >> 3) & 7, z = c & 7;
      const int bi = bix(x, yy, z);
      if (s_f[bi]) {
        const int nb[3] = {bix(x - 1, yy, z), bix(x, yy - 1, z), bix(x, yy, z - 1)};
        V dot = V(0);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          if (!s_f[nb[q]]) continue;
          const V m = s_o[nb[q]] * s_y[nb[q]];
          dot = dot + m;
        }
        s_y[bi] = (s_r[c] - dot) / s_d[c];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = lane + kW * j;
    y[(leaf << 9) | c] = s_y[bix_local(c)];
  }
}

// Backward solve L^T z = y over one level (z is 0 on non-FLUID cells).
template <typename V>
__global__ __launch_bounds__(kW) void k_ic_backward(Dom d, const int32_t *leaves,
                                                    const uint8_t *lab, const V *ld, const V *lo,
                                                    const V *y, V *z, const PcgState *st,
                                                    int force) {
#pragma clang fp contract(off)
  if ((!force && st->done) || st->status) return;
  __shared__ V s_z[1000], s_y[512], s_d[512], s_o[512];
  __shared__ uint8_t s_f[1000];
  const int64_t leaf = leaves[blockIdx.x];
  const int lane = threadIdx.x;
  int o[3];
  leaf_origin(d, leaf, o);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = lane + kW * j;
    const int64_t v = (leaf << 9) | c;
    const int bi = bix_local(c);
    s_f[bi] = fluid_lab(lab[v]);
    s_y[c] = y[v];
    s_d[c] = ld[v];
    s_o[c] = lo[v];
    s_z[bi] = V(0);
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    int64_t hv;
    int bi;
    face_cell(d, leaf, f, true, lane >> 3, lane & 7, hv, bi);
    const bool fl = face_coupled(d, o, f, true) && fluid_lab(lab[hv]);
    s_f[bi] = fl;
    s_z[bi] = fl ? z[hv] : V(0);
  }
  __syncthreads();
#pragma unroll
  for (int t = 21; t >= 0; --t) {
    const int c = plane_cell(t, lane);
    if (c >= 0) {
      const int x = c >> 6, yy = (c >> 3) & 7, zz = c & 7;
      const int bi = bix(x, yy, zz);
      if (s_f[bi]) {
        const int nb[3] = {bix(x, yy, zz + 1), bix(x, yy + 1, zz), bix(x + 1, yy, zz)};
        const V oi = s_o[c];
        V dot = V(0);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          if (!s_f[nb[q]]) continue;
          const V m = oi * s_z[nb[q]];
          dot = dot + m;
        }
        s_z[bi] = (s_y[c] - dot) / s_d[c];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = lane + kW * j;
    z[(leaf << 9) | c] = s_z[bix_local(c)];
  }
}

template <typename V>
void launch_levels(fsl_ctx *c, int force, bool factor) {
  const int nlev = int(c->ic_off.size()) - 1;
  V *ld = static_cast<V *>(c->ic_d), *lo = static_cast<V *>(c->ic_o);
  if (factor) {
    for (int l = 0; l < nlev; ++l) {
      const int n = c->ic_off[l + 1] - c->ic_off[l];
      if (n > 0)
        k_ic_factor<V><<<n, kW, 0, c->stream>>>(c->dom, c->ic_leaves + c->ic_off[l], c->lab, ld,
                                                 lo, c->pcg);
    }
    return;
  }
  // y lives in the Ap buffer: A(k) overwrites it before reading.
  V *y = static_cast<V *>(c->Ap);
  for (int l = 0; l < nlev; ++l) {
    const int n = c->ic_off[l + 1] - c->ic_off[l];
    if (n > 0)
      k_ic_forward<V><<<n, kW, 0, c->stream>>>(c->dom, c->ic_leaves + c->ic_off[l], c->lab, ld,
                                                lo, static_cast<const V *>(c->r), y, c->pcg,
                                                force);
  }
  for (int l = nlev - 1; l >= 0; --l) {
    const int n = c->ic_off[l + 1] - c->ic_off[l];
    if (n > 0)
      k_ic_backward<V><<<n, kW, 0, c->stream>>>(c->dom, c->ic_leaves + c->ic_off[l], c->lab, ld,
                                                 lo, y, static_cast<V *>(c->z), c->pcg, force);
  }
}

}  // namespace

int ic_factor(fsl_ctx *ctx) {
  fsl_ctx *c = ctx;
  if (c->n_fluid <= 0) return FSL_OK;
  if (!c->ic_d) {
    FSL_HIP(hipMalloc(&c->ic_d, size_t(c->nvox) * sizeof(double)));
    FSL_HIP(hipMalloc(&c->ic_o, size_t(c->nvox) * sizeof(double)));
    FSL_HIP(hipMalloc(reinterpret_cast<void **>(&c->ic_leaves),
                      size_t(c->nleaves) * sizeof(int32_t)));
  }
  // Fluid leaves sorted by tile-local level (the list is made on the stream).
  std::vector<int32_t> fl(size_t(c->n_fluid)), sorted(size_t(c->n_fluid));
  FSL_HIP(hipStreamSynchronize(c->stream));
  FSL_HIP(hipMemcpy(fl.data(), c->fluid_list, fl.size() * sizeof(int32_t),
                    hipMemcpyDeviceToHost));
  const Dom &d = c->dom;
  int nlev = 1;
  for (int a = 0; a < 3; ++a) nlev += d.n[a] / d.T[a] / 8 - 1;
  std::vector<int> lev(fl.size());
  c->ic_off.assign(size_t(nlev) + 1, 0);
  for (size_t q = 0; q < fl.size(); ++q) {
    int o[3];
    leaf_origin(d, fl[q], o);
    int s = 0;
    for (int a = 0; a < 3; ++a) s += (o[a] % (d.n[a] / d.T[a])) / 8;
    bool inside = s < nlev;
    for (int a = 0; a < 3; ++a) inside = inside && o[a] >= 0 && o[a] + 8 <= d.n[a];
    if (!inside) {
      c->err = "fluid leaf outside the domain";
      return FSL_ERR_STATE;
    }
    lev[q] = s;
    ++c->ic_off[size_t(s) + 1];
  }
  for (int l = 0; l < nlev; ++l) c->ic_off[size_t(l) + 1] += c->ic_off[size_t(l)];
  std::vector<int> fill(c->ic_off.begin(), c->ic_off.end() - 1);
  for (size_t q = 0; q < fl.size(); ++q) sorted[size_t(fill[size_t(lev[q])]++)] = fl[q];
  FSL_HIP(hipMemcpyAsync(c->ic_leaves, sorted.data(), sorted.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, c->stream));
  if (c->pcg_precision == FSL_PCG_FP64) launch_levels<double>(c, 0, true);
  else launch_levels<float>(c, 0, true);
  FSL_HIP(hipGetLastError());
  // The upload reads `sorted` on the stream: wait before it goes out of scope.
  FSL_HIP(hipStreamSynchronize(c->stream));
  return FSL_OK;
}

void ic_sweeps(fsl_ctx *c, int force) {
  if (c->n_fluid <= 0) return;
  if (c->pcg_precision == FSL_PCG_FP64) launch_levels<double>(c, force, false);
  else launch_levels<float>(c, force, false);
}

void ic_free(fsl_ctx *c) {
  for (void *p : {c->ic_d, c->ic_o, static_cast<void *>(c->ic_leaves)})
    if (p) (void)hipFree(p);
  c->ic_d = c->ic_o = nullptr;
  c->ic_leaves = nullptr;
}

}  // namespace fsl
