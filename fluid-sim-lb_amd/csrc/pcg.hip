// Jacobi-preconditioned CG for the pressure Poisson system, matrix-free.
//
// The reference assembles an explicit 7-point SparseStencilMatrix and runs
// PCG in three stages per iteration with three scalar all-reduces
// (solver.cc:262-366, solver_driver.cc:318-343). Here the operator is applied
// from the 1-byte solver label of each cell (marker | diag << 2): row c has
// diag(c) on the diagonal and -1 for every FLUID neighbour (solver.cc:
// 214-231). Iteration k is two kernels over the fluid leaves:
//   A(k): beta = rz_k / rz_{k-1}; p = z + beta p; Ap; p.Ap
//   B(k): alpha = rz_k / pAp; x += alpha p; r -= alpha Ap; z = r / diag; r.r; r.z
// Each kernel publishes one partial sum per workgroup, and every workgroup of
// the next kernel sums those partials itself, in the same fixed order, so all
// workgroups hold bit-identical scalars: no atomics, no fences, no extra
// reduction launches, and the host only polls convergence every `batch`
// iterations. B's partials are double-buffered by iteration parity because
// A(k) needs rz_k and rz_{k-1}.
//
// Workgroups are persistent over the leaf list: G = 8 x (blocks per XCD),
// and XCD x walks the contiguous chunk x of the (spatially ordered) list, so a
// leaf's halo neighbours are usually processed by the same XCD and hit its L2.
// A group of 256 threads handles one leaf at a time (two cells per thread),
// staging p of the leaf and its six face halos in LDS; halo values of p are
// recomputed from z and the previous p, so p is double-buffered.
//
// Precision modes:
//   FP32  - the reference's precision: float vectors, float dot products,
//           the stencil summed in the reference's column order without FMA.
//   FP64  - double everywhere.
//   MIXED - float vectors and stencil, double dot products (exact products
//           accumulated by FMA in each thread, then DPP trees and fixed-order
//           workgroup and cross-workgroup sums in double: r.z and r.r are
//           sums of positive terms, for which the rounding stays below about
//           1e-14 relative; the two-pass path's totals are
//           Neumaier-compensated), and every `refresh`
//           iterations the float iterate is folded into a double solution and
//           the true residual b - A x is recomputed in double (a reliable
//           update; CG continues with the same direction).
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "device_util.h"
#include "fsl_internal.h"

namespace fsl {

namespace {

constexpr int kG = 256;             // threads per leaf group (two cells each)
constexpr int kGroups = 4;          // leaves in flight per workgroup
constexpr int kT = kG * kGroups;    // threads per workgroup
constexpr int kBlocksPerXcd = 64;   // 2 workgroups per CU, 32 CUs per XCD

// 1 / diag for diag in 0..6 (0 for 0). JacobiPreconditioner stores
// float(1.0 / diag) for float vectors. The value is selected in registers: a
// table indexed at run time would live in global memory, and each lookup
// would wait for every load in flight, including the next leaf's prefetch.
template <typename V>
__device__ inline V dinv(int diag) {
  const V r3 = V(1.0 / 3.0), r5 = V(1.0 / 5.0), r6 = V(1.0 / 6.0);
  V v = diag == 1 ? V(1) : V(0);
  v = diag == 2 ? V(0.5) : v;
  v = diag == 3 ? r3 : v;
  v = diag == 4 ? V(0.25) : v;
  v = diag == 5 ? r5 : v;
  v = diag == 6 ? r6 : v;
  return v;
}

__device__ inline bool is_fluid(uint8_t lab) { return (lab & 3) == kFluid; }

// LDS tile of a leaf plus its six face halos: index (x+1)*100 + (y+1)*10 + z+1.
__device__ inline int tix(int x, int y, int z) { return (x + 1) * 100 + (y + 1) * 10 + (z + 1); }
__device__ inline int tix_of_local(int c) { return tix(c >> 6, (c >> 3) & 7, c & 7); }

// Halo cell h in [0, 384): face h/64 (x-, x+, y-, y+, z-, z+).
__device__ inline void halo_cell(int h, int &x, int &y, int &z) {
  const int f = h >> 6, a = (h >> 3) & 7, b = h & 7;
  const int s = (f & 1) ? 8 : -1;
  if (f < 2) { x = s; y = a; z = b; }
  else if (f < 4) { x = a; y = s; z = b; }
  else { x = a; y = b; z = s; }
}

// Halo slot s of thread t in a 256-thread leaf group: slot 0 covers the
// x and y faces (one face per wave), slot 1 the two z faces spread over the
// first 32 lanes of every wave, so each wave carries the same halo work.
// Returns -1 when the slot is idle.
__device__ inline int halo_of(int t, int s) {
  if (s == 0) return t;
  const int lane = t & 63;
  return lane < 32 ? 256 + (t >> 6) * 32 + lane : -1;
}
// Voxel of halo cell h of `leaf`: the neighbour leaf across face h / 64 and
// the mirrored layer inside it.
__device__ inline int64_t halo_voxel(const Dom &d, int64_t leaf, int h) {
  const int f = h >> 6, a = (h >> 3) & 7, b = h & 7;
  const int64_t l12 = int64_t(d.L[1]) * d.L[2];
  const int64_t step = f < 2 ? l12 : (f < 4 ? int64_t(d.L[2]) : 1);
  const int64_t nb = (f & 1) ? leaf + step : leaf - step;
  const int layer = (f & 1) ? 0 : 7;
  int off;
  if (f < 2) off = layer << 6 | a << 3 | b;
  else if (f < 4) off = a << 6 | layer << 3 | b;
  else off = a << 6 | b << 3 | layer;
  return (nb << 9) | off;
}

// Apply the operator at LDS index c.
template <typename V, bool ORD>
__device__ inline V stencil(const V *sp, const uint8_t *sf, int c, int diag) {
#pragma clang fp contract(off)
  if (ORD) {
    // Row dot over columns in ascending row-index order (solver.cc:127-148):
    // x-1, y-1, z-1, self, z+1, y+1, x+1.
    V s = V(0);
    if (sf[c - 100]) s = s + (-sp[c - 100]);
    if (sf[c - 10]) s = s + (-sp[c - 10]);
    if (sf[c - 1]) s = s + (-sp[c - 1]);
    { const V m = V(diag) * sp[c]; s = s + m; }
    if (sf[c + 1]) s = s + (-sp[c + 1]);
    if (sf[c + 10]) s = s + (-sp[c + 10]);
    if (sf[c + 100]) s = s + (-sp[c + 100]);
    return s;
  } else {
    V nb = V(0);
    nb += sf[c - 100] ? sp[c - 100] : V(0);
    nb += sf[c + 100] ? sp[c + 100] : V(0);
    nb += sf[c - 10] ? sp[c - 10] : V(0);
    nb += sf[c + 10] ? sp[c + 10] : V(0);
    nb += sf[c - 1] ? sp[c - 1] : V(0);
    nb += sf[c + 1] ? sp[c + 1] : V(0);
    return fmav(V(diag), sp[c], -nb);
  }
}

// Operator from a tile in which every non-FLUID cell holds 0 (so no flags).
template <typename V, bool ORD>
__device__ inline V stencil0(const V *sp, int c, int diag) {
#pragma clang fp contract(off)
  if (ORD) {
    // Column order of the reference's row dot (solver.cc:127-148); adding
    // -0 for a non-FLUID neighbour leaves the sum unchanged.
    V s = V(0);
    s = s + (-sp[c - 100]);
    s = s + (-sp[c - 10]);
    s = s + (-sp[c - 1]);
    { const V m = V(diag) * sp[c]; s = s + m; }
    s = s + (-sp[c + 1]);
    s = s + (-sp[c + 10]);
    s = s + (-sp[c + 100]);
    return s;
  } else {
    const V nb = ((sp[c - 100] + sp[c + 100]) + (sp[c - 10] + sp[c + 10])) + (sp[c - 1] + sp[c + 1]);
    return fmav(V(diag), sp[c], -nb);
  }
}

// Persistent walk over the leaf list: XCD x = blockIdx % 8 owns the list
// chunk x; its workgroups stride through it, blockDim / kG leaves per round
// (leaf group g = threadIdx / kG takes every (blockDim / kG)-th leaf of the
// workgroup).
struct Walk {
  int64_t w, end, step, ng;
  __device__ bool more() const { return w < end; }
  __device__ int64_t mine() const { return w + int64_t(threadIdx.x / kG) * step; }
  __device__ bool valid() const { return mine() < end; }
  __device__ void next() { w += ng * step; }
};
// NT: the workgroup size when the kernel fixes it (0: read blockDim, which
// is a memory load at the kernel's start).
template <int NT = 0>
__device__ inline Walk walk(int64_t nf) {
  const int64_t per = gridDim.x / kXcd;
  const int64_t chunk = (nf + kXcd - 1) / kXcd;
  const int64_t beg = int64_t(blockIdx.x % kXcd) * chunk;
  const int nt = NT ? NT : int(blockDim.x);
  return {beg + blockIdx.x / kXcd, beg + chunk < nf ? beg + chunk : nf, per, int64_t(nt / kG)};
}
__device__ inline int gtid() { return threadIdx.x % kG; }

// Block-reduce K values and publish this workgroup's partial sums: wave
// sums, one barrier, then thread 0 adds the waves' sums in order.
template <typename A, int K, int NT = 0>
__device__ inline void publish(A vals[K], A *partials) {
  __shared__ A s_red[K][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = ((NT ? NT : int(blockDim.x)) + 63) >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) vals[k] = wave_sum(vals[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) s_red[k][wave] = vals[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      A t = A(0);
      for (int w = 0; w < nw; ++w) t += s_red[k][w];
      vals[k] = t;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) partials[int64_t(blockIdx.x) * K + k] = vals[k];
  }
}

// Total of value k over the n published partials (stride K), summed by the
// whole workgroup in a fixed order: every workgroup gets the same bits.
template <typename A, int K>
__device__ inline double total(const A *partials, int n, int k) {
  if constexpr (sizeof(A) == 8) {
    __shared__ double s_d[16];
    double s = 0.0, c = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) neumaier_add(s, c, double(partials[q * K + k]));
    return block_sum<double>(s + c, s_d);
  } else {
    __shared__ A s_a[16];
    A s = A(0);
    for (int q = threadIdx.x; q < n; q += blockDim.x) s += partials[q * K + k];
    return double(block_sum<A>(s, s_a));
  }
}

// Totals of all K values of the n published (stride K) partials in one pass:
// one load round and one combined workgroup reduction, in A's precision, the
// same in every workgroup. A thread adds at most a few partials (n is at most
// a few thousand workgroups), so plain sums in a fixed order suffice: the
// rounding error is a few units of 1e-16 relative to the sum of magnitudes.
template <typename A, int K>
__device__ inline void total_all(const A *partials, int n, double out[K]) {
  __shared__ A s_t[K][16];
  A v[K];
#pragma unroll
  for (int j = 0; j < K; ++j) v[j] = A(0);
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
#pragma unroll
    for (int j = 0; j < K; ++j) v[j] += partials[K * q + j];
  }
#pragma unroll
  for (int j = 0; j < K; ++j) v[j] = wave_sum(v[j]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < K; ++j) s_t[j][wave] = v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < K; ++j) {
    A t = A(0);
    for (int w = 0; w < nw; ++w) t += s_t[j][w];
    out[j] = double(t);
  }
}

// The same totals in two steps, so that the loads of the partials are in
// flight while the caller issues other loads: parts_issue loads partials
// t, t + NT, ..., t + (R - 1) NT of thread t (NT threads) at once (a round
// past n re-reads partial 0 and is discarded), parts_total adds them (and,
// beyond R rounds, the rest in a loop) in the order of total_all, so the
// bits are total_all's.
template <typename A, int K, int R>
struct PartRegs {
  A v[R][K];
};
template <typename A, int K, int NT, int R>
__device__ inline void parts_issue(const A *partials, int n, PartRegs<A, K, R> &pr) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int q = int(threadIdx.x) + r * NT;
    const int qq = q < n ? q : 0;
#pragma unroll
    for (int j = 0; j < K; ++j) pr.v[r][j] = partials[K * qq + j];
  }
}
template <typename A, int K, int NT, int R>
__device__ inline void parts_total(const A *partials, int n, const PartRegs<A, K, R> &pr,
                                   double out[K]) {
  __shared__ A s_t[K][16];
  A v[K];
#pragma unroll
  for (int j = 0; j < K; ++j) v[j] = A(0);
  if (n >= R * NT) {
    // Every round holds a partial (the usual case: one partial per
    // workgroup of a full grid), so no round is masked.
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int j = 0; j < K; ++j) v[j] += pr.v[r][j];
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (int(threadIdx.x) + r * NT < n) {
#pragma unroll
        for (int j = 0; j < K; ++j) v[j] += pr.v[r][j];
      }
    }
  }
  for (int q = int(threadIdx.x) + R * NT; q < n; q += NT) {
#pragma unroll
    for (int j = 0; j < K; ++j) v[j] += partials[K * q + j];
  }
#pragma unroll
  for (int j = 0; j < K; ++j) v[j] = wave_sum(v[j]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int nw = (NT + 63) >> 6;
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < K; ++j) s_t[j][wave] = v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < K; ++j) {
    A t = A(0);
    for (int w = 0; w < nw; ++w) t += s_t[j][w];
    out[j] = double(t);
  }
}

// Mode bits of single-reduction iteration k: bit 0, x is updated (odd k);
// bit 1, x lacks alpha_{k-1} p_{k-1} (odd k, except the iteration a solve
// resumed at); bit 2, x is folded into x64 before a scheduled refresh.
inline int sr_mode_of(int k, int resume_k, int refresh) {
  return ((k & 1) ? (k == resume_k ? 1 : 3) : 0) | (refresh > 0 && (k + 1) % refresh == 0 ? 4 : 0);
}

__device__ inline bool converged(int stop_mode, double thr, double rr) {
  // Relative mode uses <=, so a zero right-hand side (x0 = 0) stops at once.
  return stop_mode == FSL_STOP_ABS_RR ? rr < thr : sqrt(rr) <= thr;
}
__device__ inline bool converged(const PcgState *st, double rr) {
  return converged(st->stop_mode, st->thr, rr);
}

// Loads p of a leaf and its halo into LDS, where cell values come from
// value(global voxel index, label) and non-FLUID cells hold 0.
template <typename V, typename F>
__device__ inline void fill_tile(const Dom &d, const uint8_t *lab, int64_t leaf, const int o[3],
                                 V *sp, uint8_t *sf, uint8_t own_lab[2], F value) {
  const int t = gtid();
  uint8_t hl[2] = {0, 0};
  int64_t hv[2] = {0, 0};
  int hix[2] = {0, 0};
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int c = t + s * kG;
    own_lab[s] = lab[(leaf << 9) | c];
    const int h = t + s * kG;
    if (h < 384) {
      int hx, hy, hz;
      halo_cell(h, hx, hy, hz);
      hv[s] = vox_of(d, o[0] + hx, o[1] + hy, o[2] + hz);
      hl[s] = lab[hv[s]];
      hix[s] = tix(hx, hy, hz);
    }
  }
  // Values are read for every cell and selected afterwards, so the data
  // loads do not wait for the label loads (storage covers every leaf, so the
  // reads stay in bounds; values of non-FLUID cells are discarded).
  V own_val[2], halo_val[2] = {V(0), V(0)};
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    own_val[s] = value((leaf << 9) | (t + s * kG), own_lab[s]);
    if (t + s * kG < 384) halo_val[s] = value(hv[s], hl[s]);
  }
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int c = t + s * kG;
    const bool f = is_fluid(own_lab[s]);
    sp[tix_of_local(c)] = f ? own_val[s] : V(0);
    sf[tix_of_local(c)] = f;
    if (t + s * kG < 384) {
      const bool hf = is_fluid(hl[s]);
      sp[hix[s]] = hf ? halo_val[s] : V(0);
      sf[hix[s]] = hf;
    }
  }
}

// PreparePCGSolver (solver.cc:262-286): x = x0 (0, or the pressure grid when
// warm starting), r = b - A x, z = M^-1 r; publishes (r.r, r.z) as B(-1) and
// b.b for the relative stopping rule.
template <typename V, typename A, bool ORD, bool MIXED>
__global__ __launch_bounds__(kT) void k_prepare(Dom d, const int32_t *flist, int64_t nf,
                                                const uint8_t *lab, const float *div,
                                                const float *pres, int warm, V *x, V *r,
                                                V *zv, double *x64, A *part_b, A *part_bb) {
#pragma clang fp contract(off)
  using W = typename std::conditional<MIXED, double, V>::type;  // residual arithmetic
  __shared__ W s_x[kGroups][1000];
  __shared__ uint8_t s_f[kGroups][1000];
  const int g = threadIdx.x / kG;
  A acc[3] = {A(0), A(0), A(0)};
  for (Walk q = walk(nf); q.more(); q.next()) {
    const bool valid = q.valid();
    const int64_t leaf = valid ? flist[q.mine()] : 0;
    int o[3];
    leaf_origin(d, leaf, o);
    uint8_t ol[2] = {0, 0};
    if (valid)
      fill_tile<W>(d, lab, leaf, o, s_x[g], s_f[g], ol,
                   [&](int64_t v, uint8_t) { return warm ? W(pres[v]) : W(0); });
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int c = gtid() + s * kG;
      const int64_t v = (leaf << 9) | c;
      if (!valid || !is_fluid(ol[s])) continue;
      const W ax = stencil<W, ORD>(s_x[g], s_f[g], tix_of_local(c), ol[s] >> 2);
      const float b = div[v];
      const V rv = V(W(b) - ax);
      const float x0 = warm ? pres[v] : 0.0f;
      if (MIXED) { x64[v] = double(x0); x[v] = V(0); }
      else x[v] = V(x0);
      r[v] = rv;
      const V z = rv * dinv<V>(ol[s] >> 2);
      if (zv) zv[v] = z;
      acc[0] += A(rv) * A(rv);
      acc[1] += A(rv) * A(z);
      acc[2] += A(b) * A(b);
    }
    __syncthreads();
  }
  A two[2] = {acc[0], acc[1]};
  publish<A, 2>(two, part_b);
  A one[1] = {acc[2]};
  publish<A, 1>(one, part_bb);
}

// z, p0 and p1 start at 0 on every cell of the active leaves (the PCG
// kernels keep them 0 off the FLUID cells).
template <typename V>
__global__ __launch_bounds__(kT) void k_zero(const int32_t *list, int64_t n, V *a, V *b, V *c) {
  for (Walk q = walk(n); q.more(); q.next()) {
    if (!q.valid()) continue;
    const int64_t base = int64_t(list[q.mine()]) << 9;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int64_t v = base | (gtid() + s * kG);
      a[v] = V(0); b[v] = V(0); c[v] = V(0);
    }
  }
}

// Scalars after preparation: thresholds and the initial stopping test.
template <typename A>
__global__ __launch_bounds__(1024) void k_prepare_scalars(const A *part_b, const A *part_bb,
                                                          int g, PcgState *st) {
  const double rr = total<A, 2>(part_b, g, 0), rz = total<A, 2>(part_b, g, 1);
  const double bb = total<A, 1>(part_bb, g, 0);
  if (threadIdx.x != 0) return;
  st->rr = rr;
  st->rz = rz;
  st->rzh[0] = rz;
  st->bnorm2 = bb;
  if (st->stop_mode == FSL_STOP_REL) st->thr = st->thr * sqrt(bb);
  st->it = 0;
  st->status = st->pc_fail ? FSL_ERR_PRECONDITIONER : (isnan(rr) ? FSL_ERR_NAN : 0);
  st->done = (st->status != 0 || converged(st, rr) || st->max_iter <= 0) ? 1 : 0;
}

// A(k): ComputeLocal3 of iteration k-1 (beta, p = z + beta p,
// solver.cc:355-366), the While test (solver_driver.cc:330) and ComputeLocal1
// (Ap, pAp, solver.cc:296-309). rz_{k-1} comes from the device scalars
// (slot (k-1)&1); rz_k is published there for B(k) and A(k+1).
// z and p hold 0 on every non-FLUID cell of the active leaves, so the halo
// needs no labels: p = z + beta p is 0 there and drops out of the stencil.
template <typename V, typename A, bool ORD>
__global__ __launch_bounds__(kT) void k_pass_a(Dom d, const int32_t *flist, int64_t nf,
                                               const uint8_t *lab, const V *zv, const V *pin,
                                               V *pout, V *Ap, const A *part_b1, A *part_a,
                                               int g_parts, int k, PcgState *st) {
#pragma clang fp contract(off)
  __shared__ V s_p[2][kGroups][1000];
  if (st->done) return;
  V beta = V(0);
  const int g = threadIdx.x / kG, t = gtid();
  A acc = A(0);
  auto pval = [&](V zz, V pp) {
    if (k == 0) return zz;
    const V m = beta * pp;
    return zz + m;
  };
  // Software pipeline: the loads of round j+1 are in flight while round j is
  // staged, synchronised and computed; the LDS tile is double-buffered, so
  // one barrier per round suffices.
  struct Pre {
    V oz[2], op[2], hz[2], hp[2];
    uint8_t ol[2];
    int hix[2];
    int64_t leaf;
    bool valid;
  };
  auto load = [&](const Walk &q, Pre &x) {
    x.valid = q.valid();
    x.leaf = x.valid ? flist[q.mine()] : 0;
    x.ol[0] = x.ol[1] = 0;
    if (!x.valid) return;
    int o[3];
    leaf_origin(d, x.leaf, o);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int64_t v = (x.leaf << 9) | (t + s * kG);
      x.ol[s] = lab[v];
      x.oz[s] = zv[v];
      x.op[s] = k ? pin[v] : V(0);
      const int h = halo_of(t, s);
      x.hz[s] = V(0);
      x.hp[s] = V(0);
      x.hix[s] = -1;
      if (h >= 0) {
        int hx, hy, hz_;
        halo_cell(h, hx, hy, hz_);
        const int64_t hv = halo_voxel(d, x.leaf, h);
        x.hz[s] = zv[hv];
        x.hp[s] = k ? pin[hv] : V(0);
        x.hix[s] = tix(hx, hy, hz_);
      }
    }
  };
  // The first leaf's loads do not depend on beta: issue them before the
  // scalar reduction, so the reduction runs under their latency.
  Walk q = walk(nf);
  Pre cur;
  if (q.more()) load(q, cur);
  if (k > 0) {
    double tot[2];
    total_all<A, 2>(part_b1, g_parts, tot);
    const double rr = tot[0];
    const A rz = A(tot[1]);
    const A rz_old = A(st->rzh[(k - 1) & 1]);
    const A b = rz / rz_old;
    int status = 0;
    if (isnan(rr)) status = FSL_ERR_NAN;
    else if (!(double(b) <= 1e5)) status = FSL_ERR_BETA;  // CHECK_LE(beta, LARGE_FLOAT), solver.cc:361
    const bool stop = status != 0 || converged(st, rr) || k >= st->max_iter;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->it = k;
      st->rr = rr;
      st->rz = double(rz);
      st->rz_old = double(rz_old);
      st->beta = double(b);
      st->rzh[k & 1] = double(rz);
      if (status) st->status = status;
      if (stop) st->done = 1;
    }
    if (stop) return;
    beta = V(b);
  }
  for (int j = 0; q.more(); ++j) {
    V *tile = s_p[j & 1][g];
    if (cur.valid) {
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        tile[tix_of_local(t + s * kG)] = pval(cur.oz[s], cur.op[s]);
        if (cur.hix[s] >= 0) tile[cur.hix[s]] = pval(cur.hz[s], cur.hp[s]);
      }
    }
    const bool valid = cur.valid;
    const int64_t leaf = cur.leaf;
    const uint8_t ol0 = cur.ol[0], ol1 = cur.ol[1];
    q.next();
    if (q.more()) load(q, cur);
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const uint8_t l = s ? ol1 : ol0;
        const int c = t + s * kG;
        const int64_t v = (leaf << 9) | c;
        const int ti = tix_of_local(c);
        const V pn = tile[ti];
        const V ap = is_fluid(l) ? stencil0<V, ORD>(tile, ti, l >> 2) : V(0);
        pout[v] = pn;
        Ap[v] = ap;
        acc += A(pn) * A(ap);
      }
    }
  }
  A vals[1] = {acc};
  publish<A, 1>(vals, part_a);
}

// B(k): ComputeLocal2 (solver.cc:323-346): alpha = rz_k / pAp, x += alpha p,
// r -= alpha Ap, z = M^-1 r; publishes (r.r, r.z).
template <typename V, typename A, bool ORD>
__global__ __launch_bounds__(kT) void k_pass_b(const int32_t *flist, int64_t nf,
                                               const uint8_t *lab, const V *p, const V *Ap, V *x,
                                               V *r, V *zv, const A *part_a, A *part_b,
                                               int g_parts, int k, PcgState *st) {
#pragma clang fp contract(off)
  if (st->done) return;
  A acc[2] = {A(0), A(0)};
  // Software pipeline: round j+1's loads are issued before round j computes.
  struct Pre {
    V p[2], a[2], x[2], r[2];
    uint8_t l[2];
    int64_t base;
    bool valid;
  };
  auto load = [&](const Walk &q, Pre &c) {
    c.valid = q.valid();
    if (!c.valid) return;
    c.base = int64_t(flist[q.mine()]) << 9;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int64_t v = c.base | (gtid() + s * kG);
      c.l[s] = lab[v];
      c.p[s] = p[v];
      c.a[s] = Ap[v];
      c.x[s] = x[v];
      c.r[s] = r[v];
    }
  };
  // Issue the first leaf's loads, then reduce pAp under their latency.
  Walk q = walk(nf);
  Pre cur;
  cur.valid = false;
  if (q.more()) load(q, cur);
  const A pAp = A(total<A, 1>(part_a, g_parts, 0));
  const A rz = A(st->rzh[k & 1]);
  const A alpha_a = rz / pAp;
  if (!(double(alpha_a) <= 1e5)) {  // CHECK_LE(alpha, LARGE_FLOAT), solver.cc:333
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->status = FSL_ERR_ALPHA;
      st->done = 1;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pAp = double(pAp);
    st->alpha = double(alpha_a);
  }
  const V alpha = V(alpha_a);
  while (q.more()) {
    const Pre c = cur;
    q.next();
    cur.valid = false;
    if (q.more()) load(q, cur);
    if (!c.valid) continue;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (!is_fluid(c.l[s])) continue;
      const int64_t v = c.base | (gtid() + s * kG);
      V xn, rn;
      if (ORD) {
        const V m = alpha * c.p[s];
        xn = m + c.x[s];
        const V m2 = (-alpha) * c.a[s];
        rn = m2 + c.r[s];
      } else {
        xn = fmav(alpha, c.p[s], c.x[s]);
        rn = fmav(-alpha, c.a[s], c.r[s]);
      }
      x[v] = xn;
      r[v] = rn;
      const V z = rn * dinv<V>(c.l[s] >> 2);
      zv[v] = z;
      acc[0] += A(rn) * A(rn);
      acc[1] += A(rn) * A(z);
    }
  }
  publish<A, 2>(acc, part_b);
}

// Block IC(0) (ic.hip): after the sweeps wrote z = M^-1 r, (r.r, r.z) are
// published as B(k)'s partials, in B's grid and summation form.
template <typename V, typename A>
__global__ __launch_bounds__(kT) void k_ic_dots(const int32_t *flist, int64_t nf,
                                                const uint8_t *lab, const V *r, const V *zv,
                                                A *part_b, const PcgState *st, int force) {
#pragma clang fp contract(off)
  if ((!force && st->done) || st->status) return;
  A acc[2] = {A(0), A(0)};
  for (Walk q = walk(nf); q.more(); q.next()) {
    if (!q.valid()) continue;
    const int64_t base = int64_t(flist[q.mine()]) << 9;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int64_t v = base | (gtid() + s * kG);
      if (!is_fluid(lab[v])) continue;
      const V rn = r[v], z = zv[v];
      acc[0] += A(rn) * A(rn);
      acc[1] += A(rn) * A(z);
    }
  }
  publish<A, 2>(acc, part_b);
}

// MIXED refresh, part 1: fold the float iterate into the double solution.
// force = 2: the fold of a convergence check enqueued without a host read,
// done only when the solve has stopped.
__global__ __launch_bounds__(kT) void k_refresh_x(const int32_t *flist, int64_t nf,
                                                  const uint8_t *lab, float *x, double *x64,
                                                  const PcgState *st, int force) {
  if ((!force && st->done) || st->status || (force == 2 && !st->done)) return;
  for (Walk q = walk(nf); q.more(); q.next()) {
    if (!q.valid()) continue;
    const int64_t base = int64_t(flist[q.mine()]) << 9;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int64_t v = base | (gtid() + s * kG);
      if (!is_fluid(lab[v])) continue;
      x64[v] += double(x[v]);
      x[v] = 0.0f;
    }
  }
}

// MIXED refresh, part 2: r = b - A x64 in double, stored as float; the
// partials (r.r, r.z) replace those of the last B pass.
__global__ __launch_bounds__(kT) void k_refresh_r(Dom d, const int32_t *flist, int64_t nf,
                                                  const uint8_t *lab, const float *div,
                                                  const double *x64, float *r, float *zv,
                                                  double *part_b, const PcgState *st, int force) {
  __shared__ double s_x[kGroups][1000];
  __shared__ uint8_t s_f[kGroups][1000];
  if ((!force && st->done) || st->status) return;
  const int g = threadIdx.x / kG;
  double acc[2] = {0.0, 0.0};
  for (Walk q = walk(nf); q.more(); q.next()) {
    const bool valid = q.valid();
    const int64_t leaf = valid ? flist[q.mine()] : 0;
    int o[3];
    leaf_origin(d, leaf, o);
    uint8_t ol[2] = {0, 0};
    if (valid)
      fill_tile<double>(d, lab, leaf, o, s_x[g], s_f[g], ol,
                        [&](int64_t v, uint8_t) { return x64[v]; });
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int c = gtid() + s * kG;
      const int64_t v = (leaf << 9) | c;
      if (!valid || !is_fluid(ol[s])) continue;
      const double rt =
          double(div[v]) - stencil<double, false>(s_x[g], s_f[g], tix_of_local(c), ol[s] >> 2);
      const float rf = float(rt);
      const float zf = rf * dinv<float>(ol[s] >> 2);
      r[v] = rf;
      if (zv) zv[v] = zf;
      acc[0] += double(rf) * double(rf);
      acc[1] += double(rf) * double(zf);
    }
    __syncthreads();
  }
  if (part_b) publish<double, 2>(acc, part_b);
}

// ------------------------------------------------------------------------
// Single-reduction iteration. One kernel per iteration k, given r_k, p_{k-1},
// x_k and the totals (gamma_k, delta_k, r.r_k) published by the previous
// kernel, with gamma = r.u, delta = u.Au, u = M^-1 r:
//   beta_k  = gamma_k / gamma_{k-1}
//   alpha_k = gamma_k / (delta_k - beta_k gamma_k / alpha_{k-1})
//   p_k = u_k + beta_k p_{k-1};  x_{k+1} = x_k + alpha_k p_k;
//   r_{k+1} = r_k - alpha_k A p_k;  u_{k+1} = M^-1 r_{k+1};
//   publish gamma_{k+1}, delta_{k+1} = u_{k+1}.A u_{k+1}, r.r_{k+1}.
// These are the iterates of the two-pass PCG (solver.cc:296-366) in exact
// arithmetic: p.Ap = delta_k - beta_k gamma_k / alpha_{k-1} follows from the
// M-orthogonality of the residuals (Chronopoulos and Gear, J. Comput. Appl.
// Math. 25, 1989). A p_k is computed directly (no recurrence for A p), so the
// residual recurrence is the two-pass one.
//
// delta_{k+1} is taken in edge form: sum over rows c of diag_c u_c^2 minus
// 2 u_c u_n over the faces (c, n) between rows, each face counted at its
// lower cell, so a leaf needs u_{k+1} on itself and on its three upper face
// layers, hence r_{k+1} and A p_k there, hence p_k on the leaf, its six face
// layers at distance 1, the three upper layers at distance 2 and nine edge
// rows: 1160 cells, staged in an LDS tile. r_k and p_{k-1} are read on those
// cells and p_k recomputed (halo cells are recomputed, not exchanged). r and
// p are double-buffered by iteration parity, as other workgroups read the
// halo values of the iteration's inputs.

// Workgroups of k_cg per CU. Each launch begins with every workgroup reading
// all partials of the previous launch and ends with a workgroup reduction
// and a store, so these per-launch costs grow with the workgroup count;
// four per CU still keep the leaf rounds fed.
constexpr int kCgMaxPerCu = 4;
// Rounds of partials each thread of k_cg loads at once: with 256 CUs and one
// 256-thread workgroup per leaf group, at most kCgMaxPerCu rounds.
constexpr int kCgPartRounds = kCgMaxPerCu;

// Tile P over [-2, 10)^3 of the leaf and tile U over [-1, 9)^3. Every stride
// is even and a cell at even z sits at an even index, so the z-pairs a
// thread owns (and the halo pairs, which start at even z) are aligned: each
// moves as one 8-byte (FP64: 16-byte) LDS access. The x strides (174 and 142
// words, against the 144 and 120 of unpadded rows) spread the pair accesses
// of a wavefront over the banks.
constexpr int kPX = 174, kPY = 12, kPSize = 11 * kPX + 11 * kPY + 12;
constexpr int kUX = 142, kUY = 12, kUSize = 9 * kUX + 9 * kUY + 12;
static_assert(kPX % 2 == 0 && kPY % 2 == 0 && kUX % 2 == 0 && kUY % 2 == 0 && kPSize % 2 == 0 &&
                  kUSize % 2 == 0,
              "aligned z-pairs");
__device__ inline int t12(int x, int y, int z) { return (x + 2) * kPX + (y + 2) * kPY + (z + 2); }
__device__ inline int t12_of_local(int c) { return t12(c >> 6, (c >> 3) & 7, c & 7); }
__device__ inline int u10(int x, int y, int z) { return (x + 1) * kUX + (y + 1) * kUY + (z + 2); }
__device__ inline int u10_of_local(int c) { return u10(c >> 6, (c >> 3) & 7, c & 7); }

// Elements i and i + 1 of an LDS tile (i even, so the pair is aligned) as
// one access.
template <typename V>
__device__ inline void lds_get2(const V *p, int i, V &a, V &b) {
  typedef V V2 __attribute__((ext_vector_type(2)));
  const V2 v = *reinterpret_cast<const V2 *>(p + i);
  a = v[0];
  b = v[1];
}
template <typename V>
__device__ inline void lds_put2(V *p, int i, V a, V b) {
  typedef V V2 __attribute__((ext_vector_type(2)));
  V2 v;
  v[0] = a;
  v[1] = b;
  *reinterpret_cast<V2 *>(p + i) = v;
}

// A p on a thread's own z-pair, cells c and c + 1 of the P tile (c even),
// with the additions of stencil_s in its order: the x and y neighbours as
// pair reads, the z neighbours outside the pair as single reads, and the
// pair itself from pc (the values the tile holds there).
template <typename V, bool ORD>
__device__ inline void stencil_pair(const V *sp, int c, const V pc[2], const int diag[2],
                                    V out[2]) {
#pragma clang fp contract(off)
  V xm[2], xp[2], ym[2], yp[2];
  lds_get2(sp, c - kPX, xm[0], xm[1]);
  lds_get2(sp, c + kPX, xp[0], xp[1]);
  lds_get2(sp, c - kPY, ym[0], ym[1]);
  lds_get2(sp, c + kPY, yp[0], yp[1]);
  const V zm[2] = {sp[c - 1], pc[0]}, zp[2] = {pc[1], sp[c + 2]};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (ORD) {
      V s = V(0);
      s = s + (-xm[i]);
      s = s + (-ym[i]);
      s = s + (-zm[i]);
      { const V m = V(diag[i]) * pc[i]; s = s + m; }
      s = s + (-zp[i]);
      s = s + (-yp[i]);
      s = s + (-xp[i]);
      out[i] = s;
    } else {
      const V nb = ((xm[i] + xp[i]) + (ym[i] + yp[i])) + (zm[i] + zp[i]);
      out[i] = fmav(V(diag[i]), pc[i], -nb);
    }
  }
}

// Halo cell m of the fused pass, relative to the leaf origin. m in [0, 384):
// face layers at distance 1 (face m/64: x-, x+, y-, y+, z-, z+); m in
// [384, 768): face layers at distance 2; m in [768, 864): edge rows at
// distance (1, 1) (axis pairs xy, xz, yz, each with signs --, -+, +-, ++;
// the free axis runs over 0..7).
__device__ inline void fhalo_cell(int m, int &x, int &y, int &z) {
  if (m < 768) {
    const int near = m < 384;
    const int h = near ? m : m - 384;
    const int f = h >> 6, a = (h >> 3) & 7, b = h & 7;
    const int s = (f & 1) ? (near ? 8 : 9) : (near ? -1 : -2);
    if (f < 2) { x = s; y = a; z = b; }
    else if (f < 4) { x = a; y = s; z = b; }
    else { x = a; y = b; z = s; }
  } else {
    const int e = (m - 768) >> 3, t = m & 7;
    const int pair = e >> 2, s1 = (e & 2) ? 8 : -1, s2 = (e & 1) ? 8 : -1;
    if (pair == 0) { x = s1; y = s2; z = t; }
    else if (pair == 1) { x = s1; y = t; z = s2; }
    else { x = t; y = s1; z = s2; }
  }
}
// Global voxel of the cell (x, y, z) (each in [-2, 10)) relative to `leaf`.
__device__ inline int64_t rel_voxel(const Dom &d, int64_t leaf, int x, int y, int z) {
  const int64_t l12 = int64_t(d.L[1]) * d.L[2];
  const int dx = x < 0 ? -1 : (x > 7 ? 1 : 0), dy = y < 0 ? -1 : (y > 7 ? 1 : 0),
            dz = z < 0 ? -1 : (z > 7 ? 1 : 0);
  const int64_t nb = leaf + dx * l12 + dy * int64_t(d.L[2]) + dz;
  return (nb << 9) | ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
}
// Halo pairs of the single-reduction pass. With the edge form of u.Au (each
// face of the fluid rows counted at its lower cell), a leaf needs u_{k+1} on
// itself and on its three upper face layers (x, y or z = 8; 192 cells), so
// A p_k there, so p_k on the leaf, its six face layers at distance 1, the
// upper face layers at distance 2 and nine edge rows: 648 halo cells. They
// are loaded as 380 pairs of z-neighbours (two values per lane, one vector
// access):
//   [0, 64)    x = -1 and x = 8 faces;     [64, 128)  y = -1 and y = 8 faces;
//   [128, 256) z faces, pairs z = (-2, -1) and (8, 9);
//   [256, 288) x = 9 face;                 [288, 320) y = 9 face;
//   [320, 332) xy edge rows (8, -1), (8, 8), (-1, 8), z pairs;
//   [332, 356) xz edge rows (x, z) = (8, -1), (8, 8), (-1, 8), loaded with
//              the cell beyond the edge in z (pairs (-2, -1) or (8, 9));
//   [356, 380) yz edge rows (y, z), likewise.
// Pair slot 0 of thread t is pair t; slot 1 is pair 256 + wave * 31 + lane
// on lanes 0-30; idle lanes load pair 0 and discard it.
struct Run {
  int x, y, z;  // first cell of the pair, relative to the leaf origin
};
__device__ inline Run halo_pair(int u) {
  if (u < 128) {
    const int i = u & 31, c = (u >> 5) & 1 ? 8 : -1;
    return u < 64 ? Run{c, i >> 2, (i & 3) * 2} : Run{i >> 2, c, (i & 3) * 2};
  }
  if (u < 256) {
    const int v = u - 128, i = v & 63;
    return Run{i >> 3, i & 7, v >> 6 ? 8 : -2};
  }
  if (u < 320) {
    const int v = u - 256, i = v & 31;
    return v < 32 ? Run{9, i >> 2, (i & 3) * 2} : Run{i >> 2, 9, (i & 3) * 2};
  }
  // The three edge rows of a pair of axes: (8, -1), (8, 8), (-1, 8).
  if (u < 332) {
    const int v = u - 320, e = v >> 2;
    return Run{e < 2 ? 8 : -1, e == 0 ? -1 : 8, (v & 3) * 2};
  }
  const int v = u < 356 ? u - 332 : u - 356, e = v >> 3, f = v & 7;
  const int a = e < 2 ? 8 : -1, b = e == 0 ? -2 : 8;
  return u < 356 ? Run{a, f, b} : Run{f, a, b};
}
__device__ inline int pair_of(int t, int s) {
  const int w = t >> 6, l = t & 63;
  return s == 0 ? t : (l < 31 ? 256 + w * 31 + l : -1);
}
// Upper distance-1 cell m in [0, 192): face m / 64 (x = 8, y = 8, z = 8).
__device__ inline void r1_cell(int m, int &x, int &y, int &z) {
  const int f = m >> 6, a = (m >> 3) & 7, b = m & 7;
  if (f == 0) { x = 8; y = a; z = b; }
  else if (f == 1) { x = a; y = 8; z = b; }
  else { x = a; y = b; z = 8; }
}
// Upper distance-1 cells of a pair: bit e of the mask is set when cell e of
// the pair is cell base + e of r1_cell.
__device__ inline void pair_r1(int u, const Run &r, int &mask, int &base) {
  mask = 0;
  base = 0;
  if (u >= 32 && u < 64) { mask = 3; base = r.y * 8 + r.z; }
  else if (u >= 96 && u < 128) { mask = 3; base = 64 + r.x * 8 + r.z; }
  else if (u >= 192 && u < 256) { mask = 1; base = 128 + r.x * 8 + r.y; }
}

// N consecutive elements from element i (i aligned to N), one vector access.
template <int N, typename T, typename Off>
__device__ inline void load_run(const T *p, Off i, T *out);
// N consecutive labels from voxel i, packed little-endian.
template <int N, typename Off>
__device__ inline uint32_t load_labels(const uint8_t *lab, Off i);

// Zero-padded stencil with tile strides SX, SY, 1.
template <typename V, bool ORD, int SX, int SY>
__device__ inline V stencil_s(const V *sp, int c, int diag) {
#pragma clang fp contract(off)
  if (ORD) {
    V s = V(0);
    s = s + (-sp[c - SX]);
    s = s + (-sp[c - SY]);
    s = s + (-sp[c - 1]);
    { const V m = V(diag) * sp[c]; s = s + m; }
    s = s + (-sp[c + 1]);
    s = s + (-sp[c + SY]);
    s = s + (-sp[c + SX]);
    return s;
  } else {
    const V nb = ((sp[c - SX] + sp[c + SX]) + (sp[c - SY] + sp[c + SY])) + (sp[c - 1] + sp[c + 1]);
    return fmav(V(diag), sp[c], -nb);
  }
}

// GR leaf groups per workgroup: the groups of a workgroup share its
// barriers, so small workgroups (GR = 1) let the leaves of one CU proceed
// independently and overlap each other's memory latency.
// Element i of p. With 32-bit offsets the byte offset is formed in 32 bits,
// so the access is a scalar base plus a 32-bit vector offset (no 64-bit
// address arithmetic per access); the caller guarantees it fits.
template <typename T>
__device__ inline T &elem(T *p, uint32_t i) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return *reinterpret_cast<T *>(reinterpret_cast<B *>(p) + uint32_t(i * uint32_t(sizeof(T))));
}
template <typename T>
__device__ inline T &elem(T *p, int64_t i) { return p[i]; }

template <int N, typename T, typename Off>
__device__ inline void load_run(const T *p, Off i, T *out) {
  if constexpr (N == 1) {
    out[0] = elem(p, i);
  } else {
    typedef T VT __attribute__((ext_vector_type(N)));
    const VT v = *reinterpret_cast<const VT *>(&elem(p, i));
#pragma unroll
    for (int e = 0; e < N; ++e) out[e] = v[e];
  }
}
template <int N, typename Off>
__device__ inline uint32_t load_labels(const uint8_t *lab, Off i) {
  if constexpr (N == 4) return *reinterpret_cast<const uint32_t *>(&elem(lab, i));
  else if constexpr (N == 2) return *reinterpret_cast<const uint16_t *>(&elem(lab, i));
  else return elem(lab, i);
}
// The same pair as one agent-scope relaxed atomic store, which the hardware
// writes through the L2 (sc1): k_cg stores r, p and x this way, so a launch
// ends with none of their lines dirty in the L2 (at 256^3 0.8 us per launch
// faster than write-back stores; at 512^3 1% slower).
template <typename T, typename Off>
__device__ inline void store_pair_wt(T *p, Off i, T a, T b) {
  static_assert(sizeof(T) == 4, "float pairs");
  union { T f[2]; unsigned long long u; } v;
  v.f[0] = a; v.f[1] = b;
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(&elem(p, i)), v.u, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T, typename Off>
__device__ inline void store_pair(T *p, Off i, T a, T b) {
  typedef T VT __attribute__((ext_vector_type(2)));
  VT v;
  v[0] = a;
  v[1] = b;
  *reinterpret_cast<VT *>(&elem(p, i)) = v;
}

// OFF32: every array the kernel indexes is under 4 GiB, so voxel offsets
// are 32-bit and each access is a scalar base plus a vector offset.
//
// Per leaf group (256 threads): thread t owns cells 2t and 2t+1 (pairs along
// z, loaded and stored as vectors) and up to two halo pairs. Phase 1 forms
// p_k on the leaf and its halo in the P tile and hands the r and labels of
// the upper distance-1 cells to a balanced mapping (48 per wave) through
// LDS; phase 2 applies A to p_k there and on the own cells and forms
// u_{k+1}; phase 3 accumulates gamma, delta (edge form) and r.r. At most
// 128 VGPRs, so that four workgroups fit on a CU (kCgMaxPerCu) in FP64 too.
template <typename V, typename A, bool ORD, int GR, bool OFF32>
__global__ __launch_bounds__(kG * GR) __attribute__((amdgpu_waves_per_eu(kCgMaxPerCu * GR)))
void k_cg(Dom d, const int32_t *flist, int64_t nf,
                                                const uint8_t *lab, const V *rin, V *rout,
                                                const V *pin, V *pout, V *x, double *x64,
                                                const A *part_in, A *part_out, int g_parts,
                                                int k, int mode, PcgState *st) {
#pragma clang fp contract(off)
  using Off = typename std::conditional<OFF32, uint32_t, int64_t>::type;
  __shared__ __align__(16) V s_p[GR][kPSize];
  __shared__ __align__(16) V s_u[GR][kUSize];
  __shared__ V s_r1[GR][192];        // r of the upper distance-1 cells
  __shared__ uint8_t s_l1[GR][192];  // their labels
  // Leaf schedule: entry j is the list entry of round j of this group
  // (the first list entry for rounds past the end, which are discarded).
  constexpr int kSched = kG;
  __shared__ int32_t s_leaf[GR][kSched];
  // 1 / diag by label (label = FLUID | diag << 2; 0 for every non-row label).
  __shared__ V s_dinv[32];
  const int g = threadIdx.x / kG, t = gtid();
  // First the loads nothing else waits for: this group's leaf schedule and
  // the partials of the previous launch, all in flight together.
  const Walk q0 = walk<kG * GR>(nf);
  const int64_t pos0 = q0.mine() + int64_t(t) * q0.ng * q0.step;
  const int32_t sched0 = flist[pos0 < q0.end ? pos0 : 0];
  const int32_t leaf0 = flist[q0.mine() < q0.end ? q0.mine() : 0];  // round 0's leaf
  PartRegs<A, 3, kCgPartRounds> pr;
  parts_issue<A, 3, kG * GR, kCgPartRounds>(part_in, g_parts, pr);
  // State of iteration k - 1 (slots (k - 1) & 1 = (k + 1) & 1, which this
  // launch does not write) and the stop test's input, read before anything
  // is stored.
  const double rzh_prev = st->rzh[(k + 1) & 1], alph_prev = st->alph[(k + 1) & 1];
  const int stop_k = st->stop_k, stop_mode = st->stop_mode, max_iter = st->max_iter;
  const double thr = st->thr;
  // Keeps the loads above ahead of the exit test below (the compiler would
  // otherwise move them after it, into the path that uses them).
  asm volatile("" ::: "memory");
  // Exit only on a stop stored by an earlier launch: workgroup 0 of this
  // launch may store its own stop before other workgroups start, and those
  // must still reach the same decision and apply the pending x update.
  if (stop_k < k) return;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6), l = t & 63;
  if (threadIdx.x < 32) {
    const int lb = threadIdx.x;
    s_dinv[lb] = is_fluid(uint8_t(lb)) ? dinv<V>(lb >> 2) : V(0);
  }
  // Per-thread constants: the two halo runs (voxel offset from the leaf's
  // first voxel, P tile position, distance-1 cells), the two distance-1
  // cells of phase 2 (tile positions) and the own pair (tile positions).
  const int64_t l12 = int64_t(d.L[1]) * d.L[2];
  int32_t hoff[2];
  int hp0[2], rmask[2], rbase[2];
  bool hon[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int u = pair_of(t, s);
    hon[s] = u >= 0;
    const Run r = halo_pair(hon[s] ? u : 0);
    const int dx = r.x < 0 ? -1 : (r.x > 7 ? 1 : 0), dy = r.y < 0 ? -1 : (r.y > 7 ? 1 : 0),
              dz = r.z < 0 ? -1 : (r.z > 7 ? 1 : 0);
    hoff[s] = int32_t((dx * l12 + dy * int64_t(d.L[2]) + dz) * kLeaf) |
              ((r.x & 7) << 6) | ((r.y & 7) << 3) | (r.z & 7);
    hp0[s] = t12(r.x, r.y, r.z);
    pair_r1(hon[s] ? u : 512, r, rmask[s], rbase[s]);
  }
  // Upper distance-1 cell of phase 2: 48 per wave.
  const bool m1ok = l < 48;
  const int m1 = w * 48 + (m1ok ? l : 0);
  int r1p, r1u;
  {
    int hx, hy, hz;
    r1_cell(m1, hx, hy, hz);
    r1p = t12(hx, hy, hz);
    r1u = u10(hx, hy, hz);
  }
  int op12 = t12_of_local(2 * t), ou10 = u10_of_local(2 * t);  // cell 2t+1 is +1

  struct Pre {
    V orr[2], op[2], ox[2], hr[2][2], hp[2][2];
    uint16_t ol, hl[2];
    Off base;
    bool valid;
  };
  // Prefetch of one leaf. All loads of a wave are unconditional (a round
  // past the end of the list re-reads the list's first leaf, at k = 0 the
  // unused p buffer is read, idle lanes repeat a run; such values are
  // discarded), so the compiler issues them together.
  auto load = [&](bool valid, int32_t leaf, Pre &c) {
    c.valid = valid;
    c.base = Off(leaf) << 9;
    const Off v = c.base + Off(2 * t);
    c.ol = uint16_t(load_labels<2>(lab, v));
    load_run<2>(rin, v, c.orr);
    load_run<2>(pin, v, c.op);
    if (mode & 1) load_run<2>(x, v, c.ox);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const Off hv = c.base + Off(hoff[s]);
      c.hl[s] = uint16_t(load_labels<2>(lab, hv));
      load_run<2>(rin, hv, c.hr[s]);
      load_run<2>(pin, hv, c.hp[s]);
    }
  };
  Walk q = q0;
  auto refill = [&](int j0) {
    const int64_t pos = q0.mine() + int64_t(j0 + t) * q0.ng * q0.step;
    s_leaf[g][t] = flist[pos < q0.end ? pos : 0];
  };
  int j = 0;  // round of q
  Pre cur;
  // Issued also when the group has no leaf (leaf0 is then a valid list entry
  // and the values are discarded): with no branch around these loads, the
  // wait for the partials below does not wait for them.
  load(q.more() && q.valid(), leaf0, cur);
  // The schedule is read from round 1 on, after the barrier in parts_total.
  s_leaf[g][t] = sched0;
  // The per-thread LDS positions are formed here, under the latency of the
  // partials: otherwise the compiler forms them after the totals, on the way
  // to the first leaf's arithmetic (0.4 us per FP64 launch at 256^3).
  asm volatile("" : "+v"(op12), "+v"(ou10), "+v"(r1p), "+v"(r1u), "+v"(hp0[0]), "+v"(hp0[1]),
               "+v"(rmask[0]), "+v"(rmask[1]), "+v"(rbase[0]), "+v"(rbase[1]));
  // Scalars of iteration k, reduced under the latency of the first leaf's loads.
  double tot[3];
  parts_total<A, 3, kG * GR, kCgPartRounds>(part_in, g_parts, pr, tot);
  const A gam = A(tot[0]), del = A(tot[1]);
  const double rr = tot[2];
  A beta = A(0), pAp = del;
  if (k > 0) {
    beta = gam / A(rzh_prev);
    const A m = beta * gam;
    pAp = del - m / A(alph_prev);
  }
  const A alpha_a = gam / pAp;
  int status = 0;
  if (isnan(rr)) status = FSL_ERR_NAN;
  else if (!(double(beta) <= 1e5)) status = FSL_ERR_BETA;  // solver.cc:361
  bool stop = status != 0 || converged(stop_mode, thr, rr) || k >= max_iter;
  if (!stop && !(double(pAp) > 0.0 && double(alpha_a) <= 1e5)) {  // solver.cc:333
    status = FSL_ERR_ALPHA;
    stop = true;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->it = k;
    st->rr = rr;
    st->rz = double(gam);
    st->beta = double(beta);
    st->pAp = double(pAp);
    st->alpha = double(alpha_a);
    st->rzh[k & 1] = double(gam);
    st->alph[k & 1] = double(alpha_a);
    if (status) st->status = status;
    if (stop) {
      st->stop_k = k;
      st->done = 1;
    }
  }
  // alpha_{k-1}: x lacks alpha_{k-1} p_{k-1} when bit 1 of mode is set.
  const V alpha_prev = V(alph_prev);
  if (stop) {
    if (status == 0 && (mode & 2)) {
      // Stopping at x_k: apply the pending alpha_{k-1} p_{k-1} (p_{k-1} is
      // the input direction).
      for (Walk w2 = q0; w2.more(); w2.next()) {
        if (!w2.valid()) continue;
        const Off v = (Off(flist[w2.mine()]) << 9) + Off(2 * t);
        const uint32_t lb = load_labels<2>(lab, v);
        V pp[2], xx[2];
        load_run<2>(pin, v, pp);
        load_run<2>(x, v, xx);
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          if (!is_fluid(uint8_t(lb >> (8 * i)))) continue;
          if (ORD) {
            const V m = alpha_prev * pp[i];
            xx[i] = m + xx[i];
          } else {
            xx[i] = fmav(alpha_prev, pp[i], xx[i]);
          }
        }
        store_pair(x, v, xx[0], xx[1]);
      }
    }
    return;
  }
  const V alpha = V(alpha_a), bv = V(beta);
  // p = u + beta p_old with u = dinv * r; the select drops whatever a
  // non-row cell holds.
  auto pnew = [&](V r, V p, uint32_t lb, V di) {
    const uint8_t lc = uint8_t(lb);
    const V u = r * di;
    const V m = bv * p;
    const V pv = k == 0 ? u : u + m;
    return is_fluid(lc) ? pv : V(0);
  };
  A acc[3] = {A(0), A(0), A(0)};
  V *tp = s_p[g];
  V *tu = s_u[g];
  V *tr1 = s_r1[g];
  uint8_t *tl1 = s_l1[g];
  while (q.more()) {
    // Phase 1: p_k on the 1160 cells into the P tile; upper distance-1 cells' r
    // and labels to LDS.
    const bool valid = cur.valid;
    const Off base = cur.base;
    V wr[2], wp[2], wx[2], wd[2], wo[2];
    uint8_t wl[2];
    // The 1 / diag of the own and halo cells, read from the table together.
    V hd[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      wl[i] = uint8_t(cur.ol >> (8 * i));
      wd[i] = s_dinv[wl[i] & 31];
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int e = 0; e < 2; ++e) hd[s][e] = s_dinv[(cur.hl[s] >> (8 * e)) & 31];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      wp[i] = pnew(cur.orr[i], cur.op[i], wl[i], wd[i]);
      wr[i] = cur.orr[i];
      wx[i] = cur.ox[i];
      wo[i] = cur.op[i];
    }
    if (valid) {
      lds_put2(tp, op12, wp[0], wp[1]);
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        if (!hon[s]) continue;
        V hv[2];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const uint32_t lb = cur.hl[s] >> (8 * e);
          hv[e] = pnew(cur.hr[s][e], cur.hp[s][e], lb, hd[s][e]);
          if ((rmask[s] >> e) & 1) {
            tr1[rbase[s] + e] = cur.hr[s][e];
            tl1[rbase[s] + e] = uint8_t(lb);
          }
        }
        lds_put2(tp, hp0[s], hv[0], hv[1]);
      }
    }
    q.next();
    ++j;
    cur.valid = false;
    if (q.more()) {
      if (j % kSched == 0) {
        __syncthreads();
        refill(j);
        __syncthreads();
      }
      load(q.valid(), s_leaf[g][j % kSched], cur);
    }
    __syncthreads();
    // Phase 2: A p_k on the own pair and the upper distance-1 cells; r_{k+1} and
    // u_{k+1} there; x, r and p of the own pair are written.
    V rn[2], un[2];
    if (valid) {
      V xn[2], aps[2];
      {
        const int dg[2] = {wl[0] >> 2, wl[1] >> 2};
        stencil_pair<V, ORD>(tp, op12, wp, dg, aps);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        rn[i] = V(0);
        un[i] = V(0);
        xn[i] = wx[i];
        if (is_fluid(wl[i])) {
          const V ap = aps[i];
          // x: alpha_{k-1} p_{k-1} (when pending) then alpha_k p_k, the two
          // axpys of consecutive iterations in order.
          if (ORD) {
            if (mode & 2) {
              const V m0 = alpha_prev * wo[i];
              xn[i] = m0 + xn[i];
            }
            const V m = alpha * wp[i];
            xn[i] = m + xn[i];
            const V m2 = (-alpha) * ap;
            rn[i] = m2 + wr[i];
          } else {
            if (mode & 2) xn[i] = fmav(alpha_prev, wo[i], xn[i]);
            xn[i] = fmav(alpha, wp[i], xn[i]);
            rn[i] = fmav(-alpha, ap, wr[i]);
          }
          un[i] = rn[i] * wd[i];
        }
      }
      lds_put2(tu, ou10, un[0], un[1]);
      const Off v = base + Off(2 * t);
      if (mode & 4) {
        // Fold before a residual refresh: x64 += x_{k+1}, x = 0 (x and x64
        // are read only at a thread's own cells, so no other workgroup sees
        // the intermediate state).
        typedef double D2 __attribute__((ext_vector_type(2)));
        D2 *xp = reinterpret_cast<D2 *>(&elem(x64, v));
        D2 acc64 = *xp;
        if (is_fluid(wl[0])) acc64[0] += double(xn[0]);
        if (is_fluid(wl[1])) acc64[1] += double(xn[1]);
        *xp = acc64;
        store_pair(x, v, V(0), V(0));
      } else if (mode & 1) {
        if constexpr (sizeof(V) == 4) store_pair_wt(x, v, xn[0], xn[1]);
        else store_pair(x, v, xn[0], xn[1]);
      }
      if constexpr (sizeof(V) == 4) {
        store_pair_wt(rout, v, rn[0], rn[1]);
        store_pair_wt(pout, v, wp[0], wp[1]);
      } else {
        store_pair(rout, v, rn[0], rn[1]);
        store_pair(pout, v, wp[0], wp[1]);
      }
      if (m1ok) {
        const uint8_t lb = tl1[m1];
        const V di = s_dinv[lb & 31];
        // A non-row cell has dinv 0: its u is 0 whatever its stencil.
        const V ap = stencil_s<V, ORD, kPX, kPY>(tp, r1p, lb >> 2);
        const V rh0 = tr1[m1];
        V rh;
        if (ORD) {
          const V m2 = (-alpha) * ap;
          rh = m2 + rh0;
        } else {
          rh = fmav(-alpha, ap, rh0);
        }
        tu[r1u] = di == V(0) ? V(0) : rh * di;
      }
    }
    __syncthreads();
    // Phase 3: gamma, delta and r.r of iteration k+1 on the own pair, with
    // u.Au in edge form: diag u_c^2 - 2 u_c (u_{c+x} + u_{c+y} + u_{c+z})
    // counts every face between two rows once, at its lower cell (A is
    // symmetric; u is 0 off the rows). FP64 accumulation, one rounding per
    // term.
    if (valid) {
      // The +x and +y neighbours as pair reads; the +z neighbour of cell 2t
      // is cell 2t+1 (un[1], the value the tile holds there).
      V ux[2], uy[2];
      lds_get2(tu, ou10 + kUX, ux[0], ux[1]);
      lds_get2(tu, ou10 + kUY, uy[0], uy[1]);
      const V uz[2] = {un[1], tu[ou10 + 2]};
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        if (!is_fluid(wl[i])) continue;
        const A uc = A(un[i]);
        const A sn = (A(ux[i]) + A(uy[i])) + A(uz[i]);
        const A tt = fmav(A(wl[i] >> 2), uc, A(-2) * sn);
        acc[0] = fmav(A(rn[i]), uc, acc[0]);
        acc[1] = fmav(uc, tt, acc[1]);
        acc[2] = fmav(A(rn[i]), A(rn[i]), acc[2]);
      }
    }
  }
  publish<A, 3, kG * GR>(acc, part_out);
}

// MIXED residual refresh (single-reduction path): r = b - A x64 in FP64 on
// the leaf and its upper distance-1 cells from x64 staged on the halo of
// k_cg (own z-pairs, 380 halo pairs, upper distance-1 cells through LDS),
// stored as float on the leaf; then u = M^-1 r and the partials (gamma,
// delta in edge form, r.r) of the refreshed residual. x has been
// folded into x64 before (by the k_cg launch before the refresh, or by
// k_refresh_x).
template <bool OFF32>
// force = 2: a convergence check enqueued without knowing whether or where
// the solve stopped. It records kn = it (stopped) or it + 1 (still running)
// as refresh_kn and, when stopped, refreshes iteration kn, whose residual
// buffer and partials are rout/part_out for even kn and rout_odd/part_odd
// for odd kn; a running solve is left as it is.
__global__ __launch_bounds__(kG) void k_refresh_cg(Dom d, const int32_t *flist, int64_t nf,
                                                   const uint8_t *lab, const float *div,
                                                   const double *x64, float *rout,
                                                   double *part_out, float *rout_odd,
                                                   double *part_odd, PcgState *st, int force) {
#pragma clang fp contract(off)
  using Off = typename std::conditional<OFF32, uint32_t, int64_t>::type;
  __shared__ __align__(16) double s_p[kPSize];
  __shared__ __align__(16) float s_u[kUSize];
  __shared__ float s_b1[192];
  __shared__ uint8_t s_l1[192];
  __shared__ float s_dinv[32];
  if ((!force && st->done) || st->status) return;
  if (force == 2) {
    const bool stopped = st->done != 0;
    const int kn = stopped ? st->it : st->it + 1;
    // Every workgroup reads done and it before workgroup 0 stores
    // refresh_kn; nothing in this launch writes them.
    if (blockIdx.x == 0 && threadIdx.x == 0) st->refresh_kn = kn;
    if (!stopped) return;
    if (kn & 1) {
      rout = rout_odd;
      part_out = part_odd;
    }
  }
  const int t = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6), l = t & 63;
  if (t < 32) s_dinv[t] = is_fluid(uint8_t(t)) ? dinv<float>(t >> 2) : 0.0f;
  const int64_t l12 = int64_t(d.L[1]) * d.L[2];
  int32_t hoff[2];
  int hp0[2], rmask[2], rbase[2];
  bool hon[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int u = pair_of(t, s);
    hon[s] = u >= 0;
    const Run r = halo_pair(hon[s] ? u : 0);
    const int dx = r.x < 0 ? -1 : (r.x > 7 ? 1 : 0), dy = r.y < 0 ? -1 : (r.y > 7 ? 1 : 0),
              dz = r.z < 0 ? -1 : (r.z > 7 ? 1 : 0);
    hoff[s] = int32_t((dx * l12 + dy * int64_t(d.L[2]) + dz) * kLeaf) |
              ((r.x & 7) << 6) | ((r.y & 7) << 3) | (r.z & 7);
    hp0[s] = t12(r.x, r.y, r.z);
    pair_r1(hon[s] ? u : 512, r, rmask[s], rbase[s]);
  }
  // Upper distance-1 cell of phase 2: 48 per wave.
  const bool m1ok = l < 48;
  const int m1 = w * 48 + (m1ok ? l : 0);
  int r1p, r1u;
  {
    int hx, hy, hz;
    r1_cell(m1, hx, hy, hz);
    r1p = t12(hx, hy, hz);
    r1u = u10(hx, hy, hz);
  }
  const int op12 = t12_of_local(2 * t), ou10 = u10_of_local(2 * t);
  double acc[3] = {0.0, 0.0, 0.0};
  // One leaf's loads, issued a round ahead of its arithmetic.
  struct RPre {
    double ox[2], hx[2][2];
    float ob[2], hb[2][2];
    uint32_t ol, hl[2];
    Off base;
  };
  auto rload = [&](int32_t leaf, RPre &c) {
    c.base = Off(leaf) << 9;
    const Off v = c.base + Off(2 * t);
    c.ol = load_labels<2>(lab, v);
    load_run<2>(x64, v, c.ox);
    load_run<2>(div, v, c.ob);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const Off hv = c.base + Off(hoff[s]);
      c.hl[s] = load_labels<2>(lab, hv);
      load_run<2>(x64, hv, c.hx[s]);
      load_run<2>(div, hv, c.hb[s]);
    }
  };
  Walk q = walk<kG>(nf);
  RPre cur;
  if (q.more()) rload(flist[q.mine()], cur);
  while (q.more()) {
    // Every round has a valid leaf: one group per workgroup.
    const Off v = cur.base + Off(2 * t);
    // Phase 1: x64 on the 1160 cells (0 off the rows) into the tile, as pairs.
    uint8_t wl[2];
    float ob[2];
    double px[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      wl[i] = uint8_t(cur.ol >> (8 * i));
      ob[i] = cur.ob[i];
      px[i] = is_fluid(wl[i]) ? cur.ox[i] : 0.0;
    }
    lds_put2(s_p, op12, px[0], px[1]);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (!hon[s]) continue;
      double hv[2];
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const uint8_t lb = uint8_t(cur.hl[s] >> (8 * e));
        hv[e] = is_fluid(lb) ? cur.hx[s][e] : 0.0;
        if ((rmask[s] >> e) & 1) {
          s_b1[rbase[s] + e] = cur.hb[s][e];
          s_l1[rbase[s] + e] = lb;
        }
      }
      lds_put2(s_p, hp0[s], hv[0], hv[1]);
    }
    q.next();
    if (q.more()) rload(flist[q.mine()], cur);
    __syncthreads();
    // Phase 2: r = b - A x64 (FP64) on the own pair and the upper distance-1 cells.
    float rn[2], un[2];
    {
      double aps[2];
      const int dg[2] = {wl[0] >> 2, wl[1] >> 2};
      stencil_pair<double, false>(s_p, op12, px, dg, aps);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        rn[i] = 0.0f;
        if (is_fluid(wl[i])) rn[i] = float(double(ob[i]) - aps[i]);
        un[i] = rn[i] * s_dinv[wl[i] & 31];
      }
    }
    lds_put2(s_u, ou10, un[0], un[1]);
    store_pair(rout, v, rn[0], rn[1]);
    if (m1ok) {
      const uint8_t lb = s_l1[m1];
      const float di = s_dinv[lb & 31];
      float uh = 0.0f;
      if (di != 0.0f)
        uh = float(double(s_b1[m1]) - stencil_s<double, false, kPX, kPY>(s_p, r1p, lb >> 2)) * di;
      s_u[r1u] = uh;
    }
    __syncthreads();
    // Phase 3: gamma, delta (edge form, as in k_cg), r.r in FP64. The next
    // round's phase 1 writes only the P tile, which this phase does not
    // read, and its barrier comes before the next phase 2 writes the U tile.
    {
      float ux[2], uy[2];
      lds_get2(s_u, ou10 + kUX, ux[0], ux[1]);
      lds_get2(s_u, ou10 + kUY, uy[0], uy[1]);
      const float uz[2] = {un[1], s_u[ou10 + 2]};
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        if (!is_fluid(wl[i])) continue;
        const double uc = double(un[i]);
        const double sn = (double(ux[i]) + double(uy[i])) + double(uz[i]);
        const double tt = fmav(double(wl[i] >> 2), uc, -2.0 * sn);
        acc[0] = fmav(double(rn[i]), uc, acc[0]);
        acc[1] = fmav(uc, tt, acc[1]);
        acc[2] = fmav(double(rn[i]), double(rn[i]), acc[2]);
      }
    }
  }
  publish<double, 3>(acc, part_out);
}

// gamma = r.u, delta = u.Au and r.r for a residual r (u = M^-1 r), from a
// leaf tile with its face halos: the totals the single-reduction kernel
// starts from, after preparation and after each residual refresh.
template <typename V, typename A, bool ORD, int GR>
__global__ __launch_bounds__(kG * GR) void k_gamma_delta(Dom d, const int32_t *flist,
                                                         int64_t nf, const uint8_t *lab,
                                                         const V *r, A *part,
                                                         const PcgState *st, int force) {
#pragma clang fp contract(off)
  __shared__ V s_u[GR][1000];
  __shared__ uint8_t s_f[GR][1000];
  if ((!force && st->done) || st->status) return;
  const int g = threadIdx.x / kG;
  A acc[3] = {A(0), A(0), A(0)};
  for (Walk q = walk(nf); q.more(); q.next()) {
    const bool valid = q.valid();
    const int64_t leaf = valid ? flist[q.mine()] : 0;
    int o[3];
    leaf_origin(d, leaf, o);
    uint8_t ol[2] = {0, 0};
    if (valid)
      fill_tile<V>(d, lab, leaf, o, s_u[g], s_f[g], ol,
                   [&](int64_t v, uint8_t l) { return r[v] * dinv<V>(l >> 2); });
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int c = gtid() + s * kG;
      if (!valid || !is_fluid(ol[s])) continue;
      const int ti = tix_of_local(c);
      const V u = s_u[g][ti];
      const V rv = r[(leaf << 9) | c];
      const V au = stencil_s<V, ORD, 100, 10>(s_u[g], ti, ol[s] >> 2);
      acc[0] += A(rv) * A(u);
      acc[1] += A(u) * A(au);
      acc[2] += A(rv) * A(rv);
    }
    __syncthreads();
  }
  publish<A, 3>(acc, part);
}

// Scalars after preparation (single-reduction path).
template <typename A>
__global__ __launch_bounds__(1024) void k_prepare_scalars_sr(const A *part_f, int g_f,
                                                             const A *part_bb, int g_bb,
                                                             PcgState *st) {
  double tot[3];
  total_all<A, 3>(part_f, g_f, tot);
  const double bb = total<A, 1>(part_bb, g_bb, 0);
  if (threadIdx.x != 0) return;
  st->rr = tot[2];
  st->rz = tot[0];
  st->bnorm2 = bb;
  if (st->stop_mode == FSL_STOP_REL) st->thr = st->thr * sqrt(bb);
  st->it = 0;
  st->status = isnan(tot[2]) ? FSL_ERR_NAN : 0;
  st->done = (st->status != 0 || converged(st, tot[2]) || st->max_iter <= 0) ? 1 : 0;
}

// After a forced refresh (single-reduction path): the true r.r decides.
// part_odd set: the check of k_refresh_cg's force 2 (iteration refresh_kn,
// stored there; its partials are part_odd when odd), which also becomes the
// solve's iteration count.
__global__ __launch_bounds__(1024) void k_refresh_scalars_sr(const double *part_f, int g,
                                                             PcgState *st,
                                                             const double *part_odd) {
  // A check of a running solve refreshed nothing.
  if (part_odd && !st->done) return;
  const int kn = st->refresh_kn;
  if (part_odd && (kn & 1)) part_f = part_odd;
  double tot[3];
  total_all<double, 3>(part_f, g, tot);
  if (threadIdx.x != 0) return;
  if (part_odd) st->it = kn;
  st->rr_true = tot[2];
  st->rr = tot[2];
  if (isnan(tot[2])) st->status = FSL_ERR_NAN;
  st->done = (st->status != 0 || converged(st, tot[2]) || st->it >= st->max_iter) ? 1 : 0;
  // The solve resumes from iteration it: its k_cg launch must run.
  if (!st->done) st->stop_k = 0x7fffffff;
}

// Stage API: after B(k), the scalars and the stopping test of iteration k+1
// (the same computation A(k+1) starts with).
template <typename A>
__global__ __launch_bounds__(1024) void k_stage_scalars(const A *part_b_new, int g, int k,
                                                        PcgState *st) {
  const double rr = total<A, 2>(part_b_new, g, 0);
  const A rz = A(total<A, 2>(part_b_new, g, 1));
  if (threadIdx.x != 0 || st->done) return;
  const A rz_old = A(st->rzh[k & 1]);
  const A b = rz / rz_old;
  st->it = k + 1;
  st->rr = rr;
  st->rz = double(rz);
  st->rz_old = double(rz_old);
  st->beta = double(b);
  if (isnan(rr)) st->status = FSL_ERR_NAN;
  else if (!(double(b) <= 1e5)) st->status = FSL_ERR_BETA;
  st->done = (st->status != 0 || converged(st, rr) || k + 1 >= st->max_iter) ? 1 : 0;
}

// Stage API: pAp of A(k) (ComputeLocal1's output, solver.cc:304-308).
template <typename A>
__global__ __launch_bounds__(1024) void k_stage_pap(const A *part_a, int g, PcgState *st) {
  const double pAp = total<A, 1>(part_a, g, 0);
  if (threadIdx.x == 0 && !st->done) st->pAp = double(A(pAp));
}

// Dot products summed in the oracle's row order (FSL_DOT_ROW_ORDER): over
// the FLUID rows in lexicographic (i, j, k) order, the row index order of
// CreateIndexAndMaskTree (solver.cc:114-151), one rounding per product and
// per sum, s = s + a_i b_i, as a sequential loop. The row list (voxel of
// each row, in row order) is built once per solve: k_row_flag marks the
// rows over the cells in lexicographic order, a scan numbers them and
// k_row_list writes each row's voxel at its number.
__global__ void k_row_flag(Dom d, const uint8_t *lab, uint32_t *flag) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n12 = int64_t(d.n[1]) * d.n[2];
  if (q >= int64_t(d.n[0]) * n12) return;
  const int i = int(q / n12), j = int((q / d.n[2]) % d.n[1]), k = int(q % d.n[2]);
  flag[q] = is_fluid(lab[vox_of(d, i, j, k)]) ? 1u : 0u;
}
__global__ void k_row_list(Dom d, const uint8_t *lab, const uint32_t *offs, uint32_t *rows) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n12 = int64_t(d.n[1]) * d.n[2];
  if (q >= int64_t(d.n[0]) * n12) return;
  const int i = int(q / n12), j = int((q / d.n[2]) % d.n[1]), k = int(q % d.n[2]);
  const int64_t v = vox_of(d, i, j, k);
  if (is_fluid(lab[v])) rows[offs[q]] = uint32_t(v);
}
// The sequential sums of a0.b0 (K = 1) and also a1.b1 (K = 2) over the row
// list. One workgroup: chunks of 1024 rows are loaded and multiplied by all
// threads into a double-buffered LDS chunk, and thread 0 adds each chunk in
// order while the next one loads (one barrier per chunk). The totals replace
// a pass's published partials: entry 0 of `part` holds them and every other
// workgroup slot holds 0, so the pass's own total over g partials returns
// them unchanged. For checks of the stage arithmetic against the oracle, not
// for speed.
constexpr int kRowChunk = 1024;
template <typename V, typename A, int K>
__global__ __launch_bounds__(kRowChunk) void k_rowdot(const uint32_t *rows,
                                                      const unsigned long long *nrows,
                                                      const V *a0, const V *b0, const V *a1,
                                                      const V *b1, A *part, int g,
                                                      const PcgState *st, int force) {
#pragma clang fp contract(off)
  __shared__ A s_m[2][K][kRowChunk];
  if ((!force && st->done) || st->status) return;
  const int t = threadIdx.x;
  const int64_t n = int64_t(*nrows);
  A sum[K];
#pragma unroll
  for (int j = 0; j < K; ++j) sum[j] = A(0);
  int buf = 0;
  for (int64_t c0 = 0; c0 < n; c0 += kRowChunk, buf ^= 1) {
    const int64_t q = c0 + t;
    A m[K];
#pragma unroll
    for (int j = 0; j < K; ++j) m[j] = A(0);
    if (q < n) {
      const uint32_t v = rows[q];
      m[0] = A(a0[v]) * A(b0[v]);
      if (K == 2) m[K - 1] = A(a1[v]) * A(b1[v]);
    }
#pragma unroll
    for (int j = 0; j < K; ++j) s_m[buf][j][t] = m[j];
    __syncthreads();
    if (t == 0) {
      const int cnt = n - c0 < kRowChunk ? int(n - c0) : kRowChunk;
      for (int u = 0; u < cnt; ++u) {
#pragma unroll
        for (int j = 0; j < K; ++j) sum[j] = sum[j] + s_m[buf][j][u];
      }
    }
  }
  // Thread 0 holds the sums.
  if (t == 0) {
#pragma unroll
    for (int j = 0; j < K; ++j) part[j] = sum[j];
  }
  for (int e = K + t; e < K * g; e += kRowChunk) part[e] = A(0);
}

// After a forced refresh: the true r.r decides whether the solve is done.
__global__ __launch_bounds__(1024) void k_refresh_scalars(const double *part_b, int g,
                                                          PcgState *st) {
  const double rr = total<double, 2>(part_b, g, 0);
  if (threadIdx.x != 0) return;
  st->rr_true = rr;
  st->rr = rr;
  if (isnan(rr)) st->status = FSL_ERR_NAN;
  st->done = (st->status != 0 || converged(st, rr) || st->it >= st->max_iter) ? 1 : 0;
}

// CopySolution (solver.cc:93-99): the pressure grid holds the solution on
// FLUID cells and 0 elsewhere (all active leaves are written).
template <typename V, bool MIXED>
__global__ __launch_bounds__(kT) void k_finish(const int32_t *list, int64_t nlist,
                                               const uint8_t *lab, const V *x, const double *x64,
                                               float *pres) {
  for (Walk q = walk(nlist); q.more(); q.next()) {
    if (!q.valid()) continue;
    const int64_t base = int64_t(list[q.mine()]) << 9;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int64_t v = base | (gtid() + s * kG);
      float out = 0.0f;
      if (is_fluid(lab[v])) out = MIXED ? float(x64[v] + double(x[v])) : float(x[v]);
      pres[v] = out;
    }
  }
}

// ||b - A x||^2 in double for the solver's own solution (stats only).
template <typename V, bool MIXED>
__global__ __launch_bounds__(kT) void k_true_residual(Dom d, const int32_t *flist, int64_t nf,
                                                      const uint8_t *lab, const float *div,
                                                      const V *x, const double *x64,
                                                      double *part) {
  __shared__ double s_x[kGroups][1000];
  __shared__ uint8_t s_f[kGroups][1000];
  const int g = threadIdx.x / kG;
  double acc = 0.0;
  for (Walk q = walk(nf); q.more(); q.next()) {
    const bool valid = q.valid();
    const int64_t leaf = valid ? flist[q.mine()] : 0;
    int o[3];
    leaf_origin(d, leaf, o);
    uint8_t ol[2] = {0, 0};
    if (valid)
      fill_tile<double>(d, lab, leaf, o, s_x[g], s_f[g], ol, [&](int64_t v, uint8_t) {
        return MIXED ? x64[v] + double(x[v]) : double(x[v]);
      });
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int c = gtid() + s * kG;
      const int64_t v = (leaf << 9) | c;
      if (!valid || !is_fluid(ol[s])) continue;
      const double rt =
          double(div[v]) - stencil<double, false>(s_x[g], s_f[g], tix_of_local(c), ol[s] >> 2);
      acc += rt * rt;
    }
    __syncthreads();
  }
  double vals[1] = {acc};
  publish<double, 1>(vals, part);
}

__global__ __launch_bounds__(1024) void k_total_scalar(const double *part, int g, PcgState *st) {
  const double t = total<double, 1>(part, g, 0);
  if (threadIdx.x == 0) st->rr_true = t;
}

// Workgroups for a list of n leaves: up to `cap` per XCD.
inline int grid_for(int64_t n, int cap = kBlocksPerXcd) {
  const int64_t chunk = (n + kXcd - 1) / kXcd;
  const int64_t per = (chunk + kGroups - 1) / kGroups;
  return int(kXcd * std::max<int64_t>(1, std::min<int64_t>(per, cap)));
}
// k_cg and k_gamma_delta run one leaf group (256 threads) per workgroup, as
// many workgroups per CU as the kernel's registers and LDS allow (the
// runtime's occupancy figure), all resident at once.
constexpr int kCgGroups = 1;
template <typename K>
int resident_per_xcd(K kernel, int threads, int max_per_cu) {
  static int cached = 0;  // one per kernel instantiation
  if (!cached) {
    int per_cu = 0, dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
        per_cu <= 0)
      return 32;
    cached = std::max(1, std::min(per_cu, max_per_cu) * std::max(1, ncu / kXcd));
  }
  return cached;
}
template <typename K>
inline int grid_cg(int64_t n, K kernel) {
  const int64_t cap = std::min(resident_per_xcd(kernel, kG * kCgGroups, kCgMaxPerCu), 256);
  const int64_t chunk = (n + kXcd - 1) / kXcd;
  return int(kXcd * std::max<int64_t>(1, std::min<int64_t>(chunk, cap)));
}

// Partial-sum buffers, carved from ctx->partials: seven slots of up to four
// doubles per workgroup.
constexpr int kPartSlots = 7;
constexpr size_t kPartSlot = size_t(kXcd) * 256 * 4 * sizeof(double);
struct Parts {
  void *a, *b[2], *bb, *r, *f[2];
};
inline Parts parts(fsl_ctx *c) {
  char *base = static_cast<char *>(c->partials);
  const size_t slot = kPartSlot;
  return {base,
          {base + slot, base + 2 * slot},
          base + 3 * slot,
          base + 4 * slot,
          {base + 5 * slot, base + 6 * slot}};
}

template <typename V, typename A, bool ORD, bool MIXED>
struct Kernels {
  static void prepare(fsl_ctx *c, int warm) {
    const int g = grid_for(c->n_fluid);
    Parts P = parts(c);
    k_zero<V><<<grid_for(c->n_act), kT, 0, c->stream>>>(c->act_list, c->n_act, (V *)c->z,
                                                         (V *)c->p[0], (V *)c->p[1]);
    k_prepare<V, A, ORD, MIXED><<<g, kT, 0, c->stream>>>(
        c->dom, c->fluid_list, c->n_fluid, c->lab, c->div, c->pres, warm, (V *)c->x, (V *)c->r,
        (V *)c->z, c->x64, (A *)P.b[1], (A *)P.bb);
    if (c->pcg_ic) ic_apply(c, P.b[1], 0);
    if (c->pcg_roworder) {
      build_row_list(c);
      rowdot_b(c, P.b[1], 1);
    }
    k_prepare_scalars<A><<<1, 1024, 0, c->stream>>>((const A *)P.b[1], (const A *)P.bb, g, c->pcg);
  }
  // Block IC(0) in place of B's Jacobi z: z = M^-1 r, then (r.r, r.z) into
  // `part` (B's partials of the iteration).
  static void ic_apply(fsl_ctx *c, void *part, int force) {
    ic_sweeps(c, force);
    k_ic_dots<V, A><<<grid_for(c->n_fluid), kT, 0, c->stream>>>(
        c->fluid_list, c->n_fluid, c->lab, (const V *)c->r, (const V *)c->z, (A *)part, c->pcg,
        force);
  }
  // B(k) and, with IC, the preconditioner apply that completes it. Every
  // caller follows B(k) with the scheduled MIXED refresh when one is due,
  // and the refresh applies IC to the refreshed residual itself.
  static void pass_b_pc(fsl_ctx *c, int k, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    pass_b(c, k, e0, e1);
    const bool refresh_due = c->pcg_precision == FSL_PCG_MIXED && (k + 1) % c->pcg_refresh == 0;
    if (c->pcg_ic && !refresh_due) ic_apply(c, parts(c).b[k & 1], 0);
    if (c->pcg_roworder) rowdot_b(c, parts(c).b[k & 1], 0);
  }
  // With events, hipExtLaunchKernelGGL records the dispatch's own start and
  // end (the kernel duration, as a profiler reports it).
  static void pass_a(fsl_ctx *c, int k, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const int g = grid_for(c->n_fluid);
    Parts P = parts(c);
    hipExtLaunchKernelGGL(k_pass_a<V, A, ORD>, dim3(g), dim3(kT), 0, c->stream, e0, e1, 0,
                          c->dom, (const int32_t *)c->fluid_list, c->n_fluid,
                          (const uint8_t *)c->lab, (const V *)c->z,
                          (const V *)c->p[(k + 1) & 1], (V *)c->p[k & 1], (V *)c->Ap,
                          (const A *)P.b[(k + 1) & 1], (A *)P.a, g, k, c->pcg);
    if (c->pcg_roworder)
      k_rowdot<V, A, 1><<<1, kRowChunk, 0, c->stream>>>(
          row_list(c), &c->flags->n_rows, (const V *)c->p[k & 1], (const V *)c->Ap, nullptr,
          nullptr, (A *)P.a, g, c->pcg, 0);
  }
  // The row list of the row-order dot products lives in the x64 buffer
  // (used by the mixed solve only, which row order excludes).
  static const uint32_t *row_list(fsl_ctx *c) { return reinterpret_cast<uint32_t *>(c->x64); }
  static void build_row_list(fsl_ctx *c) {
    const Dom &d = c->dom;
    const int64_t ncell = int64_t(d.n[0]) * d.n[1] * d.n[2];
    // Row numbers in the Ap buffer (free until the first pass A).
    uint32_t *offs = reinterpret_cast<uint32_t *>(c->Ap);
    const unsigned blocks = unsigned((ncell + 255) / 256);
    k_row_flag<<<blocks, 256, 0, c->stream>>>(d, c->lab, offs);
    (void)exclusive_scan(offs, offs, ncell, c->scan_tmp, c->stream);
    k_row_list<<<blocks, 256, 0, c->stream>>>(d, c->lab, offs, reinterpret_cast<uint32_t *>(c->x64));
  }
  // (r.r, r.z) of B(k) (or of the preconditioner apply completing it) in the
  // oracle's row order, in place of the pass's partials.
  static void rowdot_b(fsl_ctx *c, void *part, int force) {
    k_rowdot<V, A, 2><<<1, kRowChunk, 0, c->stream>>>(
        row_list(c), &c->flags->n_rows, (const V *)c->r, (const V *)c->r, (const V *)c->r,
        (const V *)c->z, (A *)part, grid_for(c->n_fluid), c->pcg, force);
  }
  static void pass_b(fsl_ctx *c, int k, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const int g = grid_for(c->n_fluid);
    Parts P = parts(c);
    hipExtLaunchKernelGGL(k_pass_b<V, A, ORD>, dim3(g), dim3(kT), 0, c->stream, e0, e1, 0,
                          (const int32_t *)c->fluid_list, c->n_fluid, (const uint8_t *)c->lab,
                          (const V *)c->p[k & 1], (const V *)c->Ap, (V *)c->x, (V *)c->r,
                          (V *)c->z, (const A *)P.a, (A *)P.b[k & 1], g, k, c->pcg);
  }
  static void stage_pap(fsl_ctx *c) {
    k_stage_pap<A><<<1, 1024, 0, c->stream>>>((const A *)parts(c).a, grid_for(c->n_fluid), c->pcg);
  }
  static void stage_scalars(fsl_ctx *c, int k) {
    const int g = grid_for(c->n_fluid);
    Parts P = parts(c);
    k_stage_scalars<A><<<1, 1024, 0, c->stream>>>((const A *)P.b[k & 1], g, k, c->pcg);
  }
  static void finish(fsl_ctx *c) {
    k_finish<V, MIXED><<<grid_for(c->n_act), kT, 0, c->stream>>>(
        c->act_list, c->n_act, c->lab, (const V *)c->x, c->x64, c->pres);
  }
  // Single-reduction path: r_k lives in rb(k & 1), p_k in p[k & 1], and the
  // totals (gamma_k, delta_k, r.r_k) in f[k & 1].
  static V *rb(fsl_ctx *c, int k) { return (V *)(k & 1 ? c->z : c->r); }
  // The k_cg variant for this context: 32-bit voxel offsets when every
  // indexed array is under 4 GiB.
  static auto cg_kernel(fsl_ctx *c) {
    // x64 (double) is indexed too: 8 bytes per voxel.
    return uint64_t(c->nvox) * 8 < (uint64_t(1) << 32)
               ? k_cg<V, A, ORD, kCgGroups, true>
               : k_cg<V, A, ORD, kCgGroups, false>;
  }
  // Workgroups of every single-reduction kernel (they publish partials for
  // one another, so they share the count).
  static int sr_grid(fsl_ctx *c) { return grid_cg(c->n_fluid, k_cg<V, A, ORD, kCgGroups, true>); }
  static void sr_prepare(fsl_ctx *c, int warm) {
    const int gp = grid_for(c->n_fluid), g = sr_grid(c);
    Parts P = parts(c);
    k_prepare<V, A, ORD, MIXED><<<gp, kT, 0, c->stream>>>(
        c->dom, c->fluid_list, c->n_fluid, c->lab, c->div, c->pres, warm, (V *)c->x, rb(c, 0),
        nullptr, c->x64, (A *)P.b[1], (A *)P.bb);
    k_gamma_delta<V, A, ORD, kCgGroups><<<g, kG * kCgGroups, 0, c->stream>>>(
        c->dom, c->fluid_list, c->n_fluid, c->lab, rb(c, 0), (A *)P.f[0], c->pcg, 1);
    k_prepare_scalars_sr<A><<<1, 1024, 0, c->stream>>>((const A *)P.f[0], g, (const A *)P.bb, gp,
                                                       c->pcg);
  }
  // mode: bit 0, x is updated in iteration k (odd k); bit 1, x lacks
  // alpha_{k-1} p_{k-1} (deferred by iteration k-1).
  static void sr_step(fsl_ctx *c, int k, int mode, hipEvent_t e0 = nullptr,
                      hipEvent_t e1 = nullptr) {
    const int g = sr_grid(c);
    Parts P = parts(c);
    hipExtLaunchKernelGGL(cg_kernel(c), dim3(g), dim3(kG * kCgGroups), 0, c->stream, e0, e1, 0, c->dom,
                          (const int32_t *)c->fluid_list, c->n_fluid, (const uint8_t *)c->lab,
                          (const V *)rb(c, k), rb(c, k + 1), (const V *)c->p[(k + 1) & 1],
                          (V *)c->p[k & 1], (V *)c->x, c->x64, (const A *)P.f[k & 1],
                          (A *)P.f[(k + 1) & 1], g, k, mode, c->pcg);
  }
  static void true_residual(fsl_ctx *c) {
    const int g = grid_for(c->n_fluid);
    Parts P = parts(c);
    k_true_residual<V, MIXED><<<g, kT, 0, c->stream>>>(c->dom, c->fluid_list, c->n_fluid, c->lab,
                                                      c->div, (const V *)c->x, c->x64,
                                                      (double *)P.r);
    k_total_scalar<<<1, 1024, 0, c->stream>>>((const double *)P.r, g, c->pcg);
  }
};

// The refresh after B(k) overwrites B(k)'s partials.
void launch_refresh(fsl_ctx *c, int k, int force) {
  const int g = grid_for(c->n_fluid);
  Parts P = parts(c);
  k_refresh_x<<<g, kT, 0, c->stream>>>(c->fluid_list, c->n_fluid, c->lab, (float *)c->x, c->x64,
                                       c->pcg, force);
  k_refresh_r<<<g, kT, 0, c->stream>>>(c->dom, c->fluid_list, c->n_fluid, c->lab, c->div,
                                       c->x64, (float *)c->r, (float *)c->z, (double *)P.b[k & 1],
                                       c->pcg, force);
  if (c->pcg_ic) Kernels<float, double, false, true>::ic_apply(c, P.b[k & 1], force);
  if (force) k_refresh_scalars<<<1, 1024, 0, c->stream>>>((const double *)P.b[k & 1], g, c->pcg);
}

// Single-reduction path, MIXED: replace r_kn (input of iteration kn) by the
// FP64 true residual, and its totals by those of the true residual.
// kn < 0: a convergence check whose iteration is taken from the state on
// the device (k_refresh_cg, force 2).
void launch_refresh_sr(fsl_ctx *c, int kn, int force) {
  const int g = Kernels<float, double, false, true>::sr_grid(c);
  Parts P = parts(c);
  const bool from_state = kn < 0;
  float *rbuf = (float *)(kn & 1 ? c->z : c->r);
  double *part = (double *)P.f[kn & 1];
  // Before a scheduled refresh the k_cg launch has folded x into x64; a
  // forced refresh (convergence check) folds here.
  if (force)
    k_refresh_x<<<grid_for(c->n_fluid), kT, 0, c->stream>>>(c->fluid_list, c->n_fluid, c->lab,
                                                          (float *)c->x, c->x64, c->pcg,
                                                          from_state ? 2 : force);
  auto kern = uint64_t(c->nvox) * 8 < (uint64_t(1) << 32) ? k_refresh_cg<true> : k_refresh_cg<false>;
  if (from_state) {
    kern<<<g, kG, 0, c->stream>>>(c->dom, c->fluid_list, c->n_fluid, c->lab, c->div, c->x64,
                                  (float *)c->r, (double *)P.f[0], (float *)c->z,
                                  (double *)P.f[1], c->pcg, 2);
    k_refresh_scalars_sr<<<1, 1024, 0, c->stream>>>((const double *)P.f[0], g, c->pcg,
                                                    (const double *)P.f[1]);
    return;
  }
  kern<<<g, kG, 0, c->stream>>>(c->dom, c->fluid_list, c->n_fluid, c->lab, c->div, c->x64, rbuf,
                                part, rbuf, part, c->pcg, force);
  if (force)
    k_refresh_scalars_sr<<<1, 1024, 0, c->stream>>>(part, g, c->pcg, nullptr);
}



template <typename F>
void dispatch(int precision, F &&fn) {
  if (precision == FSL_PCG_FP64) fn(Kernels<double, double, false, false>());
  else if (precision == FSL_PCG_MIXED) fn(Kernels<float, double, false, true>());
  else fn(Kernels<float, float, true, false>());
}


int sync_state(fsl_ctx *ctx) {
  FSL_HIP(hipMemcpyAsync(ctx->pcg_host, ctx->pcg, sizeof(PcgState), hipMemcpyDeviceToHost,
                         ctx->stream));
  FSL_HIP(hipStreamSynchronize(ctx->stream));
  return FSL_OK;
}

int clear_done(fsl_ctx *ctx) {
  ctx->pcg_host->done = 0;
  FSL_HIP(hipMemcpyAsync(&ctx->pcg->done, &ctx->pcg_host->done, sizeof(int),
                         hipMemcpyHostToDevice, ctx->stream));
  return FSL_OK;
}

}  // namespace

size_t pcg_partials_bytes() { return kPartSlots * kPartSlot; }

bool pcg_single_reduction(const fsl_solve_params *prm) {
  // IC(0) needs z = M^-1 r between the residual update and r.z, which the
  // single-reduction recurrence folds into one kernel: two passes only.
  if (prm->preconditioner == FSL_PRECOND_IC) return false;
  if (prm->variant == FSL_PCG_SINGLE_REDUCTION) return true;
  // Row-order dot products are a two-pass feature.
  if (prm->dot_order == FSL_DOT_ROW_ORDER) return false;
  if (prm->variant == FSL_PCG_TWO_PASS) return false;
  return prm->precision != FSL_PCG_FP32;
}

// Host-side solver state of a new solve; uploads it to the device.
static int pcg_setup(fsl_ctx *ctx, const fsl_solve_params *prm) {
  if (prm->precision < 0 || prm->precision > 2 || prm->stop_mode < 0 || prm->stop_mode > 1 ||
      prm->variant < 0 || prm->variant > 2 || prm->preconditioner < 0 ||
      prm->preconditioner > 1) {
    ctx->err = "invalid solver parameters";
    return FSL_ERR_ARG;
  }
  if (prm->preconditioner == FSL_PRECOND_IC && prm->variant == FSL_PCG_SINGLE_REDUCTION) {
    ctx->err = "the IC preconditioner runs with the two-pass variant only";
    return FSL_ERR_ARG;
  }
  if (prm->dot_order != FSL_DOT_TREE && prm->dot_order != FSL_DOT_ROW_ORDER) {
    ctx->err = "invalid dot_order";
    return FSL_ERR_ARG;
  }
  if (prm->dot_order == FSL_DOT_ROW_ORDER &&
      (prm->variant == FSL_PCG_SINGLE_REDUCTION || prm->precision == FSL_PCG_MIXED)) {
    ctx->err = "row-order dot products run with the two-pass FP32 or FP64 solve only";
    return FSL_ERR_ARG;
  }
  ctx->pcg_roworder = prm->dot_order == FSL_DOT_ROW_ORDER;
  ctx->pcg_ic = prm->preconditioner == FSL_PRECOND_IC;
  ctx->pcg_precision = prm->precision;
  ctx->pcg_stop = prm->stop_mode;
  // An even period: the single-reduction path's x is complete after odd
  // iterations only.
  ctx->pcg_refresh = prm->refresh_every > 0 ? prm->refresh_every + (prm->refresh_every & 1) : 32;
  PcgState *h = ctx->pcg_host;
  *h = PcgState{};
  h->stop_k = 0x7fffffff;
  h->max_iter = prm->max_iterations;
  h->stop_mode = prm->stop_mode;
  h->precision = prm->precision;
  // The reference compares in the solver's type T (solver_driver.cc:330).
  h->thr = prm->precision == FSL_PCG_FP32 && prm->stop_mode == FSL_STOP_ABS_RR
               ? double(float(prm->threshold))
               : prm->threshold;
  if (ctx->n_fluid == 0) h->done = 1;
  FSL_HIP(hipMemcpyAsync(ctx->pcg, h, sizeof(PcgState), hipMemcpyHostToDevice, ctx->stream));
  return FSL_OK;
}

int pcg_prepare(fsl_ctx *ctx, const fsl_solve_params *prm) {
  int rc = pcg_setup(ctx, prm);
  if (rc) return rc;
  ctx->pcg_sr = false;
  ctx->pcg_stage = 0;
  if (ctx->pcg_ic) {
    rc = ic_factor(ctx);
    if (rc) return rc;
  }
  if (ctx->n_fluid > 0) {
    dispatch(ctx->pcg_precision,
             [&](auto k) { decltype(k)::prepare(ctx, prm->warm_start ? 1 : 0); });
    FSL_HIP(hipGetLastError());
  }
  rc = sync_state(ctx);
  if (rc) return rc;
  ctx->pcg_ready = true;
  return ctx->pcg_host->status;
}

// Stage API, one stage per call (each synchronised on its own).
// ComputeLocal1 (solver.cc:296-309) of iteration it: p = z + beta p (the
// direction update of the previous iteration's ComputeLocal3, applied here,
// where the pass reads p), Ap and pAp.
int pcg_local1(fsl_ctx *ctx, double *pAp) {
  if (!ctx->pcg_ready) { ctx->err = "pcg not prepared"; return FSL_ERR_STATE; }
  if (ctx->pcg_sr) {
    ctx->err = "stage calls need a two-pass solve (fsl_pcg_prepare with the two-pass variant)";
    return FSL_ERR_STATE;
  }
  if (ctx->pcg_stage != 0) { ctx->err = "ComputeLocal1 out of order"; return FSL_ERR_STATE; }
  if (ctx->n_fluid == 0 || ctx->pcg_host->done) return ctx->pcg_host->status;
  const int k = ctx->pcg_host->it;
  dispatch(ctx->pcg_precision, [&](auto K) {
    decltype(K)::pass_a(ctx, k);
    decltype(K)::stage_pap(ctx);
  });
  FSL_HIP(hipGetLastError());
  int rc = sync_state(ctx);
  if (rc) return rc;
  if (pAp) *pAp = ctx->pcg_host->pAp;
  if (!ctx->pcg_host->done) ctx->pcg_stage = 1;
  return ctx->pcg_host->status;
}

// ComputeLocal2 (solver.cc:323-346): alpha = rz / pAp (checked against 1e5),
// x += alpha p, r -= alpha Ap, z = M^-1 r, r.r and r.z; then the reductions
// of solver_driver.cc:318-343 and beta = rz / rz_old.
int pcg_local2(fsl_ctx *ctx, double *alpha, double *rr, double *rz) {
  if (!ctx->pcg_ready || ctx->pcg_stage != 1) {
    ctx->err = "ComputeLocal2 out of order";
    return FSL_ERR_STATE;
  }
  const int k = ctx->pcg_host->it;
  dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::pass_b_pc(ctx, k); });
  if (ctx->pcg_precision == FSL_PCG_MIXED && (k + 1) % ctx->pcg_refresh == 0)
    launch_refresh(ctx, k, 0);
  dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::stage_scalars(ctx, k); });
  FSL_HIP(hipGetLastError());
  int rc = sync_state(ctx);
  if (rc) return rc;
  const PcgState *h = ctx->pcg_host;
  if (alpha) *alpha = h->alpha;
  if (rr) *rr = h->rr;
  if (rz) *rz = h->rz;
  ctx->pcg_stage = 2;
  // A beta above 1e5 is ComputeLocal3's check: reported there.
  return h->status == FSL_ERR_BETA ? FSL_OK : h->status;
}

// ComputeLocal3 (solver.cc:355-366): beta = rz / rz_old, checked against
// 1e5; p = z + beta p is applied by the next ComputeLocal1.
int pcg_local3(fsl_ctx *ctx, double *beta) {
  if (!ctx->pcg_ready || ctx->pcg_stage != 2) {
    ctx->err = "ComputeLocal3 out of order";
    return FSL_ERR_STATE;
  }
  if (beta) *beta = ctx->pcg_host->beta;
  ctx->pcg_stage = 0;
  return ctx->pcg_host->status;
}

// One While-loop body, launched and synchronised on its own (stage API).
int pcg_iterate_once(fsl_ctx *ctx) {
  if (!ctx->pcg_ready) { ctx->err = "pcg not prepared"; return FSL_ERR_STATE; }
  if (ctx->pcg_sr) {
    ctx->err = "stage calls need a two-pass solve (fsl_pcg_prepare with the two-pass variant)";
    return FSL_ERR_STATE;
  }
  if (ctx->pcg_stage != 0) { ctx->err = "iteration started with stage calls"; return FSL_ERR_STATE; }
  if (ctx->n_fluid == 0 || ctx->pcg_host->done) return ctx->pcg_host->status;
  const int k = ctx->pcg_host->it;
  dispatch(ctx->pcg_precision, [&](auto K) {
    decltype(K)::pass_a(ctx, k);
    decltype(K)::pass_b_pc(ctx, k);
  });
  if (ctx->pcg_precision == FSL_PCG_MIXED && (k + 1) % ctx->pcg_refresh == 0)
    launch_refresh(ctx, k, 0);
  dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::stage_scalars(ctx, k); });
  FSL_HIP(hipGetLastError());
  int rc = sync_state(ctx);
  if (rc) return rc;
  return ctx->pcg_host->status;
}

// The single-reduction solve: one k_cg launch per iteration. Batches of
// iterations run between host checks of the device state; launches past
// convergence exit at once (the done flag).
// Batches of iterations with one batch always queued behind the one whose
// state the host reads: after enqueueing batch c the host waits for the
// state after batch c - 1 (copied to a pinned slot), so the GPU never idles
// on the host's decision. Launches queued past convergence exit on the done
// flag (about 2 us each); adaptive batches cover at most 60% of the
// iterations the residual decay still predicts beyond those in flight.
// issue(k) enqueues iteration k; k ends past the last enqueued iteration.
// Returns once the state reports done or an error, with the final state in
// ctx->pcg_host.
template <typename Issue>
static int run_pipelined(fsl_ctx *ctx, int &k, int batch, bool adaptive, Issue &&issue) {
  struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  for (hipEvent_t &e : ev.e) FSL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  PcgState *slot = ctx->pcg_host + 1;
  auto post = [&](int c, int n) -> int {
    for (int b = 0; b < n; ++b, ++k) issue(k);
    FSL_HIP(hipGetLastError());
    FSL_HIP(hipMemcpyAsync(&slot[c & 1], ctx->pcg, sizeof(PcgState), hipMemcpyDeviceToHost,
                           ctx->stream));
    FSL_HIP(hipEventRecord(ev.e[c & 1], ctx->stream));
    return FSL_OK;
  };
  double rr_prev = ctx->pcg_host->rr;
  int it_prev = ctx->pcg_host->it;
  int nb = batch;
  int rc = post(0, nb);
  if (rc) return rc;
  for (int c = 1;; ++c) {
    rc = post(c, nb);
    if (rc) return rc;
    FSL_HIP(hipEventSynchronize(ev.e[(c - 1) & 1]));
    const PcgState &h = slot[(c - 1) & 1];
    if (h.done || h.status) break;
    if (!adaptive) continue;
    // r.r of CG is not monotone: the decay rate comes from the running
    // minimum (rr_prev = 0: no earlier value read yet).
    const double rr_min = rr_prev > 0.0 ? std::min(h.rr, rr_prev) : h.rr;
    if (h.it > it_prev && rr_min > 0.0 && rr_prev > 0.0) {
      const double rate = std::log(rr_min / rr_prev) / double(h.it - it_prev);
      const double target = h.stop_mode == FSL_STOP_REL ? h.thr * h.thr : h.thr;
      if (rate < 0.0 && target > 0.0) {
        const double need = std::log(target / rr_min) / rate - double(k - 1 - h.it);
        nb = int(std::min(256.0, std::max(8.0, 0.6 * need)));
      }
    }
    rr_prev = rr_min;
    it_prev = h.it;
  }
  // The queued batch exits on the done flag; the mirror gets the final state.
  return sync_state(ctx);
}

static int pcg_iterations_sr(fsl_ctx *ctx, const fsl_solve_params *prm, fsl_solve_stats *st) {
  int rc = pcg_setup(ctx, prm);
  if (rc) return rc;
  ctx->pcg_sr = true;
  if (ctx->n_fluid > 0) {
    dispatch(ctx->pcg_precision,
             [&](auto K) { decltype(K)::sr_prepare(ctx, prm->warm_start ? 1 : 0); });
    FSL_HIP(hipGetLastError());
  }
  // The batches start without reading the prepared state back: a prepare
  // that stops (converged at x0, or an error status) sets the done flag, on
  // which every launch exits, and the state is read after the first batch.
  ctx->pcg_ready = true;
  const bool adaptive = prm->batch <= 0;
  const int batch = prm->batch > 0 ? prm->batch : 16;
  const bool mixed = ctx->pcg_precision == FSL_PCG_MIXED;
  const int nb_cap = ctx->timing ? 100 : 1024;
  int nb = std::min(batch, nb_cap);
  double rr_prev = ctx->pcg_host->rr;
  int it_prev = 0;
  hipEvent_t e0, e1;
  FSL_HIP(hipEventCreate(&e0));
  FSL_HIP(hipEventCreate(&e1));
  FSL_HIP(hipEventRecord(e0, ctx->stream));
  hipEvent_t *pev = ctx->tev + 32;
  int k = 0;  // next iteration to launch
  // x is updated every other iteration: iteration k odd applies
  // alpha_{k-1} p_{k-1} + alpha_k p_k; a kernel that stops at an odd k applies
  // the pending term. After a forced refresh x is complete (resume_k).
  int resume_k = -1;
  auto mode_of = [&](int kk) { return sr_mode_of(kk, resume_k, mixed ? ctx->pcg_refresh : 0); };
  // Predicted solves (with the step graph on): the previous solve's
  // iteration count predicts this one's. That many launches (rounded up to
  // an even count) are enqueued at once, then the convergence check (mixed;
  // it acts only if the solve stopped), and the state is read once. The
  // host enqueues far faster than the launches run, so the launches past
  // convergence, which exit on the stop, run back to back. A solve that has
  // not stopped by then continues 32 launches (and a check) at a time, with
  // the same iterates as without the prediction; a stop the mixed check
  // rejects (the true residual does not confirm it) resumes with the
  // host-checked batches below.
  auto issue_sr = [&](int kk) {
    dispatch(ctx->pcg_precision,
             [&](auto K) { decltype(K)::sr_step(ctx, kk, mode_of(kk), nullptr, nullptr); });
    if (mixed && (kk + 1) % ctx->pcg_refresh == 0) launch_refresh_sr(ctx, kk + 1, 0);
  };
  if (!ctx->timing && adaptive && ctx->use_graph && ctx->pcg_last_iters > 0 && ctx->n_fluid > 0 &&
      !ctx->pcg_host->done) {
    int kend = std::max(2, ctx->pcg_last_iters + (ctx->pcg_last_iters & 1));
    for (;;) {
      for (; k < kend; ++k) issue_sr(k);
      if (mixed) launch_refresh_sr(ctx, -1, 1);
      FSL_HIP(hipGetLastError());
      rc = sync_state(ctx);
      if (rc) return rc;
      const PcgState *h = ctx->pcg_host;
      // The iteration a continuation starts from.
      k = mixed ? h->refresh_kn : (h->done ? h->it : h->it + 1);
      resume_k = k;
      if (h->done || h->status || k != kend || k >= h->max_iter) break;
      kend += 32;
    }
  }
  while (!ctx->timing && !ctx->pcg_host->done) {
    rc = run_pipelined(ctx, k, nb, adaptive, issue_sr);
    if (rc) return rc;
    const PcgState *h = ctx->pcg_host;
    if (h->status || !(mixed && h->it > 0 && h->it < h->max_iter)) break;
    // Converged on the recursive residual: confirm on the true residual of
    // x_it, and continue from iteration it if it does not hold.
    const int last = h->it;
    launch_refresh_sr(ctx, last, 1);
    FSL_HIP(hipGetLastError());
    rc = sync_state(ctx);
    if (rc) return rc;
    if (ctx->pcg_host->done) break;
    k = last;
    resume_k = last;
  }
  // With the timing hook, batches run one at a time: each launch records
  // its own events, which are read back after the batch.
  while (ctx->timing && !ctx->pcg_host->done) {
    const int k0 = k;
    for (int b = 0; b < nb; ++b, ++k) {
      hipEvent_t *e = pev + 2 * b;
      dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::sr_step(ctx, k, mode_of(k), e[0], e[1]); });
      if (mixed && (k + 1) % ctx->pcg_refresh == 0) launch_refresh_sr(ctx, k + 1, 0);
    }
    FSL_HIP(hipGetLastError());
    rc = sync_state(ctx);
    if (rc) return rc;
    const PcgState *h = ctx->pcg_host;
    {
      // Launches k0..k-1 ran; those before the stopping one did the work.
      const int worked = std::min(h->done ? h->it - k0 : nb, nb);
      for (int b = 0; b < worked; ++b) {
        float ms = 0.0f;
        FSL_HIP(hipEventElapsedTime(&ms, pev[2 * b], pev[2 * b + 1]));
        ctx->tacc[FSL_T_FUSED] += ms;
        ctx->tacc[FSL_T_FUSED_N] += 1;
      }
    }
    if (h->status) break;
    if (!h->done) {
      // r.r of CG is not monotone: the decay rate is taken from the running
      // minimum, and a batch covers at most 60% of the predicted remainder
      // (launches past convergence cost about 2 us each).
      const double rr_min = rr_prev > 0.0 ? std::min(h->rr, rr_prev) : h->rr;
      if (adaptive && h->it > it_prev && rr_min > 0.0 && rr_prev > 0.0) {
        const double rate = std::log(rr_min / rr_prev) / double(h->it - it_prev);
        const double target = h->stop_mode == FSL_STOP_REL ? h->thr * h->thr : h->thr;
        if (rate < 0.0 && target > 0.0) {
          const double need = std::log(target / rr_min) / rate;
          nb = int(std::min(double(std::min(nb_cap, 256)), std::max(16.0, 0.6 * need)));
        }
      }
      rr_prev = rr_min;
      it_prev = h->it;
      continue;
    }
    if (mixed && h->it > 0 && h->it < h->max_iter) {
      // Converged on the recursive residual: confirm on the true residual of
      // x_it, and continue from iteration it if it does not hold.
      const int last = h->it;
      launch_refresh_sr(ctx, last, 1);
      FSL_HIP(hipGetLastError());
      rc = sync_state(ctx);
      if (rc) return rc;
      if (!ctx->pcg_host->done) {
        k = last;
        resume_k = last;
      }
    }
  }
  FSL_HIP(hipEventRecord(e1, ctx->stream));
  FSL_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  FSL_HIP(hipEventElapsedTime(&ms, e0, e1));
  FSL_HIP(hipEventDestroy(e0));
  FSL_HIP(hipEventDestroy(e1));
  if (st) st->seconds = double(ms) * 1e-3;
  if (!ctx->pcg_host->status) ctx->pcg_last_iters = ctx->pcg_host->it;
  return ctx->pcg_host->status;
}

int pcg_solve_iterations(fsl_ctx *ctx, const fsl_solve_params *prm, fsl_solve_stats *st) {
  if (pcg_single_reduction(prm)) return pcg_iterations_sr(ctx, prm, st);
  int rc = pcg_prepare(ctx, prm);
  if (rc) return rc;
  const int batch = prm->batch > 0 ? prm->batch : 16;
  // Batch sizing: after each host check, the next batch covers most of the
  // iterations the observed residual decay predicts are still needed, so a
  // solve takes a few host round trips instead of one per `batch`.
  // Iterations launched past convergence exit at once (the done flag).
  const bool adaptive = prm->batch <= 0;
  double rr_prev = ctx->pcg_host->rr;
  int k_prev = 0;
  hipEvent_t e0, e1;
  FSL_HIP(hipEventCreate(&e0));
  FSL_HIP(hipEventCreate(&e1));
  FSL_HIP(hipEventRecord(e0, ctx->stream));
  const bool mixed = ctx->pcg_precision == FSL_PCG_MIXED;
  // With the measurement hook on, each pass records its own start and end
  // events (at most 50 iterations per batch, four events per iteration).
  const int nb_cap = ctx->timing ? 50 : 1024;
  int nb = std::min(batch, nb_cap);
  hipEvent_t *pev = ctx->tev + 32;
  int k = 0;  // next iteration to launch
  // Without the timing hook: pipelined batches of A(k), B(k) pairs. A(k)
  // holds the stopping test of iteration k, and every pass after a stopping
  // A(k) exits on the done flag.
  while (!ctx->timing && !ctx->pcg_host->done) {
    rc = run_pipelined(ctx, k, nb, adaptive, [&](int kk) {
      dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::pass_a(ctx, kk); });
      dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::pass_b_pc(ctx, kk); });
      if (mixed && (kk + 1) % ctx->pcg_refresh == 0) launch_refresh(ctx, kk, 0);
    });
    if (rc) return rc;
    const PcgState *h = ctx->pcg_host;
    if (h->status || !(mixed && h->it > 0 && h->it < h->max_iter)) break;
    // Converged on the recursive residual: confirm on the true residual
    // (refresh after B(it-1)), and continue from A(it) if it does not hold.
    const int last = h->it;
    launch_refresh(ctx, last - 1, 1);
    FSL_HIP(hipGetLastError());
    rc = sync_state(ctx);
    if (rc) return rc;
    if (ctx->pcg_host->done) break;
    ctx->pcg_host->it = last;
    k = last;
  }
  // With the timing hook, batches run one at a time, each pass recording its
  // own events; a batch ends with the stopping test A(k).
  while (ctx->timing && !ctx->pcg_host->done) {
    const int k0 = k;
    for (int b = 0; b < nb; ++b, ++k) {
      hipEvent_t *e = pev + 4 * b;
      dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::pass_a(ctx, k, e[0], e[1]); });
      dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::pass_b_pc(ctx, k, e[2], e[3]); });
      if (mixed && (k + 1) % ctx->pcg_refresh == 0) launch_refresh(ctx, k, 0);
    }
    // The stopping test of iteration k runs at the start of A(k).
    dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::pass_a(ctx, k); });
    FSL_HIP(hipGetLastError());
    rc = sync_state(ctx);
    if (rc) return rc;
    const PcgState *h = ctx->pcg_host;
    {
      // A(j) and B(j) did work for j < it (done) or j < k (running).
      const int worked = std::min((h->done ? h->it : k) - k0, nb);
      for (int b = 0; b < worked; ++b) {
        hipEvent_t *e = pev + 4 * b;
        float ms[2] = {0.0f, 0.0f};
        FSL_HIP(hipEventElapsedTime(&ms[0], e[0], e[1]));
        FSL_HIP(hipEventElapsedTime(&ms[1], e[2], e[3]));
        ctx->tacc[FSL_T_PASS_A] += ms[0];
        ctx->tacc[FSL_T_PASS_B] += ms[1];
        ctx->tacc[FSL_T_PASS_A_N] += 1;
        ctx->tacc[FSL_T_PASS_B_N] += 1;
      }
    }
    if (h->status) break;
    if (!h->done) {
      if (adaptive && k > k_prev && h->rr > 0.0 && rr_prev > 0.0) {
        const double rate = std::log(h->rr / rr_prev) / double(k - k_prev);
        const double target = h->stop_mode == FSL_STOP_REL ? h->thr * h->thr : h->thr;
        if (rate < 0.0 && target > 0.0) {
          const double need = std::log(target / h->rr) / rate;
          nb = int(std::min(double(nb_cap), std::max(16.0, 0.8 * need)));
        }
      }
      rr_prev = h->rr;
      k_prev = k;
      // A(k) ran for real: the next batch starts at B(k).
      dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::pass_b_pc(ctx, k); });
      if (mixed && (k + 1) % ctx->pcg_refresh == 0) launch_refresh(ctx, k, 0);
      ++k;
      continue;
    }
    if (mixed && h->it > 0 && h->it < h->max_iter) {
      // Converged on the recursive residual: confirm on the true residual
      // (refresh after B(it-1)), and continue from A(it) if it does not hold.
      const int last = h->it;
      launch_refresh(ctx, last - 1, 1);
      FSL_HIP(hipGetLastError());
      rc = sync_state(ctx);
      if (rc) return rc;
      if (!ctx->pcg_host->done) {
        ctx->pcg_host->it = last;
        k = last;
      }
    }
  }
  FSL_HIP(hipEventRecord(e1, ctx->stream));
  FSL_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  FSL_HIP(hipEventElapsedTime(&ms, e0, e1));
  FSL_HIP(hipEventDestroy(e0));
  FSL_HIP(hipEventDestroy(e1));
  if (st) st->seconds = double(ms) * 1e-3;
  return ctx->pcg_host->status;
}

int pcg_run(fsl_ctx *ctx, const fsl_solve_params *prm, fsl_solve_stats *st) {
  const int rc = pcg_solve_iterations(ctx, prm, st);
  if (rc) return rc;
  return pcg_finish(ctx, st);
}

int pcg_true_residual(fsl_ctx *ctx, double *rr_true) {
  *rr_true = 0.0;
  if (ctx->n_fluid == 0) return FSL_OK;
  dispatch(ctx->pcg_precision, [&](auto k) { decltype(k)::true_residual(ctx); });
  FSL_HIP(hipGetLastError());
  int rc = sync_state(ctx);
  if (rc) return rc;
  *rr_true = ctx->pcg_host->rr_true;
  return FSL_OK;
}

int pcg_finish_enqueue(fsl_ctx *ctx) {
  if (!ctx->pcg_ready) { ctx->err = "pcg not prepared"; return FSL_ERR_STATE; }
  if (ctx->n_act > 0) {
    dispatch(ctx->pcg_precision, [&](auto k) { decltype(k)::finish(ctx); });
    FSL_HIP(hipGetLastError());
  }
  if (ctx->n_fluid > 0) {
    dispatch(ctx->pcg_precision, [&](auto k) { decltype(k)::true_residual(ctx); });
    FSL_HIP(hipGetLastError());
  }
  if (!ctx->ev_finish) FSL_HIP(hipEventCreateWithFlags(&ctx->ev_finish, hipEventDisableTiming));
  FSL_HIP(hipMemcpyAsync(ctx->pcg_host, ctx->pcg, sizeof(PcgState), hipMemcpyDeviceToHost,
                         ctx->stream));
  FSL_HIP(hipEventRecord(ctx->ev_finish, ctx->stream));
  return FSL_OK;
}

int pcg_finish_collect(fsl_ctx *ctx, fsl_solve_stats *st) {
  FSL_HIP(hipEventSynchronize(ctx->ev_finish));
  const double rr_true = ctx->n_fluid > 0 ? ctx->pcg_host->rr_true : 0.0;
  if (st) {
    st->iterations = ctx->pcg_host->it;
    st->rr = ctx->pcg_host->rr;
    st->bnorm = sqrt(ctx->pcg_host->bnorm2);
    st->true_rel_residual = st->bnorm > 0 ? sqrt(rr_true) / st->bnorm : sqrt(rr_true);
    st->n_fluid_leaves = int(ctx->n_fluid);
  }
  return ctx->pcg_host->status;
}

int pcg_finish(fsl_ctx *ctx, fsl_solve_stats *st) {
  const int rc = pcg_finish_enqueue(ctx);
  if (rc) return rc;
  return pcg_finish_collect(ctx, st);
}

}  // namespace fsl
