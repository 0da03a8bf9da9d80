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
//   MIXED - float vectors and stencil, double dot products (exact products,
//           Neumaier-compensated across workgroups), and every `refresh`
//           iterations the float iterate is folded into a double solution and
//           the true residual b - A x is recomputed in double (a reliable
//           update; CG continues with the same direction).
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "device_util.h"
#include "fsl_internal.h"

namespace fsl {

namespace {

constexpr int kG = 256;             // threads per leaf group (two cells each)
constexpr int kGroups = 4;          // leaves in flight per workgroup
constexpr int kT = kG * kGroups;    // threads per workgroup
constexpr int kBlocksPerXcd = 64;   // 2 workgroups per CU, 32 CUs per XCD

template <typename V>
__device__ inline V dinv(int diag);
template <>
__device__ inline float dinv<float>(int diag) {
  // JacobiPreconditioner stores float(1.0 / diag); diag is 1..6.
  constexpr float t[8] = {0.0f, 1.0f, 0.5f, float(1.0 / 3.0), 0.25f, 0.2f, float(1.0 / 6.0), 0.0f};
  return t[diag & 7];
}
template <>
__device__ inline double dinv<double>(int diag) {
  return diag ? 1.0 / double(diag) : 0.0;
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
    return __builtin_fma(V(diag), sp[c], -nb);
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
    return __builtin_fma(V(diag), sp[c], -nb);
  }
}

// Persistent walk over the leaf list: XCD x = blockIdx % 8 owns the list
// chunk x; its workgroups stride through it, kGroups leaves per round (leaf
// group g = threadIdx / kG takes every kGroups-th leaf of the workgroup).
struct Walk {
  int64_t w, end, step;
  __device__ bool more() const { return w < end; }
  __device__ int64_t mine() const { return w + int64_t(threadIdx.x / kG) * step; }
  __device__ bool valid() const { return mine() < end; }
  __device__ void next() { w += int64_t(kGroups) * step; }
};
__device__ inline Walk walk(int64_t nf) {
  const int64_t per = gridDim.x / kXcd;
  const int64_t chunk = (nf + kXcd - 1) / kXcd;
  const int64_t beg = int64_t(blockIdx.x % kXcd) * chunk;
  return {beg + blockIdx.x / kXcd, beg + chunk < nf ? beg + chunk : nf, per};
}
__device__ inline int gtid() { return threadIdx.x % kG; }

// Block-reduce K values and publish this workgroup's partial sums.
template <typename A, int K>
__device__ inline void publish(A vals[K], A *partials) {
  __shared__ A s_red[K][16];
#pragma unroll
  for (int k = 0; k < K; ++k) vals[k] = block_sum<A>(vals[k], s_red[k]);
  if (threadIdx.x == 0) {
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

// Totals of both values of the n published (stride 2) partials in one pass:
// one load round and one combined workgroup reduction, in A's precision
// (Neumaier-compensated for double), the same in every workgroup.
template <typename A>
__device__ inline void total2(const A *partials, int n, double out[2]) {
  __shared__ A s_t[2][16];
  A v0, v1;
  if constexpr (sizeof(A) == 8) {
    double a0 = 0.0, c0 = 0.0, a1 = 0.0, c1 = 0.0;
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
      neumaier_add(a0, c0, double(partials[2 * q]));
      neumaier_add(a1, c1, double(partials[2 * q + 1]));
    }
    v0 = a0 + c0;
    v1 = a1 + c1;
  } else {
    v0 = A(0);
    v1 = A(0);
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
      v0 += partials[2 * q];
      v1 += partials[2 * q + 1];
    }
  }
  v0 = wave_sum(v0);
  v1 = wave_sum(v1);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  if (lane == 0) { s_t[0][wave] = v0; s_t[1][wave] = v1; }
  __syncthreads();
  A t0 = A(0), t1 = A(0);
  for (int w = 0; w < nw; ++w) { t0 += s_t[0][w]; t1 += s_t[1][w]; }
  out[0] = double(t0);
  out[1] = double(t1);
}


__device__ inline bool converged(const PcgState *st, double rr) {
  // Relative mode uses <=, so a zero right-hand side (x0 = 0) stops at once.
  return st->stop_mode == FSL_STOP_ABS_RR ? rr < st->thr : sqrt(rr) <= st->thr;
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
      zv[v] = z;
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
  st->status = isnan(rr) ? FSL_ERR_NAN : 0;
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
    total2<A>(part_b1, g_parts, tot);
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
        xn = __builtin_fma(alpha, c.p[s], c.x[s]);
        rn = __builtin_fma(-alpha, c.a[s], c.r[s]);
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

// MIXED refresh, part 1: fold the float iterate into the double solution.
__global__ __launch_bounds__(kT) void k_refresh_x(const int32_t *flist, int64_t nf,
                                                  const uint8_t *lab, float *x, double *x64,
                                                  const PcgState *st, int force) {
  if ((!force && st->done) || st->status) return;
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
      zv[v] = zf;
      acc[0] += double(rf) * double(rf);
      acc[1] += double(rf) * double(zf);
    }
    __syncthreads();
  }
  publish<double, 2>(acc, part_b);
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

// Workgroups for a list of n leaves: up to kBlocksPerXcd per XCD.
inline int grid_for(int64_t n) {
  const int64_t chunk = (n + kXcd - 1) / kXcd;
  const int64_t per = (chunk + kGroups - 1) / kGroups;
  return int(kXcd * std::max<int64_t>(1, std::min<int64_t>(per, kBlocksPerXcd)));
}

// Partial-sum buffers, carved from ctx->partials (sized for 8 x 1024 x 3 doubles).
struct Parts {
  void *a, *b[2], *bb, *r;
};
inline Parts parts(fsl_ctx *c) {
  char *base = static_cast<char *>(c->partials);
  const size_t slot = size_t(kXcd) * kBlocksPerXcd * 2 * sizeof(double);
  return {base, {base + slot, base + 2 * slot}, base + 3 * slot, base + 4 * slot};
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
    k_prepare_scalars<A><<<1, 1024, 0, c->stream>>>((const A *)P.b[1], (const A *)P.bb, g, c->pcg);
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
  }
  static void pass_b(fsl_ctx *c, int k, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    const int g = grid_for(c->n_fluid);
    Parts P = parts(c);
    hipExtLaunchKernelGGL(k_pass_b<V, A, ORD>, dim3(g), dim3(kT), 0, c->stream, e0, e1, 0,
                          (const int32_t *)c->fluid_list, c->n_fluid, (const uint8_t *)c->lab,
                          (const V *)c->p[k & 1], (const V *)c->Ap, (V *)c->x, (V *)c->r,
                          (V *)c->z, (const A *)P.a, (A *)P.b[k & 1], g, k, c->pcg);
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
  if (force) k_refresh_scalars<<<1, 1024, 0, c->stream>>>((const double *)P.b[k & 1], g, c->pcg);
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

size_t pcg_partials_bytes() { return size_t(5) * kXcd * kBlocksPerXcd * 2 * sizeof(double); }

int pcg_prepare(fsl_ctx *ctx, const fsl_solve_params *prm) {
  if (prm->precision < 0 || prm->precision > 2 || prm->stop_mode < 0 || prm->stop_mode > 1) {
    ctx->err = "invalid solver parameters";
    return FSL_ERR_ARG;
  }
  ctx->pcg_precision = prm->precision;
  ctx->pcg_stop = prm->stop_mode;
  ctx->pcg_refresh = prm->refresh_every > 0 ? prm->refresh_every : 32;
  PcgState *h = ctx->pcg_host;
  *h = PcgState{};
  h->max_iter = prm->max_iterations;
  h->stop_mode = prm->stop_mode;
  h->precision = prm->precision;
  // The reference compares in the solver's type T (solver_driver.cc:330).
  h->thr = prm->precision == FSL_PCG_FP32 && prm->stop_mode == FSL_STOP_ABS_RR
               ? double(float(prm->threshold))
               : prm->threshold;
  if (ctx->n_fluid == 0) h->done = 1;
  FSL_HIP(hipMemcpyAsync(ctx->pcg, h, sizeof(PcgState), hipMemcpyHostToDevice, ctx->stream));
  if (ctx->n_fluid > 0) {
    dispatch(ctx->pcg_precision,
             [&](auto k) { decltype(k)::prepare(ctx, prm->warm_start ? 1 : 0); });
    FSL_HIP(hipGetLastError());
  }
  int rc = sync_state(ctx);
  if (rc) return rc;
  ctx->pcg_ready = true;
  return ctx->pcg_host->status;
}

// One While-loop body, launched and synchronised on its own (stage API).
int pcg_iterate_once(fsl_ctx *ctx) {
  if (!ctx->pcg_ready) { ctx->err = "pcg not prepared"; return FSL_ERR_STATE; }
  if (ctx->n_fluid == 0 || ctx->pcg_host->done) return ctx->pcg_host->status;
  const int k = ctx->pcg_host->it;
  dispatch(ctx->pcg_precision, [&](auto K) {
    decltype(K)::pass_a(ctx, k);
    decltype(K)::pass_b(ctx, k);
  });
  if (ctx->pcg_precision == FSL_PCG_MIXED && (k + 1) % ctx->pcg_refresh == 0)
    launch_refresh(ctx, k, 0);
  dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::stage_scalars(ctx, k); });
  FSL_HIP(hipGetLastError());
  int rc = sync_state(ctx);
  if (rc) return rc;
  return ctx->pcg_host->status;
}

int pcg_run(fsl_ctx *ctx, const fsl_solve_params *prm, fsl_solve_stats *st) {
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
  while (!ctx->pcg_host->done) {
    const int k0 = k;
    for (int b = 0; b < nb; ++b, ++k) {
      hipEvent_t *e = pev + 4 * b;
      const bool t = ctx->timing;
      dispatch(ctx->pcg_precision, [&](auto K) {
        decltype(K)::pass_a(ctx, k, t ? e[0] : nullptr, t ? e[1] : nullptr);
      });
      dispatch(ctx->pcg_precision, [&](auto K) {
        decltype(K)::pass_b(ctx, k, t ? e[2] : nullptr, t ? e[3] : nullptr);
      });
      if (mixed && (k + 1) % ctx->pcg_refresh == 0) launch_refresh(ctx, k, 0);
    }
    // The stopping test of iteration k runs at the start of A(k).
    dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::pass_a(ctx, k); });
    FSL_HIP(hipGetLastError());
    rc = sync_state(ctx);
    if (rc) return rc;
    const PcgState *h = ctx->pcg_host;
    if (ctx->timing) {
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
      dispatch(ctx->pcg_precision, [&](auto K) { decltype(K)::pass_b(ctx, k); });
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
  if (ctx->pcg_host->status) return ctx->pcg_host->status;
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

int pcg_finish(fsl_ctx *ctx, fsl_solve_stats *st) {
  if (!ctx->pcg_ready) { ctx->err = "pcg not prepared"; return FSL_ERR_STATE; }
  if (ctx->n_act > 0) {
    dispatch(ctx->pcg_precision, [&](auto k) { decltype(k)::finish(ctx); });
    FSL_HIP(hipGetLastError());
  }
  double rr_true = 0.0;
  int rc = pcg_true_residual(ctx, &rr_true);
  if (rc) return rc;
  if (st) {
    st->iterations = ctx->pcg_host->it;
    st->rr = ctx->pcg_host->rr;
    st->bnorm = sqrt(ctx->pcg_host->bnorm2);
    st->true_rel_residual = st->bnorm > 0 ? sqrt(rr_true) / st->bnorm : sqrt(rr_true);
    st->n_fluid_leaves = int(ctx->n_fluid);
  }
  return ctx->pcg_host->status;
}

}  // namespace fsl
