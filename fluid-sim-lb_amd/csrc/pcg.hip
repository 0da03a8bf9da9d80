// Jacobi-preconditioned CG for the pressure Poisson system, matrix-free.
//
// The reference assembles an explicit 7-point SparseStencilMatrix and runs
// PCG in three stages per iteration with three scalar all-reduces
// (solver.cc:262-366, solver_driver.cc:318-343). Here the operator is applied
// from the 1-byte solver label of each cell (marker | diag << 2): row c has
// diag(c) on the diagonal and -1 for every FLUID neighbour (solver.cc:
// 214-231). One iteration is two passes over the fluid leaves:
//   A: p = z + beta p (z = r / diag), Ap, p.Ap        (reads r, p, label)
//   B: x += alpha p, r -= alpha Ap, r.r, r.z          (reads x, p, r, Ap, label)
// with alpha and beta computed on the device by the last workgroup of each
// pass, so the host only polls convergence every `batch` iterations. Each
// workgroup owns one 8^3 leaf and stages p of the leaf plus its six face
// halos in LDS; halo values of p are recomputed from r and the previous p, so
// p is double-buffered.
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
#include <algorithm>
#include <type_traits>

#include "device_util.h"
#include "fsl_internal.h"

namespace fsl {

namespace {

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

// Halo cell handled by thread t < 384: face t/64 (x-, x+, y-, y+, z-, z+).
__device__ inline void halo_cell(int t, int &x, int &y, int &z) {
  const int f = t >> 6, a = (t >> 3) & 7, b = t & 7;
  const int s = (f & 1) ? 8 : -1;
  if (f < 2) { x = s; y = a; z = b; }
  else if (f < 4) { x = a; y = s; z = b; }
  else { x = a; y = b; z = s; }
}

// Apply the operator at local cell (x,y,z) from the LDS tile.
template <typename V, bool ORD>
__device__ inline V stencil(const V *sp, const uint8_t *sf, int x, int y, int z, int diag) {
#pragma clang fp contract(off)
  const int c = tix(x, y, z);
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
    return V(diag) * sp[c] - nb;
  }
}

template <typename T>
__device__ inline T ld_coherent(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Block-reduce K values, publish this workgroup's partials, and return true
// in the last workgroup to finish, which receives the totals (summed over
// workgroups in list order, so the result is deterministic).
template <typename A, int K>
__device__ bool reduce_last(A vals[K], A *partials, int64_t w, int64_t nwork, unsigned *counter,
                            double out[K]) {
  __shared__ A s_red[K][16];
  __shared__ int s_last;
#pragma unroll
  for (int k = 0; k < K; ++k) vals[k] = block_sum<A>(vals[k], s_red[k]);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) partials[w * K + k] = vals[k];
    __threadfence();
    const unsigned prev = atomicAdd(counter, 1u);
    s_last = prev == unsigned(nwork - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if constexpr (sizeof(A) == 8) {
      double s = 0.0, c = 0.0;
      for (int64_t q = threadIdx.x; q < nwork; q += blockDim.x)
        neumaier_add(s, c, double(ld_coherent(&partials[q * K + k])));
      out[k] = double(block_sum<double>(s + c, reinterpret_cast<double *>(s_red[k])));
    } else {
      A s = A(0);
      for (int64_t q = threadIdx.x; q < nwork; q += blockDim.x) s += ld_coherent(&partials[q * K + k]);
      out[k] = double(block_sum<A>(s, s_red[k]));
    }
  }
  if (threadIdx.x == 0) *counter = 0u;
  return true;
}

__device__ inline bool converged(const PcgState *st, double rr) {
  return st->stop_mode == FSL_STOP_ABS_RR ? rr < st->thr : sqrt(rr) < st->thr;
}

// PreparePCGSolver (solver.cc:262-286): x = x0 (the pressure grid), r = b - A x,
// z = M^-1 r, rz, rr; also ||b||^2 for the relative stopping rule.
template <typename V, typename A, bool ORD, bool MIXED>
__global__ __launch_bounds__(512) void k_prepare(Dom d, const int32_t *flist, int64_t nf,
                                                 const uint8_t *lab, const float *div,
                                                 const float *pres, int warm, V *x, V *r,
                                                 double *x64, A *partials, PcgState *st) {
#pragma clang fp contract(off)
  using W = typename std::conditional<MIXED, double, V>::type;  // residual arithmetic
  __shared__ W s_x[1000];
  __shared__ uint8_t s_f[1000];
  const int64_t w = xcd_work(nf);
  if (w >= nf) return;
  const int64_t leaf = flist[w];
  int o[3];
  leaf_origin(d, leaf, o);
  const int t = threadIdx.x, lx = t >> 6, ly = (t >> 3) & 7, lz = t & 7;
  const int64_t v = (leaf << 9) | t;
  const uint8_t l = lab[v];
  const bool f = is_fluid(l);
  const float x0 = warm && f ? pres[v] : 0.0f;
  s_x[tix(lx, ly, lz)] = W(x0);
  s_f[tix(lx, ly, lz)] = f;
  if (t < 384) {
    int hx, hy, hz;
    halo_cell(t, hx, hy, hz);
    const int64_t hv = vox_of(d, o[0] + hx, o[1] + hy, o[2] + hz);
    const bool hf = is_fluid(lab[hv]);
    s_x[tix(hx, hy, hz)] = warm && hf ? W(pres[hv]) : W(0);
    s_f[tix(hx, hy, hz)] = hf;
  }
  __syncthreads();
  A vals[3] = {A(0), A(0), A(0)};
  if (f) {
    const W ax = stencil<W, ORD>(s_x, s_f, lx, ly, lz, l >> 2);
    const W bw = W(div[v]);
    const V rv = V(bw - ax);
    if (MIXED) { x64[v] = double(x0); x[v] = V(0); }
    else x[v] = V(x0);
    r[v] = rv;
    const V z = rv * dinv<V>(l >> 2);
    vals[0] = A(rv) * A(rv);
    vals[1] = A(rv) * A(z);
    vals[2] = A(div[v]) * A(div[v]);
  }
  double tot[3];
  if (reduce_last<A, 3>(vals, partials, w, nf, &st->cnt_x, tot) && threadIdx.x == 0) {
    st->rr = tot[0];
    st->rz = tot[1];
    st->bnorm2 = tot[2];
    if (st->stop_mode == FSL_STOP_REL) st->thr = st->thr * sqrt(tot[2]);
    st->it = 0;
    st->status = isnan(tot[0]) ? FSL_ERR_NAN : 0;
    st->done = (st->status != 0 || converged(st, tot[0]) || st->max_iter <= 0) ? 1 : 0;
    st->pAp = 0; st->alpha = 0; st->beta = 0; st->rz_old = 0;
  }
}

// Pass A: p = z + beta p (p = z on the first iteration); Ap = A p; p.Ap;
// alpha = rz / pAp (ComputeLocal1 + ComputeLocal3, solver.cc:296-309, 355-366).
template <typename V, typename A, bool ORD>
__global__ __launch_bounds__(512) void k_pass_a(Dom d, const int32_t *flist, int64_t nf,
                                                const uint8_t *lab, const V *r, V *p0, V *p1,
                                                V *Ap, A *partials, PcgState *st) {
#pragma clang fp contract(off)
  __shared__ V s_p[1000];
  __shared__ uint8_t s_f[1000];
  if (st->done) return;
  const int64_t w = xcd_work(nf);
  if (w >= nf) return;
  const bool first = st->it == 0;
  const V beta = V(st->beta);
  const V *pin = st->pad ? p1 : p0;
  V *pout = st->pad ? p0 : p1;
  const int64_t leaf = flist[w];
  int o[3];
  leaf_origin(d, leaf, o);
  const int t = threadIdx.x, lx = t >> 6, ly = (t >> 3) & 7, lz = t & 7;
  const int64_t v = (leaf << 9) | t;
  const uint8_t l = lab[v];
  const bool f = is_fluid(l);
  V pn = V(0);
  if (f) {
    const V z = r[v] * dinv<V>(l >> 2);
    if (first) pn = z;
    else { const V m = beta * pin[v]; pn = z + m; }
  }
  s_p[tix(lx, ly, lz)] = pn;
  s_f[tix(lx, ly, lz)] = f;
  if (t < 384) {
    int hx, hy, hz;
    halo_cell(t, hx, hy, hz);
    const int64_t hv = vox_of(d, o[0] + hx, o[1] + hy, o[2] + hz);
    const uint8_t hl = lab[hv];
    const bool hf = is_fluid(hl);
    V hp = V(0);
    if (hf) {
      const V z = r[hv] * dinv<V>(hl >> 2);
      if (first) hp = z;
      else { const V m = beta * pin[hv]; hp = z + m; }
    }
    s_p[tix(hx, hy, hz)] = hp;
    s_f[tix(hx, hy, hz)] = hf;
  }
  __syncthreads();
  V ap = V(0);
  if (f) ap = stencil<V, ORD>(s_p, s_f, lx, ly, lz, l >> 2);
  pout[v] = pn;
  Ap[v] = ap;
  A vals[1] = {f ? A(pn) * A(ap) : A(0)};
  double tot[1];
  if (reduce_last<A, 1>(vals, partials, w, nf, &st->cnt_a, tot) && threadIdx.x == 0) {
    const A pAp = A(tot[0]);
    const A alpha = A(st->rz) / pAp;
    st->pAp = double(pAp);
    st->alpha = double(alpha);
    if (!(double(alpha) <= 1e5)) {  // CHECK_LE(alpha, LARGE_FLOAT), solver.cc:333
      st->status = FSL_ERR_ALPHA;
      st->done = 1;
    }
  }
}

// Pass B: x += alpha p, r -= alpha Ap, z = M^-1 r, rr, rz; beta = rz/rz_old
// (ComputeLocal2 + the While test, solver.cc:323-346, solver_driver.cc:330).
template <typename V, typename A, bool ORD>
__global__ __launch_bounds__(512) void k_pass_b(const int32_t *flist, int64_t nf,
                                                const uint8_t *lab, const V *p0, const V *p1,
                                                const V *Ap, V *x, V *r, A *partials,
                                                PcgState *st) {
#pragma clang fp contract(off)
  if (st->done) return;
  const int64_t w = xcd_work(nf);
  if (w >= nf) return;
  const V alpha = V(st->alpha);
  const V *p = st->pad ? p0 : p1;  // the buffer pass A wrote
  const int64_t v = (int64_t(flist[w]) << 9) | threadIdx.x;
  const uint8_t l = lab[v];
  A vals[2] = {A(0), A(0)};
  if (is_fluid(l)) {
    V xn, rn;
    if (ORD) {
      const V m = alpha * p[v];
      xn = m + x[v];
      const V m2 = (-alpha) * Ap[v];
      rn = m2 + r[v];
    } else {
      xn = __builtin_fma(alpha, p[v], x[v]);
      rn = __builtin_fma(-alpha, Ap[v], r[v]);
    }
    x[v] = xn;
    r[v] = rn;
    const V z = rn * dinv<V>(l >> 2);
    vals[0] = A(rn) * A(rn);
    vals[1] = A(rn) * A(z);
  }
  double tot[2];
  if (reduce_last<A, 2>(vals, partials, w, nf, &st->cnt_b, tot) && threadIdx.x == 0) {
    const A rz_old = A(st->rz), rz = A(tot[1]);
    const A beta = rz / rz_old;
    st->rr = tot[0];
    st->rz_old = double(rz_old);
    st->rz = double(rz);
    st->beta = double(beta);
    st->it += 1;
    st->pad ^= 1;
    if (isnan(tot[0])) st->status = FSL_ERR_NAN;
    else if (!(double(beta) <= 1e5)) st->status = FSL_ERR_BETA;  // solver.cc:361
    st->done = (st->status != 0 || converged(st, tot[0]) || st->it >= st->max_iter) ? 1 : 0;
  }
}

// MIXED refresh, part 1: fold the float iterate into the double solution.
__global__ __launch_bounds__(512) void k_refresh_x(const int32_t *flist, int64_t nf,
                                                   const uint8_t *lab, float *x, double *x64,
                                                   const PcgState *st, int force) {
  if (!force && st->done) return;
  if (st->status) return;
  const int64_t w = xcd_work(nf);
  if (w >= nf) return;
  const int64_t v = (int64_t(flist[w]) << 9) | threadIdx.x;
  if (!is_fluid(lab[v])) return;
  x64[v] += double(x[v]);
  x[v] = 0.0f;
}

// MIXED refresh, part 2: r = b - A x64 in double, stored as float; rr and rz
// replace the recursive values and beta is recomputed from the true rz.
__global__ __launch_bounds__(512) void k_refresh_r(Dom d, const int32_t *flist, int64_t nf,
                                                   const uint8_t *lab, const float *div,
                                                   const double *x64, float *r,
                                                   double *partials, PcgState *st, int force) {
  __shared__ double s_x[1000];
  __shared__ uint8_t s_f[1000];
  if (!force && st->done) return;
  if (st->status) return;
  const int64_t w = xcd_work(nf);
  if (w >= nf) return;
  const int64_t leaf = flist[w];
  int o[3];
  leaf_origin(d, leaf, o);
  const int t = threadIdx.x, lx = t >> 6, ly = (t >> 3) & 7, lz = t & 7;
  const int64_t v = (leaf << 9) | t;
  const uint8_t l = lab[v];
  const bool f = is_fluid(l);
  s_x[tix(lx, ly, lz)] = f ? x64[v] : 0.0;
  s_f[tix(lx, ly, lz)] = f;
  if (t < 384) {
    int hx, hy, hz;
    halo_cell(t, hx, hy, hz);
    const int64_t hv = vox_of(d, o[0] + hx, o[1] + hy, o[2] + hz);
    const bool hf = is_fluid(lab[hv]);
    s_x[tix(hx, hy, hz)] = hf ? x64[hv] : 0.0;
    s_f[tix(hx, hy, hz)] = hf;
  }
  __syncthreads();
  double vals[2] = {0.0, 0.0};
  if (f) {
    const double rt = double(div[v]) - stencil<double, false>(s_x, s_f, lx, ly, lz, l >> 2);
    const float rf = float(rt);
    r[v] = rf;
    vals[0] = rt * rt;
    vals[1] = double(rf) * double(rf * dinv<float>(l >> 2));
  }
  double tot[2];
  if (reduce_last<double, 2>(vals, partials, w, nf, &st->cnt_r, tot) && threadIdx.x == 0) {
    st->rr_true = tot[0];
    st->rr = tot[0];
    st->rz = tot[1];
    st->beta = st->rz_old != 0.0 ? tot[1] / st->rz_old : 0.0;
    if (isnan(tot[0])) st->status = FSL_ERR_NAN;
    st->done = (st->status != 0 || converged(st, tot[0]) || st->it >= st->max_iter) ? 1 : 0;
  }
}

// CopySolution (solver.cc:93-99): the pressure grid holds the solution on
// FLUID cells and 0 elsewhere.
template <typename V, bool MIXED>
__global__ __launch_bounds__(512) void k_finish(const int32_t *list, int64_t nlist,
                                                const uint8_t *lab, const V *x,
                                                const double *x64, float *pres) {
  const int64_t w = xcd_work(nlist);
  if (w >= nlist) return;
  const int64_t v = (int64_t(list[w]) << 9) | threadIdx.x;
  float out = 0.0f;
  if (is_fluid(lab[v])) out = MIXED ? float(x64[v] + double(x[v])) : float(x[v]);
  pres[v] = out;
}

// ||b - A x||^2 in double for the solver's own solution (stats only).
template <typename V, bool MIXED>
__global__ __launch_bounds__(512) void k_true_residual(Dom d, const int32_t *flist, int64_t nf,
                                                       const uint8_t *lab, const float *div,
                                                       const V *x, const double *x64,
                                                       double *partials, PcgState *st) {
  __shared__ double s_x[1000];
  __shared__ uint8_t s_f[1000];
  const int64_t w = xcd_work(nf);
  if (w >= nf) return;
  const int64_t leaf = flist[w];
  int o[3];
  leaf_origin(d, leaf, o);
  const int t = threadIdx.x, lx = t >> 6, ly = (t >> 3) & 7, lz = t & 7;
  const int64_t v = (leaf << 9) | t;
  const uint8_t l = lab[v];
  const bool f = is_fluid(l);
  auto sol = [&](int64_t q) { return MIXED ? x64[q] + double(x[q]) : double(x[q]); };
  s_x[tix(lx, ly, lz)] = f ? sol(v) : 0.0;
  s_f[tix(lx, ly, lz)] = f;
  if (t < 384) {
    int hx, hy, hz;
    halo_cell(t, hx, hy, hz);
    const int64_t hv = vox_of(d, o[0] + hx, o[1] + hy, o[2] + hz);
    const bool hf = is_fluid(lab[hv]);
    s_x[tix(hx, hy, hz)] = hf ? sol(hv) : 0.0;
    s_f[tix(hx, hy, hz)] = hf;
  }
  __syncthreads();
  double vals[1] = {0.0};
  if (f) {
    const double rt = double(div[v]) - stencil<double, false>(s_x, s_f, lx, ly, lz, l >> 2);
    vals[0] = rt * rt;
  }
  double tot[1];
  if (reduce_last<double, 1>(vals, partials, w, nf, &st->cnt_r, tot) && threadIdx.x == 0)
    st->rr_true = tot[0];
}

template <typename V, typename A, bool ORD, bool MIXED>
struct Kernels {
  static void prepare(fsl_ctx *c, int warm) {
    k_prepare<V, A, ORD, MIXED><<<xcd_grid(c->n_fluid), 512, 0, c->stream>>>(
        c->dom, c->fluid_list, c->n_fluid, c->lab, c->div, c->pres, warm, (V *)c->x, (V *)c->r,
        c->x64, (A *)c->partials, c->pcg);
  }
  static void pass_a(fsl_ctx *c) {
    k_pass_a<V, A, ORD><<<xcd_grid(c->n_fluid), 512, 0, c->stream>>>(
        c->dom, c->fluid_list, c->n_fluid, c->lab, (const V *)c->r, (V *)c->p[0], (V *)c->p[1],
        (V *)c->Ap, (A *)c->partials, c->pcg);
  }
  static void pass_b(fsl_ctx *c) {
    k_pass_b<V, A, ORD><<<xcd_grid(c->n_fluid), 512, 0, c->stream>>>(
        c->fluid_list, c->n_fluid, c->lab, (const V *)c->p[0], (const V *)c->p[1],
        (const V *)c->Ap, (V *)c->x, (V *)c->r, (A *)c->partials, c->pcg);
  }
  static void iterate(fsl_ctx *c) {
    pass_a(c);
    pass_b(c);
  }
  static void finish(fsl_ctx *c) {
    k_finish<V, MIXED><<<xcd_grid(c->n_act), 512, 0, c->stream>>>(
        c->act_list, c->n_act, c->lab, (const V *)c->x, c->x64, c->pres);
  }
  static void true_residual(fsl_ctx *c) {
    k_true_residual<V, MIXED><<<xcd_grid(c->n_fluid), 512, 0, c->stream>>>(
        c->dom, c->fluid_list, c->n_fluid, c->lab, c->div, (const V *)c->x, c->x64,
        (double *)c->partials, c->pcg);
  }
};

void launch_refresh(fsl_ctx *c, int force) {
  k_refresh_x<<<xcd_grid(c->n_fluid), 512, 0, c->stream>>>(c->fluid_list, c->n_fluid, c->lab,
                                                            (float *)c->x, c->x64, c->pcg, force);
  k_refresh_r<<<xcd_grid(c->n_fluid), 512, 0, c->stream>>>(
      c->dom, c->fluid_list, c->n_fluid, c->lab, c->div, c->x64, (float *)c->r,
      (double *)c->partials, c->pcg, force);
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

}  // namespace

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
  FSL_HIP(hipMemcpyAsync(ctx->pcg, h, sizeof(PcgState), hipMemcpyHostToDevice, ctx->stream));
  if (ctx->n_fluid > 0) {
    dispatch(ctx->pcg_precision,
             [&](auto k) { decltype(k)::prepare(ctx, prm->warm_start ? 1 : 0); });
    FSL_HIP(hipGetLastError());
  }
  int rc = sync_state(ctx);
  if (rc) return rc;
  if (ctx->n_fluid == 0) {
    ctx->pcg_host->done = 1;
    FSL_HIP(hipMemcpyAsync(ctx->pcg, ctx->pcg_host, sizeof(PcgState), hipMemcpyHostToDevice,
                           ctx->stream));
  }
  ctx->pcg_ready = true;
  return ctx->pcg_host->status;
}

int pcg_iterate_once(fsl_ctx *ctx) {
  if (!ctx->pcg_ready) { ctx->err = "pcg not prepared"; return FSL_ERR_STATE; }
  if (ctx->n_fluid == 0) return FSL_OK;
  dispatch(ctx->pcg_precision, [&](auto k) { decltype(k)::iterate(ctx); });
  FSL_HIP(hipGetLastError());
  if (ctx->pcg_precision == FSL_PCG_MIXED && ctx->pcg_host->it + 1 > 0 &&
      (ctx->pcg_host->it + 1) % ctx->pcg_refresh == 0) {
    launch_refresh(ctx, 0);
    FSL_HIP(hipGetLastError());
  }
  int rc = sync_state(ctx);
  if (rc) return rc;
  return ctx->pcg_host->status;
}

int pcg_run(fsl_ctx *ctx, const fsl_solve_params *prm, fsl_solve_stats *st) {
  int rc = pcg_prepare(ctx, prm);
  if (rc) return rc;
  const int batch = prm->batch > 0 ? prm->batch : 16;
  hipEvent_t e0, e1;
  FSL_HIP(hipEventCreate(&e0));
  FSL_HIP(hipEventCreate(&e1));
  FSL_HIP(hipEventRecord(e0, ctx->stream));
  int it_host = ctx->pcg_host->it;
  const bool mixed = ctx->pcg_precision == FSL_PCG_MIXED;
  // With the measurement hook on, every pass is bracketed by events (at most
  // 32 iterations per batch, three events per iteration).
  const int nb = ctx->timing ? std::min(batch, 32) : batch;
  hipEvent_t *pev = ctx->tev + 32;
  while (!ctx->pcg_host->done) {
    const int it_before = ctx->pcg_host->it;
    for (int b = 0; b < nb; ++b) {
      if (ctx->timing) FSL_HIP(hipEventRecord(pev[3 * b], ctx->stream));
      dispatch(ctx->pcg_precision, [&](auto k) { decltype(k)::pass_a(ctx); });
      if (ctx->timing) FSL_HIP(hipEventRecord(pev[3 * b + 1], ctx->stream));
      dispatch(ctx->pcg_precision, [&](auto k) { decltype(k)::pass_b(ctx); });
      if (ctx->timing) FSL_HIP(hipEventRecord(pev[3 * b + 2], ctx->stream));
      ++it_host;
      if (mixed && it_host % ctx->pcg_refresh == 0) launch_refresh(ctx, 0);
    }
    FSL_HIP(hipGetLastError());
    rc = sync_state(ctx);
    if (rc) return rc;
    if (ctx->timing) {
      const int worked = std::min(ctx->pcg_host->it - it_before, nb);
      for (int b = 0; b < worked; ++b) {
        float a = 0.0f, bb = 0.0f;
        FSL_HIP(hipEventElapsedTime(&a, pev[3 * b], pev[3 * b + 1]));
        FSL_HIP(hipEventElapsedTime(&bb, pev[3 * b + 1], pev[3 * b + 2]));
        ctx->tacc[FSL_T_PASS_A] += a;
        ctx->tacc[FSL_T_PASS_B] += bb;
        ctx->tacc[FSL_T_PASS_A_N] += 1;
        ctx->tacc[FSL_T_PASS_B_N] += 1;
      }
    }
    it_host = ctx->pcg_host->it;
    if (ctx->pcg_host->status) break;
    if (ctx->pcg_host->done && mixed) {
      // Converged on the recursive residual: confirm on the true residual,
      // and continue if it does not hold yet.
      launch_refresh(ctx, 1);
      FSL_HIP(hipGetLastError());
      rc = sync_state(ctx);
      if (rc) return rc;
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
