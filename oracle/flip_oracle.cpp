// CPU oracle: a restatement of the reference FLIP step, for tests only.
//
// TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, and only as the checker (or the timed
// CPU baseline). The product path (fluid-sim-lb_amd/) never links or calls it.
//
// The reference cannot be built here (OpenVDB 4.0.1 + patch, Canary, cereal,
// glog are absent), so every stage below restates the reference's source with
// dense arrays standing in for OpenVDB trees. A dense array reproduces a VDB
// grid exactly: reads of inactive voxels return the background 0, set()
// activates, and the per-voxel "on" flag is kept because activity matters
// (AddGravity touches active faces only, flip_simulation.cc:505-515).
//
// Iteration orders that change floating-point rounding follow the reference:
//  * particle-tree traversal is VDB order (root tiles of 64^3 sorted by origin,
//    then 8^3 leaves in x-major offset order, then voxels x-major), which fixes
//    the P2G accumulation order (flip_simulation.cc:826-891);
//  * StepInGrid re-adds particles in traversal order (particles.cc:489-544);
//  * the Laplacian row dot runs over columns in ascending order, i.e. the
//    row-index order of CreateIndexAndMaskTree (solver.cc:114-151).
// Build flags: -O2 -ffp-contract=off, so no multiply-add is fused (the
// reference is built for baseline x86-64, which has no FMA).
//
// Parity unpinned (lives in un-vendored OpenVDB 4.0.1 math/ConjGradient.h):
// Vector::dot summation order (restated here as a sequential sum), the Jacobi
// apply form (restated as z = r * float(1.0/diag), OpenVDB's published
// JacobiPreconditioner), computeResidual/axpy rounding (restated as
// b - row.dot(x) and a*x + y).

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include <omp.h>

// Scalar type of grids, particles and the API: float restates the reference;
// the double build (-DORACLE_REAL=double, liboracle64.so) is the same
// algorithm in FP64, for finite-difference checks of the adjoint.
#ifndef ORACLE_REAL
#define ORACLE_REAL float
#endif
typedef ORACLE_REAL Real;

namespace {

enum CellType { AIR = 0, FLUID = 1, SOLID = 2 };

// Dense stand-in for an OpenVDB tree with background 0 over [-P, n+P-1].
template <typename T>
struct Dense {
  static const int P = 2;
  int nx = 0, ny = 0, nz = 0;
  std::vector<T> val;
  std::vector<uint8_t> on;
  void init(int x, int y, int z) {
    nx = x; ny = y; nz = z;
    size_t s = size_t(x + 2 * P) * (y + 2 * P) * (z + 2 * P);
    val.assign(s, T(0));
    on.assign(s, 0);
  }
  bool inside(int i, int j, int k) const {
    return i >= -P && j >= -P && k >= -P && i < nx + P && j < ny + P && k < nz + P;
  }
  size_t idx(int i, int j, int k) const {
    return (size_t(i + P) * (ny + 2 * P) + (j + P)) * (nz + 2 * P) + (k + P);
  }
  T get(int i, int j, int k) const { return inside(i, j, k) ? val[idx(i, j, k)] : T(0); }
  bool isOn(int i, int j, int k) const { return inside(i, j, k) && on[idx(i, j, k)]; }
  void set(int i, int j, int k, T v) {
    if (!inside(i, j, k)) { fprintf(stderr, "oracle: set out of range\n"); abort(); }
    size_t s = idx(i, j, k);
    val[s] = v; on[s] = 1;
  }
  void clear() {
    std::fill(val.begin(), val.end(), T(0));
    std::fill(on.begin(), on.end(), 0);
  }
};

// Vec3 grid: one activity flag per voxel, three components.
struct DenseVec {
  Dense<Real> c[3];
  void init(int x, int y, int z) { for (auto &g : c) g.init(x, y, z); }
  bool isOn(int i, int j, int k) const { return c[0].isOn(i, j, k); }
  void get(int i, int j, int k, Real v[3]) const {
    for (int d = 0; d < 3; ++d) v[d] = c[d].get(i, j, k);
  }
  void set(int i, int j, int k, const Real v[3]) {
    for (int d = 0; d < 3; ++d) c[d].set(i, j, k, v[d]);
  }
  void clear() { for (auto &g : c) g.clear(); }
};

struct Particle {
  Real pos[3], vel[3], gv[3];
  int cell[3];
};

// VDB traversal key for Tree3<T,3,3>: 64^3 internal nodes ordered by origin
// (std::map<Coord> in the root), then leaf offset, then voxel offset.
inline uint64_t vdb_key(int i, int j, int k) {
  uint64_t ri = uint64_t((i >> 6) + 1024), rj = uint64_t((j >> 6) + 1024),
           rk = uint64_t((k >> 6) + 1024);
  uint64_t leaf = uint64_t(((i >> 3) & 7) << 6 | ((j >> 3) & 7) << 3 | ((k >> 3) & 7));
  uint64_t vox = uint64_t((i & 7) << 6 | (j & 7) << 3 | (k & 7));
  return (((ri * 2048 + rj) * 2048 + rk) << 18) | (leaf << 9) | vox;
}
inline uint64_t leaf_of_key(uint64_t key) { return key >> 9; }

struct Oracle {
  int n[3];
  Real gravity = 9.8f;  // already scaled by N_y/128 (flip_app_sim.cc:74)
  Real fx = 0, fy = 1, fz = 0;
  Real flip = 0.95f;
  Dense<int32_t> solid, marker;
  DenseVec v, vu, w;   // grid_velocity_, grid_velocity_update_, grid_weight_
  Dense<Real> div, pressure;
  std::vector<Particle> parts;   // kept in VDB traversal order
  // Solver outputs.
  int n_rows = 0, n_cols = 0, iters = 0;
  double rr_final = 0;
  Dense<int32_t> idx;            // solver row index (solver.cc:114-151)
  std::vector<double> rr_hist;
  // Per-iteration scalars of the last solve, in the solver's precision
  // (widened): rz after PreparePCGSolver and after each ComputeLocal2, and
  // pAp, alpha, beta of each iteration.
  std::vector<double> rz_hist, pap_hist, alpha_hist, beta_hist;

  void init(int x, int y, int z) {
    n[0] = x; n[1] = y; n[2] = z;
    solid.init(x, y, z); marker.init(x, y, z);
    v.init(x, y, z); vu.init(x, y, z); w.init(x, y, z);
    div.init(x, y, z); pressure.init(x, y, z); idx.init(x, y, z);
  }
  bool in_bbox(int i, int j, int k) const {
    return i >= 0 && j >= 0 && k >= 0 && i < n[0] && j < n[1] && k < n[2];
  }
  // Stable sort of the particles by VDB traversal key. Cells in [0, n] (all
  // a sort meets after StepInGrid's clamps) are ranked densely (root tile,
  // then leaf, then voxel, the order of vdb_key) and counting-sorted; any
  // other cell falls back to a comparison sort. Both keep equal keys in
  // buffer order.
  void sort_particles() {
    const int64_t nn[3] = {n[0] / 64 + 1, n[1] / 64 + 1, n[2] / 64 + 1};
    bool dense = true;
    for (const Particle &p : parts)
      for (int d = 0; d < 3; ++d)
        if (p.cell[d] < 0 || p.cell[d] > n[d]) dense = false;
    if (!dense) {
      std::stable_sort(parts.begin(), parts.end(), [](const Particle &a, const Particle &b) {
        return vdb_key(a.cell[0], a.cell[1], a.cell[2]) < vdb_key(b.cell[0], b.cell[1], b.cell[2]);
      });
      return;
    }
    auto rank = [&](const int c[3]) {
      const int64_t root = ((int64_t(c[0] >> 6) * nn[1]) + (c[1] >> 6)) * nn[2] + (c[2] >> 6);
      const int64_t leaf = ((c[0] >> 3) & 7) << 6 | ((c[1] >> 3) & 7) << 3 | ((c[2] >> 3) & 7);
      const int64_t vox = (c[0] & 7) << 6 | (c[1] & 7) << 3 | (c[2] & 7);
      return (root << 18) | (leaf << 9) | vox;
    };
    const size_t nb = size_t(nn[0] * nn[1] * nn[2]) << 18;
    std::vector<uint32_t> start(nb + 1, 0);
    std::vector<uint32_t> key(parts.size());
    for (size_t q = 0; q < parts.size(); ++q) {
      key[q] = uint32_t(rank(parts[q].cell));
      ++start[key[q] + 1];
    }
    for (size_t b = 0; b < nb; ++b) start[b + 1] += start[b];
    std::vector<Particle> out(parts.size());
    for (size_t q = 0; q < parts.size(); ++q) out[start[key[q]]++] = parts[q];
    parts.swap(out);
  }

  // FlipSimulation::SetBoxBoundary (flip_simulation.cc:255-308), one partition.
  void set_box_boundary(bool tank) {
    int s[3] = {0, 0, 0}, e[3] = {n[0] - 1, n[1] - 1, n[2] - 1};
    for (int j = s[1] - 1; j <= e[1] + 1; ++j)
      for (int k = s[2] - 1; k <= e[2] + 1; ++k) { solid.set(-1, j, k, 1); solid.set(n[0], j, k, 1); }
    for (int i = s[0] - 1; i <= e[0] + 1; ++i)
      for (int k = s[2] - 1; k <= e[2] + 1; ++k) {
        solid.set(i, -1, k, 1);
        if (!tank) solid.set(i, n[1], k, 1);  // SetTankBoundary skips upper y (:360)
      }
    for (int i = s[0] - 1; i <= e[0] + 1; ++i)
      for (int j = s[1] - 1; j <= e[1] + 1; ++j) { solid.set(i, j, -1, 1); solid.set(i, j, n[2], 1); }
  }
  // AddSolidCuboid (flip_simulation.cc:315-324): fill box ∩ bbox.expand(1).
  void add_solid_box(const int lo[3], const int hi[3]) {
    int a[3], b[3];
    for (int d = 0; d < 3; ++d) { a[d] = std::max(lo[d], -1); b[d] = std::min(hi[d], n[d]); }
    for (int i = a[0]; i <= b[0]; ++i)
      for (int j = a[1]; j <= b[1]; ++j)
        for (int k = a[2]; k <= b[2]; ++k) solid.set(i, j, k, 1);
  }

  // CFL (flip_simulation.cc:459-488).
  Real cfl() const {
    int e[3] = {n[0] - 1, n[1] - 1, n[2] - 1};
    Real v_inf = 0.0f;
    for (int i = 0; i <= e[0]; ++i)
      for (int j = 0; j <= e[1]; ++j)
        for (int k = 0; k <= e[2]; ++k) {
          if (!v.isOn(i, j, k)) continue;
          if (j != e[1] && k != e[2]) v_inf = std::fmax(v_inf, std::fabs(v.c[0].get(i, j, k)));
          if (i != e[0] && k != e[2]) v_inf = std::fmax(v_inf, std::fabs(v.c[1].get(i, j, k)));
          if (i != e[0] && j != e[1]) v_inf = std::fmax(v_inf, std::fabs(v.c[2].get(i, j, k)));
        }
    Real vv = Real(std::fmax(std::sqrt(0.5 * double(gravity)), double(v_inf)));
    if (vv < 1e-2) vv = 1e-2f;
    return Real(1.0 / double(vv));
  }

  // Particles::StepInGrid<VectorGridT,true,1,1> (particles.cc:478-546).
  void step_in_grid(Real dt) {
    for (auto &p : parts) {
      for (int d = 0; d < 3; ++d) {
        Real m = p.gv[d] * dt;
        p.pos[d] = p.pos[d] + m;
      }
      int id[3];
      for (int d = 0; d < 3; ++d) id[d] = int(std::floor(p.pos[d]));
      for (int d = 0; d < 3; ++d)
        if (id[d] < 0) { id[d] = 0; p.pos[d] = 0; }
      for (int d = 0; d < 3; ++d)
        if (id[d] >= n[d]) { id[d] = n[d]; p.pos[d] = Real(n[d]); }
      for (int d = 0; d < 3; ++d) p.cell[d] = id[d];
    }
    sort_particles();  // AddParticleToData in buffer order == stable sort
  }
  // Particles::DeleteParticlesInSolids (particles.cc:202-219).
  void delete_in_solids() {
    std::vector<Particle> keep;
    keep.reserve(parts.size());
    for (auto &p : parts)
      if (!solid.isOn(p.cell[0], p.cell[1], p.cell[2])) keep.push_back(p);
    parts.swap(keep);
  }
  // Particles::DeleteOutsideParticles (particles.cc:273-293).
  void delete_outside() {
    std::vector<Particle> keep;
    keep.reserve(parts.size());
    for (auto &p : parts)
      if (in_bbox(p.cell[0], p.cell[1], p.cell[2])) keep.push_back(p);
    parts.swap(keep);
  }
  // ClearGridData (flip_simulation.cc:379-389); solids are kept.
  void clear_grid() {
    v.clear(); vu.clear(); w.clear(); div.clear(); marker.clear(); pressure.clear();
  }

  // Staggered lower-sample position for component D (flip_simulation.cc:936-973).
  template <int D>
  static void lower_sample(const Real pos[3], const int c[3], Real lower[3], int off[3]) {
    lower[0] = Real(c[0]); lower[1] = Real(c[1]); lower[2] = Real(c[2]);
    off[0] = off[1] = off[2] = 0;
    for (int a = 0; a < 3; ++a) {
      if (a == D) continue;
      lower[a] = lower[a] + 0.5f;
      if (pos[a] < lower[a]) { lower[a] = lower[a] - 1.0f; off[a] = -1; }
    }
  }

  // TransferParticlesToGrid (flip_simulation.cc:808-923).
  void p2g() {
    const int B = 10, S = B * B * B;
    std::vector<Real> vb(3 * S), wb(3 * S);
    std::vector<char> mb(3 * S);
    size_t p0 = 0;
    while (p0 < parts.size()) {
      uint64_t lk = leaf_of_key(vdb_key(parts[p0].cell[0], parts[p0].cell[1], parts[p0].cell[2]));
      size_t p1 = p0;
      while (p1 < parts.size() &&
             leaf_of_key(vdb_key(parts[p1].cell[0], parts[p1].cell[1], parts[p1].cell[2])) == lk)
        ++p1;
      int org[3] = {(parts[p0].cell[0] & ~7) - 1, (parts[p0].cell[1] & ~7) - 1,
                    (parts[p0].cell[2] & ~7) - 1};
      std::fill(vb.begin(), vb.end(), 0.0f);
      std::fill(wb.begin(), wb.end(), 0.0f);
      std::fill(mb.begin(), mb.end(), 0);
      for (size_t q = p0; q < p1; ++q) {
        const Particle &p = parts[q];
        const int *c = p.cell;
        if (q == p0 || std::memcmp(parts[q - 1].cell, c, sizeof(int) * 3) != 0)
          if (in_bbox(c[0], c[1], c[2])) marker.set(c[0], c[1], c[2], FLUID);
        int lc[3] = {c[0] - org[0], c[1] - org[1], c[2] - org[2]};
        scatter<0>(p, lc, &vb[0], &wb[0], &mb[0]);
        scatter<1>(p, lc, &vb[S], &wb[S], &mb[S]);
        scatter<2>(p, lc, &vb[2 * S], &wb[2 * S], &mb[2 * S]);
      }
      for (int i = 0; i < B; ++i)
        for (int j = 0; j < B; ++j)
          for (int k = 0; k < B; ++k) {
            int o = (i * B + j) * B + k;
            if (!mb[o] && !mb[o + S] && !mb[o + 2 * S]) continue;
            int x = org[0] + i, y = org[1] + j, z = org[2] + k;
            Real vv[3], ww[3];
            v.get(x, y, z, vv);
            for (int d = 0; d < 3; ++d) vv[d] = vv[d] + vb[o + d * S];
            v.set(x, y, z, vv);
            w.get(x, y, z, ww);
            for (int d = 0; d < 3; ++d) ww[d] = ww[d] + wb[o + d * S];
            w.set(x, y, z, ww);
          }
      p0 = p1;
    }
    // Normalize (flip_simulation.cc:900-922).
    for (size_t s = 0; s < v.c[0].on.size(); ++s) {
      if (!v.c[0].on[s]) continue;
      for (int d = 0; d < 3; ++d) {
        Real ws = w.c[d].val[s];
        if (double(ws) < 1e-5) v.c[d].val[s] = 0;  // THRESH is a double literal
        else v.c[d].val[s] = v.c[d].val[s] / ws;
      }
    }
  }
  // ScatterWeight<D> (flip_simulation.cc:925-1009).
  template <int D>
  void scatter(const Particle &p, const int lc[3], Real *vb, Real *wb, char *mb) const {
    Real lower[3]; int off[3];
    lower_sample<D>(p.pos, p.cell, lower, off);
    int st[3] = {lc[0] + off[0], lc[1] + off[1], lc[2] + off[2]};
    Real x = p.pos[0] - lower[0], y = p.pos[1] - lower[1], z = p.pos[2] - lower[2];
    Real xy = x * y, yz = y * z, xz = z * x, xyz = xy * z;
    Real wt[2][2][2];
    wt[1][1][1] = xyz;
    wt[1][1][0] = xy - xyz;
    wt[1][0][1] = xz - xyz;
    wt[1][0][0] = x - xz - wt[1][1][0];
    wt[0][1][1] = yz - xyz;
    wt[0][1][0] = y - yz - wt[1][1][0];
    wt[0][0][1] = z - yz - wt[1][0][1];
    wt[0][0][0] = 1 - x - y + xy - wt[0][0][1];
    for (int di = 0; di < 2; ++di)
      for (int dj = 0; dj < 2; ++dj)
        for (int dk = 0; dk < 2; ++dk) {
          int o = ((st[0] + di) * 10 + st[1] + dj) * 10 + st[2] + dk;
          Real m = wt[di][dj][dk] * p.vel[D];
          vb[o] = vb[o] + m;
          wb[o] = wb[o] + wt[di][dj][dk];
          mb[o] = 1;
        }
  }

  // ApplyBoundaryConditions (flip_simulation.cc:656-713).
  void apply_bc() {
    for (int i = -Dense<int>::P; i < n[0] + Dense<int>::P; ++i)
      for (int j = -Dense<int>::P; j < n[1] + Dense<int>::P; ++j)
        for (int k = -Dense<int>::P; k < n[2] + Dense<int>::P; ++k) {
          if (!solid.isOn(i, j, k)) continue;
          marker.set(i, j, k, SOLID);
          const int nb[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
          for (int d = 0; d < 3; ++d) {  // lower faces live at c
            if (solid.get(i - nb[d][0], j - nb[d][1], k - nb[d][2]) == 0 && v.isOn(i, j, k))
              v.c[d].set(i, j, k, 0.0f);
          }
          for (int d = 0; d < 3; ++d) {  // upper faces live at c + e_d
            int a = i + nb[d][0], b = j + nb[d][1], c = k + nb[d][2];
            if (solid.get(a, b, c) == 0 && v.isOn(a, b, c)) v.c[d].set(a, b, c, 0.0f);
          }
        }
  }
  // SaveGridVelocities (flip_simulation.cc:490-496).
  void save_velocities() {
    for (size_t s = 0; s < v.c[0].on.size(); ++s)
      if (v.c[0].on[s])
        for (int d = 0; d < 3; ++d) { vu.c[d].val[s] = v.c[d].val[s]; vu.c[d].on[s] = 1; }
  }
  // AddGravity (flip_simulation.cc:505-515).
  void add_gravity(Real dt) {
    for (size_t s = 0; s < v.c[0].on.size(); ++s)
      if (v.c[0].on[s]) {
        Real a = dt * fx, b = dt * fy, c = dt * fz;
        v.c[0].val[s] = v.c[0].val[s] + a;
        v.c[1].val[s] = v.c[1].val[s] - b;
        v.c[2].val[s] = v.c[2].val[s] + c;
      }
  }
  // ReinitializePhiSimple touches only phi, which never feeds velocity or
  // pressure; it is not restated.

  // ComputeDivergence (flip_simulation.cc:715-736).
  void divergence() {
    for (int i = 0; i < n[0]; ++i)
      for (int j = 0; j < n[1]; ++j)
        for (int k = 0; k < n[2]; ++k)
          if (marker.get(i, j, k) == FLUID) {
            Real d = 0;
            d += v.c[0].get(i + 1, j, k) - v.c[0].get(i, j, k);
            d += v.c[1].get(i, j + 1, k) - v.c[1].get(i, j, k);
            d += v.c[2].get(i, j, k + 1) - v.c[2].get(i, j, k);
            div.set(i, j, k, d);
          }
  }

  // UpdateVelocityFromPressure (flip_simulation.cc:738-778).
  void pressure_update() {
    for (int i = 0; i <= n[0]; ++i)
      for (int j = 0; j <= n[1]; ++j)
        for (int k = 0; k <= n[2]; ++k) {
          Real vv[3];
          v.get(i, j, k, vv);
          if (marker.get(i, j, k) == SOLID) continue;
          bool update = false;
          const Real pc = pressure.get(i, j, k);
          if ((marker.get(i - 1, j, k) == FLUID || marker.get(i, j, k) == FLUID) &&
              marker.get(i - 1, j, k) != SOLID) {
            vv[0] += (pc - pressure.get(i - 1, j, k)); update = true;
          }
          if ((marker.get(i, j - 1, k) == FLUID || marker.get(i, j, k) == FLUID) &&
              marker.get(i, j - 1, k) != SOLID) {
            vv[1] += (pc - pressure.get(i, j - 1, k)); update = true;
          }
          if ((marker.get(i, j, k - 1) == FLUID || marker.get(i, j, k) == FLUID) &&
              marker.get(i, j, k - 1) != SOLID) {
            vv[2] += (pc - pressure.get(i, j, k - 1)); update = true;
          }
          if (update) v.set(i, j, k, vv);
        }
  }
  // GetVelocityUpdate (flip_simulation.cc:780-801). Returns 0 on success, -1
  // where the reference would CHECK-fail.
  int velocity_update() {
    for (int i = 0; i <= n[0]; ++i)
      for (int j = 0; j <= n[1]; ++j)
        for (int k = 0; k <= n[2]; ++k) {
          if (v.isOn(i, j, k) && !vu.isOn(i, j, k)) return -1;
          if (v.isOn(i, j, k) && vu.isOn(i, j, k)) {
            Real a[3], b[3];
            v.get(i, j, k, a); vu.get(i, j, k, b);
            for (int d = 0; d < 3; ++d) a[d] = a[d] - b[d];
            vu.set(i, j, k, a);
          }
        }
    return 0;
  }
  // VectorGrid::ExtrapolateToBoundary (vector_grid.cc:302-381), one partition.
  static void extrapolate(DenseVec &g, const int n[3]) {
    auto copy2 = [&](int oi, int oj, int ok, int ii, int ij, int ik, int keep) {
      Real out[3], in[3];
      g.get(oi, oj, ok, out); g.get(ii, ij, ik, in);
      for (int d = 0; d < 3; ++d) if (d != keep) out[d] = in[d];
      g.set(oi, oj, ok, out);
    };
    for (int j = 0; j <= n[1]; ++j) for (int k = 0; k <= n[2]; ++k) copy2(-1, j, k, 0, j, k, 0);
    for (int j = 0; j <= n[1]; ++j) for (int k = 0; k <= n[2]; ++k) copy2(n[0], j, k, n[0] - 1, j, k, 0);
    for (int i = 0; i <= n[0]; ++i) for (int k = 0; k <= n[2]; ++k) copy2(i, -1, k, i, 0, k, 1);
    for (int i = 0; i <= n[0]; ++i) for (int k = 0; k <= n[2]; ++k) copy2(i, n[1], k, i, n[1] - 1, k, 1);
    for (int i = 0; i <= n[0]; ++i) for (int j = 0; j <= n[1]; ++j) copy2(i, j, -1, i, j, 0, 2);
    for (int i = 0; i <= n[0]; ++i) for (int j = 0; j <= n[1]; ++j) copy2(i, j, n[2], i, j, n[2] - 1, 2);
  }
  // InterpolateComponent<D> (flip_simulation.cc:1093-1166).
  template <int D>
  static Real interp(const Real pos[3], const int c[3], const int lc[3], const Real *buf) {
    Real lower[3]; int off[3];
    lower_sample<D>(pos, c, lower, off);
    int i = lc[0] + off[0], j = lc[1] + off[1], k = lc[2] + off[2];
    Real w0 = pos[0] - lower[0], w1 = pos[1] - lower[1], w2 = pos[2] - lower[2];
    auto at = [&](int a, int b, int cc) { return buf[(a * 10 + b) * 10 + cc]; };
    Real v000 = at(i, j, k), v001 = at(i, j, k + 1), v010 = at(i, j + 1, k),
          v011 = at(i, j + 1, k + 1), v100 = at(i + 1, j, k), v101 = at(i + 1, j, k + 1),
          v110 = at(i + 1, j + 1, k), v111 = at(i + 1, j + 1, k + 1);
    Real t;
    t = w2 * (v001 - v000); Real v00 = v000 + t;
    t = w2 * (v011 - v010); Real v01 = v010 + t;
    t = w2 * (v101 - v100); Real v10 = v100 + t;
    t = w2 * (v111 - v110); Real v11 = v110 + t;
    t = w1 * (v01 - v00); Real v0 = v00 + t;
    t = w1 * (v11 - v10); Real v1 = v10 + t;
    t = w0 * (v1 - v0); return v0 + t;
  }
  // UpdateParticlesFromGrid (flip_simulation.cc:1024-1091, 1178-1189).
  void g2p() {
    extrapolate(v, n);
    extrapolate(vu, n);
    const int S = 1000;
    std::vector<Real> vb(3 * S), db(3 * S);
    size_t p0 = 0;
    while (p0 < parts.size()) {
      uint64_t lk = leaf_of_key(vdb_key(parts[p0].cell[0], parts[p0].cell[1], parts[p0].cell[2]));
      size_t p1 = p0;
      while (p1 < parts.size() &&
             leaf_of_key(vdb_key(parts[p1].cell[0], parts[p1].cell[1], parts[p1].cell[2])) == lk)
        ++p1;
      int org[3] = {(parts[p0].cell[0] & ~7) - 1, (parts[p0].cell[1] & ~7) - 1,
                    (parts[p0].cell[2] & ~7) - 1};
      for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j)
          for (int k = 0; k < 10; ++k) {
            int o = (i * 10 + j) * 10 + k;
            for (int d = 0; d < 3; ++d) {
              vb[o + d * S] = v.c[d].get(org[0] + i, org[1] + j, org[2] + k);
              db[o + d * S] = vu.c[d].get(org[0] + i, org[1] + j, org[2] + k);
            }
          }
      for (size_t q = p0; q < p1; ++q) {
        Particle &p = parts[q];
        int lc[3] = {p.cell[0] - org[0], p.cell[1] - org[1], p.cell[2] - org[2]};
        Real gv[3] = {interp<0>(p.pos, p.cell, lc, &vb[0]), interp<1>(p.pos, p.cell, lc, &vb[S]),
                       interp<2>(p.pos, p.cell, lc, &vb[2 * S])};
        Real dv[3] = {interp<0>(p.pos, p.cell, lc, &db[0]), interp<1>(p.pos, p.cell, lc, &db[S]),
                       interp<2>(p.pos, p.cell, lc, &db[2 * S])};
        for (int d = 0; d < 3; ++d) {
          Real a = dv[d] + p.vel[d];
          Real b = a - gv[d];
          Real c = flip * b;
          p.vel[d] = gv[d] + c;
          p.gv[d] = gv[d];
        }
      }
      p0 = p1;
    }
  }

  // ---- PCG pressure solve: PCGDriver::Solve (solver_driver.cc:13-69) ----
  struct Csr {
    std::vector<int64_t> start;
    std::vector<int32_t> col;
    std::vector<Real> val;
  };
  // Solver::InitializeMatrixAndVectors + CreateIndexAndMaskTree +
  // CreateLaplacianMatrix (solver.cc:75-91, 114-151, 197-233).
  void assemble(Csr &A, std::vector<int32_t> &rows_ijk) {
    idx.clear();
    n_rows = 0; n_cols = 0;
    for (int i = 0; i < n[0]; ++i) for (int j = 0; j < n[1]; ++j) for (int k = 0; k < n[2]; ++k)
      if (marker.get(i, j, k) == FLUID) ++n_rows;
    for (int i = -1; i <= n[0]; ++i) for (int j = -1; j <= n[1]; ++j) for (int k = -1; k <= n[2]; ++k)
      if (marker.get(i, j, k) == FLUID) ++n_cols;
    int ui = 0, ue = n_rows;
    rows_ijk.assign(size_t(n_rows) * 3, 0);
    for (int i = -1; i <= n[0]; ++i)
      for (int j = -1; j <= n[1]; ++j)
        for (int k = -1; k <= n[2]; ++k) {
          if (!(marker.isOn(i, j, k) && marker.get(i, j, k) == FLUID)) continue;
          if (in_bbox(i, j, k)) {
            idx.set(i, j, k, ui);
            rows_ijk[3 * ui] = i; rows_ijk[3 * ui + 1] = j; rows_ijk[3 * ui + 2] = k;
            ++ui;
          } else {
            idx.set(i, j, k, ue++);
          }
        }
    A.start.assign(size_t(n_rows) + 1, 0);
    A.col.clear(); A.val.clear();
    const int off[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (int r = 0; r < n_rows; ++r) {
      int i = rows_ijk[3 * r], j = rows_ijk[3 * r + 1], k = rows_ijk[3 * r + 2];
      std::vector<std::pair<int32_t, Real>> ent;
      Real diag = 0;
      for (int q = 0; q < 6; ++q) {
        int m = marker.get(i + off[q][0], j + off[q][1], k + off[q][2]);
        if (m == FLUID || m == AIR) diag += 1;
        if (m == FLUID) ent.push_back({idx.get(i + off[q][0], j + off[q][1], k + off[q][2]), -1.0f});
      }
      ent.push_back({r, diag});
      std::sort(ent.begin(), ent.end());  // RowEditor keeps columns sorted
      for (auto &e : ent) { A.col.push_back(e.first); A.val.push_back(e.second); }
      A.start[r + 1] = int64_t(A.col.size());
    }
  }
  // Threads of the solver's row-parallel loops (SpMV, axpys, Jacobi: each row
  // is computed alone, so the results do not depend on the count) and of the
  // FP64 dot products (fixed chunks, see dot()).
  int threads = 1;

  template <typename T>
  static T row_dot(const Csr &A, int64_t r, const std::vector<T> &x) {
    T s = 0;
    for (int64_t q = A.start[r]; q < A.start[r + 1]; ++q) {
      T m = T(A.val[q]) * x[A.col[q]];
      s = s + m;
    }
    return s;
  }
  // Dot product. In FP32 (the reference's precision) a sequential sum in row
  // order. In FP64 the rows are cut into 256 fixed chunks, each summed in
  // order, and the chunk sums added in order: the result does not depend on
  // the thread count.
  template <typename T>
  T dot(const std::vector<T> &a, const std::vector<T> &b, size_t len) const {
    if (sizeof(T) == 4 || len < 4096) {
      T s = 0;
      for (size_t q = 0; q < len; ++q) { T m = a[q] * b[q]; s = s + m; }
      return s;
    }
    constexpr int kChunks = 256;
    T part[kChunks];
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int c = 0; c < kChunks; ++c) {
      const size_t q0 = len * size_t(c) / kChunks, q1 = len * size_t(c + 1) / kChunks;
      T s = 0;
      for (size_t q = q0; q < q1; ++q) { T m = a[q] * b[q]; s = s + m; }
      part[c] = s;
    }
    T s = 0;
    for (int c = 0; c < kChunks; ++c) s = s + part[c];
    return s;
  }
  // ---- Block IC(0), the IC branch of BuildPreconditioner (solver.cc:103-106).
  // Restates OpenVDB 4.0.1's published pcg::IncompleteCholeskyPreconditioner
  // (math/ConjGradient.h; OpenVDB is not vendored) on the general sparse
  // matrix, with openvdb.patch:209-224: a partition factors its own interior
  // rows only (ghost columns are skipped). Here the partitions are the tiles
  // of one global system, so an entry couples two rows only when both cells
  // lie in the same tile. The patch leaves the innermost update loop
  // unchecked for ghost columns (rows past the factor's end when a partition
  // has ghosts); those entries are skipped here as well.
  //   L = the lower triangle of A (block entries);
  //   for k ascending: L_kk < 1e-5 fails the build; L_kk = sqrt(L_kk);
  //     L_ik = A_ki / L_kk for i > k in row k of A;
  //     for j > k in row k of A, for i >= j in row j of A: L_ij -= L_ik L_jk.
  //   apply: L y = r by rows ascending, then L^T z = y by rows descending;
  //     each row's dot runs over its stored columns in ascending order (the
  //     diagonal contributes its product with the not yet written 0).
  template <typename T>
  struct Tri {
    std::vector<int64_t> start;
    std::vector<int32_t> col;  // ascending per row
    std::vector<T> val;
    int64_t find(int64_t r, int32_t c) const {
      for (int64_t q = start[r]; q < start[r + 1]; ++q)
        if (col[q] == c) return q;
      return -1;
    }
    T get(int64_t r, int32_t c) const {
      const int64_t q = find(r, c);
      return q < 0 ? T(0) : val[q];
    }
  };
  int tile_of_row(const std::vector<int32_t> &rows, int r, const int tiles[3]) const {
    int t = 0;
    for (int a = 2; a >= 0; --a) t = t * tiles[a] + rows[3 * r + a] / (n[a] / tiles[a]);
    return t;
  }
  template <typename T>
  bool ic_factor(const Csr &A, const std::vector<int32_t> &rows, const int tiles[3], Tri<T> &L,
                 Tri<T> &U) const {
    std::vector<int> tile(static_cast<size_t>(n_rows));
    for (int r = 0; r < n_rows; ++r) tile[size_t(r)] = tile_of_row(rows, r, tiles);
    auto coupled = [&](int a, int32_t b) { return b < n_rows && tile[size_t(a)] == tile[size_t(b)]; };
    L.start.assign(size_t(n_rows) + 1, 0);
    L.col.clear(); L.val.clear();
    for (int r = 0; r < n_rows; ++r) {
      for (int64_t e = A.start[r]; e < A.start[r + 1]; ++e)
        if (A.col[e] <= r && coupled(r, A.col[e])) { L.col.push_back(A.col[e]); L.val.push_back(T(A.val[e])); }
      L.start[r + 1] = int64_t(L.col.size());
    }
    bool ok = true;
    for (int k = 0; k < n_rows && ok; ++k) {
      T dk = L.get(k, k);
      if (dk < T(1.e-5f)) { ok = false; break; }
      dk = std::sqrt(dk);
      L.val[size_t(L.find(k, k))] = dk;
      for (int64_t e = A.start[k]; e < A.start[k + 1]; ++e) {
        const int32_t ii = A.col[e];
        if (ii < k + 1 || !coupled(k, ii)) continue;
        L.val[size_t(L.find(ii, k))] = T(A.val[e]) / dk;
      }
      for (int64_t e = A.start[k]; e < A.start[k + 1]; ++e) {
        const int32_t j = A.col[e];
        if (j < k + 1 || !coupled(k, j)) continue;
        const T a_jk = L.get(j, k);
        for (int64_t f = A.start[j]; f < A.start[j + 1]; ++f) {
          const int32_t i = A.col[f];
          if (i < j || !coupled(j, i)) continue;
          const int64_t q = L.find(i, j);
          T a_ij = L.val[size_t(q)];
          const T a_ik = L.get(i, k);
          const T m = a_ik * a_jk;
          a_ij = a_ij - m;
          L.val[size_t(q)] = a_ij;
        }
      }
    }
    // U = L^T, rows ascending by column.
    U.start.assign(size_t(n_rows) + 1, 0);
    for (size_t q = 0; q < L.col.size(); ++q) ++U.start[size_t(L.col[q]) + 1];
    for (int r = 0; r < n_rows; ++r) U.start[size_t(r) + 1] += U.start[size_t(r)];
    U.col.assign(L.col.size(), 0);
    U.val.assign(L.col.size(), T(0));
    std::vector<int64_t> fill(U.start.begin(), U.start.end() - 1);
    for (int r = 0; r < n_rows; ++r)
      for (int64_t q = L.start[r]; q < L.start[r + 1]; ++q) {
        const int64_t s = fill[size_t(L.col[q])]++;
        U.col[size_t(s)] = r;
        U.val[size_t(s)] = L.val[size_t(q)];
      }
    return ok;
  }
  template <typename T>
  static void ic_apply(const Tri<T> &L, const Tri<T> &U, const std::vector<T> &r,
                       std::vector<T> &z, std::vector<T> &tmp) {
    const int64_t nr = int64_t(L.start.size()) - 1;
    tmp.assign(size_t(nr), T(0));
    z.assign(size_t(nr), T(0));
    for (int64_t i = 0; i < nr; ++i) {
      T dot = 0, diag = 0;
      for (int64_t q = L.start[i]; q < L.start[i + 1]; ++q) {
        if (L.col[q] == i) diag = L.val[q];
        const T m = L.val[q] * tmp[size_t(L.col[q])];
        dot = dot + m;
      }
      tmp[size_t(i)] = (r[size_t(i)] - dot) / diag;
    }
    for (int64_t i = nr - 1; i >= 0; --i) {
      T dot = 0, diag = 0;
      for (int64_t q = U.start[i]; q < U.start[i + 1]; ++q) {
        if (U.col[q] == i) diag = U.val[q];
        const T m = U.val[q] * z[size_t(U.col[q])];
        dot = dot + m;
      }
      z[size_t(i)] = (tmp[size_t(i)] - dot) / diag;
    }
  }

  // Jacobi-PCG in the reference's stage structure (solver.cc:262-366,
  // solver_driver.cc:318-337). stop_mode 0: rr < threshold (absolute, the
  // reference rule); 1: sqrt(rr) < threshold * ||b|| (relative).
  // precond 1: block IC(0) over `tiles` in place of Jacobi.
  template <typename T>
  int solve(int stop_mode, double threshold, int max_iter, int precond = 0,
            const int *tiles = nullptr) {
    Csr A;
    std::vector<int32_t> rows;
    assemble(A, rows);
    const int one[3] = {1, 1, 1};
    Tri<T> L, U;
    std::vector<T> ic_tmp;
    if (precond == 1 && !ic_factor<T>(A, rows, tiles ? tiles : one, L, U)) return -7;
    std::vector<T> b(n_rows), x(n_cols, T(0)), r(n_rows), z(n_rows), p(n_cols, T(0)),
        Ap(n_rows), dinv(n_rows);
    for (int q = 0; q < n_rows; ++q)
      b[q] = T(div.get(rows[3 * q], rows[3 * q + 1], rows[3 * q + 2]));
    // x from the (cleared) pressure grid over bbox+1 (solver.cc:86-88).
    for (int i = -1; i <= n[0]; ++i) for (int j = -1; j <= n[1]; ++j) for (int k = -1; k <= n[2]; ++k)
      if (idx.isOn(i, j, k)) x[idx.get(i, j, k)] = T(pressure.get(i, j, k));
    for (int q = 0; q < n_rows; ++q) {
      Real dg = 0;
      for (int64_t e = A.start[q]; e < A.start[q + 1]; ++e) if (A.col[e] == q) dg = A.val[e];
      dinv[q] = static_cast<T>(1.0 / double(dg));
      if (dg == 0) return -2;  // Jacobi diagonal 0: FLUID cell enclosed by SOLID
    }
    double bnorm = 0;
    for (int q = 0; q < n_rows; ++q) bnorm += double(b[q]) * double(b[q]);
    bnorm = std::sqrt(bnorm);
    const double thr = stop_mode == 0 ? threshold : threshold * bnorm;
    // Relative mode uses <=, so a zero right-hand side (x0 = 0) stops at once.
    auto done = [&](T rr) { return stop_mode == 0 ? double(rr) < thr : std::sqrt(double(rr)) <= thr; };
    // PreparePCGSolver.
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int q = 0; q < n_rows; ++q) { T m = row_dot(A, q, x); r[q] = b[q] - m; }
    auto precondition = [&]() {
      if (precond == 1) ic_apply<T>(L, U, r, z, ic_tmp);
      else {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int q = 0; q < n_rows; ++q) z[q] = dinv[q] * r[q];
      }
    };
    precondition();
    for (int q = 0; q < n_rows; ++q) p[q] = z[q];
    T rz = dot(r, z, n_rows), rr = dot(r, r, n_rows), rz_old = 0;
    rr_hist.clear();
    rr_hist.push_back(double(rr));
    rz_hist.assign(1, double(rz));
    pap_hist.clear();
    alpha_hist.clear();
    beta_hist.clear();
    iters = 0;
    while (!(iters >= max_iter || done(rr))) {
      if (std::isnan(double(rr))) return -3;
#pragma omp parallel for num_threads(threads) schedule(static)
      for (int q = 0; q < n_rows; ++q) Ap[q] = row_dot(A, q, p);
      T pAp = 0;
      for (int q = 0; q < n_rows; ++q) { T m = p[q] * Ap[q]; pAp = pAp + m; }
      T alpha = rz / pAp;
      pap_hist.push_back(double(pAp));
      alpha_hist.push_back(double(alpha));
      if (!(double(alpha) <= 1e5)) return -4;  // CHECK_LE(alpha, LARGE_FLOAT)
#pragma omp parallel for num_threads(threads) schedule(static)
      for (int q = 0; q < n_cols; ++q) { T m = alpha * p[q]; x[q] = m + x[q]; }
#pragma omp parallel for num_threads(threads) schedule(static)
      for (int q = 0; q < n_rows; ++q) { T m = -alpha * Ap[q]; r[q] = m + r[q]; }
      precondition();
      rr = dot(r, r, n_rows);
      rz_old = rz;
      rz = dot(r, z, n_rows);
      T beta = rz / rz_old;
      rz_hist.push_back(double(rz));
      beta_hist.push_back(double(beta));
      if (!(double(beta) <= 1e5)) return -5;  // CHECK_LE(beta, LARGE_FLOAT)
#pragma omp parallel for num_threads(threads) schedule(static)
      for (int q = 0; q < n_rows; ++q) { T m = beta * p[q]; p[q] = z[q] + m; }
      ++iters;
      rr_hist.push_back(double(rr));
    }
    rr_final = double(rr);
    // CopySolution (solver.cc:93-99): clear, then interior rows only.
    pressure.clear();
    for (int q = 0; q < n_rows; ++q)
      pressure.set(rows[3 * q], rows[3 * q + 1], rows[3 * q + 2], Real(x[q]));
    last_x64.assign(x.begin(), x.begin() + n_rows);
    return 0;
  }
  std::vector<double> last_x64;  // solution in row order, before the FP32 store
};

// Box-format helpers: dense over [-1, n] per axis, x-major.
inline size_t box_idx(const int n[3], int i, int j, int k) {
  return (size_t(i + 1) * (n[1] + 2) + (j + 1)) * (n[2] + 2) + (k + 1);
}

}  // namespace

extern "C" {

void *or_create(int nx, int ny, int nz) {
  Oracle *o = new Oracle();
  o->init(nx, ny, nz);
  o->threads = omp_get_max_threads();
  return o;
}
void or_destroy(void *h) { delete static_cast<Oracle *>(h); }
void or_set_params(void *h, Real gravity_scaled, Real fx, Real fy, Real fz, Real flip) {
  Oracle *o = static_cast<Oracle *>(h);
  o->gravity = gravity_scaled; o->fx = fx; o->fy = fy; o->fz = fz; o->flip = flip;
}
void or_set_box_boundary(void *h, int tank) { static_cast<Oracle *>(h)->set_box_boundary(tank != 0); }
void or_add_solid_box(void *h, const int *lo, const int *hi) {
  static_cast<Oracle *>(h)->add_solid_box(lo, hi);
}
// Particles are added in the given order (AddParticleToData); their cells are
// floor(pos), as the reference asserts (flip_simulation.cc:932-934).
void or_set_particles(void *h, int64_t n, const Real *pos, const Real *vel, const Real *gv) {
  Oracle *o = static_cast<Oracle *>(h);
  o->parts.resize(size_t(n));
  for (int64_t q = 0; q < n; ++q) {
    Particle &p = o->parts[size_t(q)];
    for (int d = 0; d < 3; ++d) {
      p.pos[d] = pos[3 * q + d]; p.vel[d] = vel[3 * q + d]; p.gv[d] = gv[3 * q + d];
      p.cell[d] = int(std::floor(p.pos[d]));
    }
  }
  o->sort_particles();
}
int64_t or_num_particles(void *h) { return int64_t(static_cast<Oracle *>(h)->parts.size()); }
void or_get_particles(void *h, Real *pos, Real *vel, Real *gv, int32_t *cell) {
  Oracle *o = static_cast<Oracle *>(h);
  for (size_t q = 0; q < o->parts.size(); ++q)
    for (int d = 0; d < 3; ++d) {
      const Particle &p = o->parts[q];
      if (pos) pos[3 * q + d] = p.pos[d];
      if (vel) vel[3 * q + d] = p.vel[d];
      if (gv) gv[3 * q + d] = p.gv[d];
      if (cell) cell[3 * q + d] = p.cell[d];
    }
}
Real or_cfl(void *h) { return static_cast<Oracle *>(h)->cfl(); }
// MoveParticles + DeleteParticlesInSolids (flip_app_sim.cc:496-517).
void or_move_particles(void *h, Real dt) {
  Oracle *o = static_cast<Oracle *>(h);
  o->step_in_grid(dt);
  o->delete_in_solids();
}
void or_clear(void *h) { static_cast<Oracle *>(h)->clear_grid(); }
void or_p2g(void *h) { static_cast<Oracle *>(h)->p2g(); }
void or_apply_bc(void *h) { static_cast<Oracle *>(h)->apply_bc(); }
void or_save_velocities(void *h) { static_cast<Oracle *>(h)->save_velocities(); }
void or_delete_outside(void *h) { static_cast<Oracle *>(h)->delete_outside(); }
void or_add_gravity(void *h, Real dt) { static_cast<Oracle *>(h)->add_gravity(dt); }
void or_divergence(void *h) { static_cast<Oracle *>(h)->divergence(); }
// precision 0: float, 1: double.
int or_solve(void *h, int precision, int stop_mode, double threshold, int max_iter,
             int *iters, double *rr) {
  Oracle *o = static_cast<Oracle *>(h);
  int rc = precision == 0 ? o->solve<Real>(stop_mode, threshold, max_iter)
                          : o->solve<double>(stop_mode, threshold, max_iter);
  if (iters) *iters = o->iters;
  if (rr) *rr = o->rr_final;
  return rc;
}
// As or_solve; precond 0 Jacobi, 1 block IC(0) over tiles[3] (null: one tile).
int or_solve_pc(void *h, int precision, int stop_mode, double threshold, int max_iter,
                int precond, const int *tiles, int *iters, double *rr) {
  Oracle *o = static_cast<Oracle *>(h);
  int rc = precision == 0 ? o->solve<Real>(stop_mode, threshold, max_iter, precond, tiles)
                          : o->solve<double>(stop_mode, threshold, max_iter, precond, tiles);
  if (iters) *iters = o->iters;
  if (rr) *rr = o->rr_final;
  return rc;
}
// The block IC(0) factor of the current system in the solver's precision:
// diag[row] = L_rr, and z = M^-1 r for r given in row order (z may be null).
// Returns n_rows, or -7 when the build fails.
int or_ic_factor(void *h, const int *tiles, double *diag, const double *r, double *z) {
  Oracle *o = static_cast<Oracle *>(h);
  Oracle::Csr A;
  std::vector<int32_t> rows;
  o->assemble(A, rows);
  const int one[3] = {1, 1, 1};
  Oracle::Tri<Real> L, U;
  if (!o->ic_factor<Real>(A, rows, tiles ? tiles : one, L, U)) return -7;
  for (int q = 0; q < o->n_rows; ++q) diag[q] = double(L.get(q, q));
  if (r && z) {
    std::vector<Real> rv(size_t(o->n_rows)), zv, tmp;
    for (int q = 0; q < o->n_rows; ++q) rv[size_t(q)] = Real(r[q]);
    Oracle::ic_apply<Real>(L, U, rv, zv, tmp);
    for (int q = 0; q < o->n_rows; ++q) z[q] = double(zv[size_t(q)]);
  }
  return o->n_rows;
}
int or_solver_sizes(void *h, int *n_rows, int *n_cols) {
  Oracle *o = static_cast<Oracle *>(h);
  *n_rows = o->n_rows; *n_cols = o->n_cols;
  return 0;
}
// Scalar history of the last solve: which 0 r.r (after prepare and after
// each iteration), 1 pAp, 2 alpha (per iteration), 3 r.z (after prepare and
// after each iteration), 4 beta (per iteration).
int or_scalar_history(void *h, int which, double *out, int cap) {
  Oracle *o = static_cast<Oracle *>(h);
  const std::vector<double> *v[5] = {&o->rr_hist, &o->pap_hist, &o->alpha_hist, &o->rz_hist,
                                     &o->beta_hist};
  if (which < 0 || which > 4) return -1;
  const int m = int(v[which]->size());
  for (int q = 0; q < m && q < cap; ++q) out[q] = (*v[which])[size_t(q)];
  return m;
}

// Threads of the solver's parallel loops (1: everything sequential).
void or_set_threads(void *h, int threads) {
  static_cast<Oracle *>(h)->threads = threads < 1 ? 1 : threads;
}

int or_residual_history(void *h, double *out, int cap) {
  Oracle *o = static_cast<Oracle *>(h);
  int m = int(o->rr_hist.size());
  for (int q = 0; q < m && q < cap; ++q) out[q] = o->rr_hist[size_t(q)];
  return m;
}
// Solution vector in row order, in the solver's own precision, widened.
int or_solution(void *h, double *out, int cap) {
  Oracle *o = static_cast<Oracle *>(h);
  int m = int(o->last_x64.size());
  for (int q = 0; q < m && q < cap; ++q) out[q] = o->last_x64[size_t(q)];
  return m;
}
void or_pressure_update(void *h) { static_cast<Oracle *>(h)->pressure_update(); }
int or_velocity_update(void *h) { return static_cast<Oracle *>(h)->velocity_update(); }
void or_extrapolate(void *h) {
  Oracle *o = static_cast<Oracle *>(h);
  Oracle::extrapolate(o->v, o->n);
  Oracle::extrapolate(o->vu, o->n);
}
void or_g2p(void *h) { static_cast<Oracle *>(h)->g2p(); }

// FlipDriver::RunOneStep (flip_app_sim.cc:211-254) for one partition with no
// sources. Returns the solver status.
int or_run_one_step(void *h, Real dt, int precision, int stop_mode, double threshold,
                    int max_iter, int *iters) {
  Oracle *o = static_cast<Oracle *>(h);
  o->step_in_grid(dt);
  o->delete_in_solids();
  o->clear_grid();
  o->p2g();
  o->apply_bc();
  o->save_velocities();
  o->delete_outside();
  o->add_gravity(dt);
  o->apply_bc();
  o->divergence();
  int rc = precision == 0 ? o->solve<Real>(stop_mode, threshold, max_iter)
                          : o->solve<double>(stop_mode, threshold, max_iter);
  if (iters) *iters = o->iters;
  if (rc != 0) return rc;
  o->pressure_update();
  if (o->velocity_update() != 0) return -6;
  o->g2p();
  return 0;
}

// WhileAndComputeDt's reduction and clamps (flip_app_sim.cc:374, 395-407) for
// one partition. Returns dt and advances *step.
Real or_compute_dt(void *h, int cfl, Real frame_step, Real *step) {
  Oracle *o = static_cast<Oracle *>(h);
  Real local_dt = Real(cfl) * o->cfl();
  Real global_dt = Real(1.0 / double(frame_step));
  global_dt = std::min(global_dt, local_dt);
  if (global_dt < frame_step / 10) global_dt = frame_step / 10;
  if (*step + global_dt > frame_step) global_dt = frame_step - *step;
  if (double(*step + global_dt) > 0.9 * double(frame_step)) global_dt = frame_step - *step;
  *step += global_dt;
  return global_dt;
}

// Box-format access. which: 0 grid_velocity, 1 grid_velocity_update, 2 weight.
void or_get_vec(void *h, int which, Real *out, uint8_t *active) {
  Oracle *o = static_cast<Oracle *>(h);
  const DenseVec &g = which == 0 ? o->v : which == 1 ? o->vu : o->w;
  for (int i = -1; i <= o->n[0]; ++i) for (int j = -1; j <= o->n[1]; ++j) for (int k = -1; k <= o->n[2]; ++k) {
    size_t b = box_idx(o->n, i, j, k);
    for (int d = 0; d < 3; ++d) out[3 * b + d] = g.c[d].get(i, j, k);
    if (active) active[b] = g.isOn(i, j, k) ? 1 : 0;
  }
}
void or_set_vec(void *h, int which, const Real *in, const uint8_t *active) {
  Oracle *o = static_cast<Oracle *>(h);
  DenseVec &g = which == 0 ? o->v : which == 1 ? o->vu : o->w;
  g.clear();
  for (int i = -1; i <= o->n[0]; ++i) for (int j = -1; j <= o->n[1]; ++j) for (int k = -1; k <= o->n[2]; ++k) {
    size_t b = box_idx(o->n, i, j, k);
    if (active[b]) g.set(i, j, k, &in[3 * b]);
  }
}
// which: 0 divergence, 1 pressure.
void or_get_scalar(void *h, int which, Real *out) {
  Oracle *o = static_cast<Oracle *>(h);
  const Dense<Real> &g = which == 0 ? o->div : o->pressure;
  for (int i = -1; i <= o->n[0]; ++i) for (int j = -1; j <= o->n[1]; ++j) for (int k = -1; k <= o->n[2]; ++k)
    out[box_idx(o->n, i, j, k)] = g.get(i, j, k);
}
// which: 0 marker, 1 solid, 2 solver row index (-1 where inactive).
void or_get_int(void *h, int which, int32_t *out) {
  Oracle *o = static_cast<Oracle *>(h);
  for (int i = -1; i <= o->n[0]; ++i) for (int j = -1; j <= o->n[1]; ++j) for (int k = -1; k <= o->n[2]; ++k) {
    int32_t val;
    if (which == 0) val = o->marker.get(i, j, k);
    else if (which == 1) val = o->solid.get(i, j, k);
    else val = o->idx.isOn(i, j, k) ? o->idx.get(i, j, k) : -1;
    out[box_idx(o->n, i, j, k)] = val;
  }
}
void or_set_marker(void *h, const int32_t *in) {
  Oracle *o = static_cast<Oracle *>(h);
  o->marker.clear();
  for (int i = -1; i <= o->n[0]; ++i) for (int j = -1; j <= o->n[1]; ++j) for (int k = -1; k <= o->n[2]; ++k) {
    int32_t m = in[box_idx(o->n, i, j, k)];
    if (m != AIR) o->marker.set(i, j, k, m);
  }
}
void or_set_scalar(void *h, int which, const Real *in) {
  Oracle *o = static_cast<Oracle *>(h);
  Dense<Real> &g = which == 0 ? o->div : o->pressure;
  g.clear();
  for (int i = -1; i <= o->n[0]; ++i) for (int j = -1; j <= o->n[1]; ++j) for (int k = -1; k <= o->n[2]; ++k) {
    Real x = in[box_idx(o->n, i, j, k)];
    if (x != 0.0f) g.set(i, j, k, x);
  }
}
// Particles per cell in box format (cells are floor(pos) after clamping).
void or_cell_counts(void *h, int32_t *out) {
  Oracle *o = static_cast<Oracle *>(h);
  size_t s = size_t(o->n[0] + 2) * (o->n[1] + 2) * (o->n[2] + 2);
  std::fill(out, out + s, 0);
  for (auto &p : o->parts) out[box_idx(o->n, p.cell[0], p.cell[1], p.cell[2])] += 1;
}

}  // extern "C"
