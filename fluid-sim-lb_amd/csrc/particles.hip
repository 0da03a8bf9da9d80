// Particle-side kernels: seeding, advection + counting sort by cell, leaf
// activation, particle-to-grid (P2G) and grid-to-particle (G2P).
//
// Particles live in nine SoA float arrays sorted by cell. Cells are ordered
// as the reference's particle tree is traversed (VDB order: root tiles of
// 64^3 by origin, leaves by offset in the tile, voxels x-major; sort_index),
// so a leaf's particles are one contiguous range and a cell's particles the
// sub-range [vstart[s], vstart[s+1]) at its sort index s. Within a cell,
// particles keep their previous order (the sort is stable), as the
// reference's StepInGrid re-adds particles in traversal order
// (particles.cc:489-544): the device order is the reference's particle
// order, and results are deterministic run to run.
#include "device_util.h"
#include "fsl_internal.h"

namespace fsl {

namespace {

constexpr int kThreads = 256;

// ---------------------------------------------------------------- seeding
__device__ inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint32_t hi0 = __umulhi(M0, c[0]), lo0 = M0 * c[0];
    const uint32_t hi1 = __umulhi(M1, c[2]), lo1 = M1 * c[2];
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}
__device__ inline float jitter(int i, uint32_t x) {
  const float u = float(x >> 8) * (1.0f / 16777216.0f);
  const float p = __fadd_rn(float(i), u);
  const float top = __uint_as_float(__float_as_uint(float(i + 1)) - 1u);
  return fminf(p, top);
}

// One thread per fluid cell (mask over [0,n)^3, x-major); writes `ppc`
// particles at offs[cell] * ppc.
__global__ void k_seed(Dom d, const uint8_t *mask, const uint32_t *offs, int ppc, uint32_t seed,
                       float v0, float v1, float v2, Particles P) {
  const int64_t cell = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t ncell = int64_t(d.n[0]) * d.n[1] * d.n[2];
  if (cell >= ncell || !mask[cell]) return;
  const int i = int(cell / (int64_t(d.n[1]) * d.n[2]));
  const int j = int((cell / d.n[2]) % d.n[1]);
  const int k = int(cell % d.n[2]);
  const int64_t base = int64_t(offs[cell]) * ppc;
  for (int s = 0; s < ppc; ++s) {
    uint32_t c[4] = {uint32_t(cell), uint32_t(s), 0u, 0u};
    philox10(c, seed, 0u);
    const int64_t q = base + s;
    P.a[0][q] = jitter(i, c[0]);
    P.a[1][q] = jitter(j, c[1]);
    P.a[2][q] = jitter(k, c[2]);
    P.a[3][q] = v0; P.a[4][q] = v1; P.a[5][q] = v2;
    P.a[6][q] = v0; P.a[7][q] = v1; P.a[8][q] = v2;
  }
}

// AddContributionFromSources (flip_simulation.cc:1206-1266), particles only
// (the marker and face velocities the reference also sets there are cleared
// by ClearGridData before anything reads them). Cell m of the source's cell
// box (x-major) makes up to 32 attempts, each a point cell + (u0, u1, u2)
// with u uniform in [0, 1) from Philox4x32-10 (counter (cell, attempt, 2, 0),
// key (k0, k1)) in place of glibc random(); points inside the box (bounds
// inclusive) are kept until ppc of them are.
__device__ inline void source_cell(const SourceBox &s, int64_t m, int c[3]) {
  const int64_t e1 = s.c1[1] - s.c0[1] + 1, e2 = s.c1[2] - s.c0[2] + 1;
  c[0] = s.c0[0] + int(m / (e1 * e2));
  c[1] = s.c0[1] + int((m / e2) % e1);
  c[2] = s.c0[2] + int(m % e2);
}
template <typename F>
__device__ inline void source_points(const Dom &d, const SourceBox &s, const int c[3], F keep) {
  const uint32_t lin = uint32_t((int64_t(c[0]) * d.n[1] + c[1]) * d.n[2] + c[2]);
  int kept = 0;
  for (int a = 0; a < 32 && kept < s.ppc; ++a) {
    uint32_t x[4] = {lin, uint32_t(a), 2u, 0u};
    philox10(x, s.k0, s.k1);
    float p[3];
    bool in = true;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      p[e] = __fadd_rn(float(c[e]), float(x[e] >> 8) * (1.0f / 16777216.0f));
      in = in && p[e] >= s.lo[e] && p[e] <= s.hi[e];
    }
    if (!in) continue;
    keep(kept, p);
    ++kept;
  }
}
__global__ void k_source_count(Dom d, SourceBox s, int64_t ncell, uint32_t *cnt) {
  const int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= ncell) return;
  int c[3];
  source_cell(s, m, c);
  uint32_t n = 0;
  source_points(d, s, c, [&](int, const float *) { ++n; });
  cnt[m] = n;
}
__global__ void k_source_emit(Dom d, SourceBox s, int64_t ncell, const uint32_t *off,
                              Particles P, int64_t base) {
  const int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= ncell) return;
  int c[3];
  source_cell(s, m, c);
  const int64_t o = base + off[m];
  source_points(d, s, c, [&](int j, const float *p) {
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      P.a[e][o + j] = p[e];
      P.a[3 + e][o + j] = s.vel[e];
      P.a[6 + e][o + j] = s.vel[e];
    }
  });
}

__global__ void k_mask_to_count(const uint8_t *mask, int64_t n, uint32_t *cnt) {
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q < n) cnt[q] = mask[q] ? 1u : 0u;
}

// ------------------------------------------------------ advection and sort
// Runs of equal keys within a wavefront. Particles arrive sorted by cell from
// the previous step and move at most about one cell, so neighbouring lanes
// mostly share a key: one atomic per run instead of one per particle.
struct KeyRun {
  int head;    // lane of the first particle of this lane's run
  int count;   // run length (valid in the head lane)
  int rank;    // position of this lane in its run
};
__device__ inline KeyRun key_run(uint32_t key) {
  const int lane = __lane_id();
  const uint32_t prev = __shfl_up(key, 1);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long le = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int h = 63 - __clzll(heads & le);
  const unsigned long long gt = heads & ~le;
  const int nxt = gt ? __ffsll((long long)gt) - 1 : 64;
  return {h, nxt - h, lane - h};
}

// Particles::StepInGrid (particles.cc:504-543) + DeleteParticlesInSolids
// (particles.cc:202-219). Particles outside the partition box were deleted by
// DeleteOutsideParticles after the previous P2G; they are dropped here.
// np_dev and dt_dev, when set, hold the particle count and dt on the device
// (np is then an upper bound that sizes the launch).
__global__ void k_advect(Dom d, int64_t np, const uint32_t *np_dev, Particles P, float dt,
                         const float *dt_dev, const uint8_t *solid, uint32_t *key,
                         uint32_t *vcount, int drop_outside, KillBoxes kill) {
#pragma clang fp contract(off)
  if (np_dev) np = *np_dev;
  if (dt_dev) dt = *dt_dev;
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t k = kDead;
  if (q < np) {
    float p[3] = {P.a[0][q], P.a[1][q], P.a[2][q]};
    bool alive = true;
    if (drop_outside) {
      const int a = int(floorf(p[0])), b = int(floorf(p[1])), c = int(floorf(p[2]));
      if (!in_bbox(d, a, b, c)) alive = false;
    }
    if (alive) {
      int id[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float m = P.a[6 + a][q] * dt;
        p[a] = p[a] + m;
        id[a] = int(floorf(p[a]));
        if (id[a] < 0) { id[a] = 0; p[a] = 0.0f; }
      }
#pragma unroll
      for (int a = 0; a < 3; ++a)
        if (id[a] >= d.n[a]) { id[a] = d.n[a]; p[a] = float(d.n[a]); }
      P.a[0][q] = p[0]; P.a[1][q] = p[1]; P.a[2][q] = p[2];
      const int64_t v = vox_of(d, id[0], id[1], id[2]);
      bool killed = false;
      for (int b = 0; b < kill.n; ++b)
        killed = killed || (p[0] >= kill.lo[b][0] && p[0] <= kill.hi[b][0] &&
                            p[1] >= kill.lo[b][1] && p[1] <= kill.hi[b][1] &&
                            p[2] >= kill.lo[b][2] && p[2] <= kill.hi[b][2]);
      // The sort key ranks cells in VDB traversal order (the leaf's rank,
      // then the voxel offset), so the sorted order is the reference's
      // particle order.
      if (!solid[v] && !killed) k = uint32_t(sort_index(d, v));
    }
    key[q] = k;
  }
  const KeyRun r = key_run(k);
  if (r.rank == 0 && k != kDead) atomicAdd(&vcount[k], uint32_t(r.count));
}

// Buffer loads of four 32-bit values: the sort and P2G read a cell's run of
// particles (or particle indices) four at a time through a descriptor based
// at the leaf's first particle. Word 3 of a raw (untyped, stride 0)
// descriptor on gfx9-family chips: data format 32 bits, no swizzle.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kRawBufferWord3 = 0x00020000;

// Counting-sort scatter. vcount returns to zero, ready for the next step.
// A run of equal keys takes one atomic and fills consecutive slots in lane
// order, ascending from the cell's next free slot, so a cell whose particles
// arrive in one run is already in its previous order.
__global__ void k_scatter(int64_t np, const uint32_t *np_dev, const uint32_t *key,
                          const uint32_t *vstart, uint32_t *vcount, uint32_t *perm) {
  if (np_dev) np = *np_dev;
  // Four independent windows of particles per thread (q, q + T, q + 2T,
  // q + 3T for T threads), so four atomics are in flight per wavefront.
  const int64_t nt = int64_t(gridDim.x) * blockDim.x;
  const int64_t q0 = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t k[4], left[4];
  KeyRun r[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t q = q0 + j * nt;
    k[j] = q < np ? key[q] : kDead;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) r[j] = key_run(k[j]);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    left[j] = 0;  // slots of the cell still free before this run
    if (r[j].rank == 0 && k[j] != kDead) left[j] = atomicSub(&vcount[k[j]], uint32_t(r[j].count));
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) left[j] = __shfl(left[j], r[j].head);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (k[j] == kDead) continue;
    const uint32_t b = vstart[k[j]], total = vstart[k[j] + 1] - b;
    perm[b + (total - left[j]) + uint32_t(r[j].rank)] = uint32_t(q0 + j * nt);
  }
}

// Restore the previous order inside each cell (stable sort): one workgroup
// per leaf in sort order (leaf rank r owns sort indices r * 512 + t), one
// thread per cell; leaves without particles leave at once. A cell's old
// indices are sorted when runs arrived out of order (a cell filled by
// several runs). A cell of at most kCellRegs particles is read into
// registers four indices per load (a buffer descriptor based at the leaf's
// first slot; perm carries kParticlePad spare elements for the loads past the
// leaf's last slot), checked and sorted there (odd-even transposition,
// padding sorts last) and written back only when it was out of order; larger
// cells take the insertion sort in memory.
constexpr uint32_t kCellRegs = 16;
__global__ __launch_bounds__(512) void k_cell_sort(const uint32_t *vstart, uint32_t *perm) {
  const int64_t s0 = int64_t(blockIdx.x) << 9;
  const uint32_t first = vstart[s0], last = vstart[s0 + 512];
  if (last - first < 2) return;
  const uint32_t b = vstart[s0 + threadIdx.x], e = vstart[s0 + threadIdx.x + 1];
  const uint32_t m = e - b;
  if (m < 2) return;
  if (m <= kCellRegs) {
    const __amdgpu_buffer_rsrc_t R = __builtin_amdgcn_make_buffer_rsrc(
        perm + first, 0, int((last - first + 3) * 4), kRawBufferWord3);
    uint32_t v[kCellRegs];
#pragma unroll
    for (uint32_t j = 0; j < kCellRegs / 4; ++j) {
      u32x4 x = {~0u, ~0u, ~0u, ~0u};
      if (4 * j < m) x = __builtin_amdgcn_raw_buffer_load_b128(R, int(b - first + 4 * j) * 4, 0, 0);
      v[4 * j] = x.x; v[4 * j + 1] = x.y; v[4 * j + 2] = x.z; v[4 * j + 3] = x.w;
    }
    bool in_order = true;
#pragma unroll
    for (uint32_t i = 1; i < kCellRegs; ++i)
      if (i < m) in_order = in_order && v[i - 1] < v[i];
    if (in_order) return;
#pragma unroll
    for (uint32_t i = 0; i < kCellRegs; ++i)
      if (i >= m) v[i] = 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t round = 0; round < kCellRegs; ++round)
#pragma unroll
      for (uint32_t i = round & 1; i + 1 < kCellRegs; i += 2) {
        const uint32_t lo = v[i] < v[i + 1] ? v[i] : v[i + 1];
        const uint32_t hi = v[i] < v[i + 1] ? v[i + 1] : v[i];
        v[i] = lo;
        v[i + 1] = hi;
      }
#pragma unroll
    for (uint32_t i = 0; i < kCellRegs; ++i)
      if (i < m) perm[b + i] = v[i];
    return;
  }
  bool sorted = true;
  for (uint32_t x = b + 1; x < e && sorted; ++x) sorted = perm[x - 1] < perm[x];
  if (sorted) return;
  for (uint32_t x = b + 1; x < e; ++x) {
    const uint32_t val = perm[x];
    uint32_t y = x;
    while (y > b && perm[y - 1] > val) { perm[y] = perm[y - 1]; --y; }
    perm[y] = val;
  }
}

// Moves arrays 0..NA-1 into sorted order (NA = 6 leaves out the grid
// velocities gx, gy, gz, which only advection reads and G2P rewrites).
// With n_dev (the sorted count on the device) the new count is also stored
// in *count_out.
template <int NA>
__global__ void k_gather(uint32_t n, const uint32_t *n_dev, uint32_t *count_out,
                         const uint32_t *perm, Particles in, Particles out) {
  if (n_dev) {
    n = *n_dev;
    if (blockIdx.x == 0 && threadIdx.x == 0) *count_out = n;
  }
  // Four consecutive destinations per thread: one 16-byte index load and one
  // 16-byte store per array (every array is its own allocation, so element
  // 4k is 16-byte aligned); the tail goes element by element.
  const uint32_t s = (blockIdx.x * blockDim.x + threadIdx.x) * 4u;
  if (s >= n) return;
  if (s + 3u < n) {
    const uint4 q = *reinterpret_cast<const uint4 *>(perm + s);
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      const float4 v = make_float4(in.a[a][q.x], in.a[a][q.y], in.a[a][q.z], in.a[a][q.w]);
      *reinterpret_cast<float4 *>(out.a[a] + s) = v;
    }
    return;
  }
  for (uint32_t i = s; i < n; ++i) {
    const uint32_t q = perm[i];
#pragma unroll
    for (int a = 0; a < NA; ++a) out.a[a][i] = in.a[a][q];
  }
}

// ------------------------------------------------------------- leaf lists
__global__ void k_leaf_has(int64_t nleaves, const int32_t *lrank, const uint32_t *vstart,
                           uint32_t *lhas) {
  const int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l >= nleaves) return;
  const int64_t r = lrank[l];
  lhas[l] = vstart[(r + 1) * kLeaf] > vstart[r * kLeaf] ? 1u : 0u;
}
// A leaf is active when it or one of its 26 neighbours holds particles:
// everything the reference writes within a step lies within two cells of a
// particle's cell, hence in such a leaf.
__global__ void k_leaf_dilate(Dom d, const uint32_t *lhas, uint32_t *lact) {
  const int64_t l = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (l >= leaf_count(d)) return;
  const int a = int(l / (int64_t(d.L[1]) * d.L[2])), b = int((l / d.L[2]) % d.L[1]),
            c = int(l % d.L[2]);
  uint32_t act = 0;
  for (int x = a - 1; x <= a + 1; ++x)
    for (int y = b - 1; y <= b + 1; ++y)
      for (int z = c - 1; z <= c + 1; ++z) {
        if (x < 0 || y < 0 || z < 0 || x >= d.L[0] || y >= d.L[1] || z >= d.L[2]) continue;
        act |= lhas[(int64_t(x) * d.L[1] + y) * d.L[2] + z];
      }
  lact[l] = act;
}

// Leaf lists are emitted in Morton (Z-curve) order of the leaf coordinates:
// a contiguous chunk of a list is then a compact 3-D block, so the per-XCD
// chunks of the leaf kernels re-read their halos from their own L2.
__global__ void k_gather_flags(int64_t n, const int32_t *order, const uint32_t *flag,
                               uint32_t *fm) {
  const int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m < n) fm[m] = flag[order[m]] ? 1u : 0u;
}
__global__ void k_compact(int64_t n, const int32_t *order, const uint32_t *fm,
                          const uint32_t *pos, int32_t *list, uint32_t *count_out) {
  const int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m < n && fm[m]) list[pos[m]] = order[m];
  if (count_out && m == n - 1) *count_out = pos[m] + fm[m];
}

// ------------------------------------------------------------------ P2G
// TransferParticlesToGrid (flip_simulation.cc:808-923): a particle in cell c
// reaches, per component, the 18 samples c + (sd, sa - 1, sb - 1) along (D,
// a, b) with sd in {0, 1} and sa, sb in {0, 1, 2} (flip_simulation.cc:
// 936-1009). Each thread sums its cell's particles over these 18 slots in
// registers before anything is written.
template <int D>
struct Axes;  // the two axes other than D
template <> struct Axes<0> { static constexpr int a = 1, b = 2; };
template <> struct Axes<1> { static constexpr int a = 0, b = 2; };
template <> struct Axes<2> { static constexpr int a = 0, b = 1; };

// Accumulate one particle into the 18 sample slots (slot = sd*9 + sa*3 + sb;
// sample offset from the cell: +sd along D, sa-1 along a, sb-1 along b).
// The trilinear weights of ScatterWeight (flip_simulation.cc:936-1009) are
// formed as products of per-axis factors: along D, (1 - t, t) on the slots
// sd = 0, 1; along a (and b), (1 - t, t) on the slots (0, 1) when the
// particle lies below the cell's centre line (the lower sample is in cell
// c - 1) or on the slots (1, 2) otherwise. The reference's factored form
// (flip_simulation.cc:977-990) is the same polynomial; the two roundings
// differ in the last bits of a weight. Each slot takes two fused
// multiply-adds, w * vel as (fd * vel) * (fa * fb) and w as fd * (fa * fb),
// so the weight itself is never rounded on its own.
template <int D>
__device__ inline void p2g_particle(const float pos[3], const int c[3], float vel, float avw[18],
                                    float aw[18], uint32_t &cls) {
#pragma clang fp contract(off)
  constexpr int A = Axes<D>::a, B = Axes<D>::b;
  const float td = pos[D] - float(c[D]);
  const float fd[2] = {1.0f - td, td};
  float fa[3], fb[3];
  int ha, hb;
  {
    const float mid = float(c[A]) + 0.5f;
    ha = pos[A] < mid ? 0 : 1;
    const float t = pos[A] - (ha ? mid : mid - 1.0f), u = 1.0f - t;
    fa[0] = ha ? 0.0f : u;
    fa[1] = ha ? u : t;
    fa[2] = ha ? t : 0.0f;
  }
  {
    const float mid = float(c[B]) + 0.5f;
    hb = pos[B] < mid ? 0 : 1;
    const float t = pos[B] - (hb ? mid : mid - 1.0f), u = 1.0f - t;
    fb[0] = hb ? 0.0f : u;
    fb[1] = hb ? u : t;
    fb[2] = hb ? t : 0.0f;
  }
  cls |= 1u << (ha * 2 + hb);
  float fab[3][3];
#pragma unroll
  for (int sa = 0; sa < 3; ++sa)
#pragma unroll
    for (int sb = 0; sb < 3; ++sb) fab[sa][sb] = fa[sa] * fb[sb];
  const float vd[2] = {fd[0] * vel, fd[1] * vel};
#pragma unroll
  for (int sd = 0; sd < 2; ++sd)
#pragma unroll
    for (int sa = 0; sa < 3; ++sa)
#pragma unroll
      for (int sb = 0; sb < 3; ++sb) {
        const int s = sd * 9 + sa * 3 + sb;
        avw[s] = fmaf(vd[sd], fab[sa][sb], avw[s]);
        aw[s] = fmaf(fd[sd], fab[sa][sb], aw[s]);
      }
}

// Bitmask of particle classes (ha, hb) whose stencil reaches slot (sa, sb).
__host__ __device__ constexpr uint32_t slot_classes(int sa, int sb) {
  uint32_t m = 0;
  for (int ha = 0; ha < 2; ++ha)
    for (int hb = 0; hb < 2; ++hb)
      if (sa - ha >= 0 && sa - ha <= 1 && sb - hb >= 0 && sb - hb <= 1) m |= 1u << (ha * 2 + hb);
  return m;
}

// Ring of a leaf's 10^3 region (region coordinate r <-> local voxel r - 1):
// the 488 cells with a coordinate 0 or 9, i.e. the samples of neighbouring
// leaves that the leaf's own cells reach. Ring entry i: [0, 200) x = 0 or 9
// (y, z in 0..9); [200, 360) y = 0 or 9 (x in 1..8, z in 0..9); [360, 488)
// z = 0 or 9 (x, y in 1..8).
constexpr int kRing = 488;
__device__ inline void ring_cell(int i, int &x, int &y, int &z) {
  if (i < 200) {
    x = i < 100 ? 0 : 9;
    const int j = i % 100;
    y = j / 10; z = j % 10;
  } else if (i < 360) {
    const int j = i - 200, m = j % 80;
    y = j < 80 ? 0 : 9;
    x = 1 + m / 10; z = m % 10;
  } else {
    const int j = i - 360, m = j % 64;
    z = j < 64 ? 0 : 9;
    x = 1 + m / 8; y = 1 + m % 8;
  }
}
__device__ inline int ring_index(int x, int y, int z) {
  if (x == 0 || x == 9) return (x ? 100 : 0) + y * 10 + z;
  if (y == 0 || y == 9) return 200 + (y ? 80 : 0) + (x - 1) * 10 + z;
  return 360 + (z ? 64 : 0) + (x - 1) * 8 + (y - 1);
}

// P2G reads the particle arrays through buffer descriptors based at a leaf's
// first particle, so offsets stay 32-bit at any particle count; the record
// count covers the leaf's run plus the three particles a four-wide load can
// reach past it (the arrays carry kParticlePad spare elements).
struct LeafRuns {
  __amdgpu_buffer_rsrc_t a[6];  // px py pz and the three particle values
  uint32_t first;               // the leaf's first particle
};

// One cell's particles, all three components in one pass (each particle's
// position and velocity are loaded once), summed in particle order into the
// 18 slots of each component in registers; then committed to the region's
// LDS samples. Loads take four consecutive particles per array (one 16-byte
// load per array and trip: the load count, not the bytes, limits this loop);
// particles past the cell's end are not accumulated.
//
// Commit: the x offset of slot (sd, sa, sb) is sd for component x and sa - 1
// for y and z. Grouped by that offset (-1, 0, +1), the cells of one x plane
// (one wavefront) write only to one plane of samples, a different plane for
// each wavefront, so a group needs no workgroup barrier; inside a
// wavefront, LDS operations are performed in issue order and within one
// slot every lane targets a different sample, so the slots of a group are
// applied one after another in a fixed order (the empty asm keeps the
// compiler from reordering LDS accesses across slots). The touch flags are
// only ever set to 1. r is the cell's region index (x * 100 + y * 10 + z).
__device__ inline void p2g_own(const LeafRuns &R, const int c[3], int r, uint32_t beg,
                               uint32_t end, float (*s_vw)[1000], float (*s_w)[1000],
                               uint8_t *s_touch) {
  float avw[3][18], aw[3][18];
  uint32_t cls[3] = {0u, 0u, 0u};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int q = 0; q < 18; ++q) { avw[a][q] = 0.0f; aw[a][q] = 0.0f; }
  for (uint32_t q = beg; q < end; q += 4) {
    float pp[4][3], pv[4][3];
    const int off = int(q - R.first) * 4;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const u32x4 x = __builtin_amdgcn_raw_buffer_load_b128(R.a[a], off, 0, 0);
      const u32x4 y = __builtin_amdgcn_raw_buffer_load_b128(R.a[3 + a], off, 0, 0);
      pp[0][a] = __uint_as_float(x.x); pp[1][a] = __uint_as_float(x.y);
      pp[2][a] = __uint_as_float(x.z); pp[3][a] = __uint_as_float(x.w);
      pv[0][a] = __uint_as_float(y.x); pv[1][a] = __uint_as_float(y.y);
      pv[2][a] = __uint_as_float(y.z); pv[3][a] = __uint_as_float(y.w);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (q + u >= end) break;
      p2g_particle<0>(pp[u], c, pv[u][0], avw[0], aw[0], cls[0]);
      p2g_particle<1>(pp[u], c, pv[u][1], avw[1], aw[1], cls[1]);
      p2g_particle<2>(pp[u], c, pv[u][2], avw[2], aw[2], cls[2]);
    }
  }
  constexpr int st[3] = {100, 10, 1};
  const bool any = end > beg;
#pragma unroll
  for (int dx = -1; dx <= 1; ++dx) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      if (dx < 0 && k >= 6) continue;  // x has no slot at offset -1: six rounds
      if (any) {
#pragma unroll
        for (int D = 0; D < 3; ++D) {
          int sd, sa, sb;
          if (D == 0) {
            if (dx < 0) continue;
            sd = dx; sa = k / 3; sb = k % 3;
          } else {
            if (k >= 6) continue;
            sa = dx + 1; sd = k / 3; sb = k % 3;
          }
          const int slot = sd * 9 + sa * 3 + sb;
          const int A = D == 0 ? 1 : 0, B = D == 2 ? 1 : 2;
          const int g = r + sd * st[D] + (sa - 1) * st[A] + (sb - 1) * st[B];
          s_vw[D][g] += avw[D][slot];
          s_w[D][g] += aw[D][slot];
          if (cls[D] & slot_classes(sa, sb)) s_touch[g] = 1;
        }
      }
      asm volatile("" ::: "memory");
    }
    __syncthreads();
  }
}

// TransferParticlesToGrid (flip_simulation.cc:808-923) restated as a gather
// in two kernels. k_p2g_own: one workgroup of 256 threads per leaf with
// particles; thread t sums cell t, then cell t + 256 (x in 0..3, then 4..7),
// into the samples of the leaf's 10^3 region (each particle is visited
// once); the own samples' partial sums go to vx/vy/vz, ws* and vflag, the
// ring samples (those of neighbouring leaves) to the leaf's ring block.
// 256 threads let two leaves share a CU, so one leaf's commit overlaps the
// other's sums. k_p2g_combine then completes every active leaf's samples
// from the ring blocks of its neighbours, in a fixed order, and normalises.
// ADJ = true: the transpose of G2P's interpolation, i.e. the un-normalised
// sums sum_p w * value_p with the particle values taken from P.a[3..5]; no
// weights, labels or activity are written.
constexpr int kP2gThreads = 256;
template <bool ADJ>
__global__ __launch_bounds__(kP2gThreads) void k_p2g_own(
    Dom d, LeafList L, const uint32_t *lhas, const uint32_t *vstart, Particles P, float *vx,
    float *vy, float *vz, float *wx, float *wy, float *wz, uint8_t *vflag, float *ring,
    uint8_t *ring_t, int64_t rstride) {
  constexpr int T = kP2gThreads, NC = 512 / T;
  __shared__ float s_vw[3][1000];
  __shared__ float s_w[3][1000];
  __shared__ uint8_t s_touch[1000];
  for_each_leaf(L.list, L.count, L.n, [&](int64_t leaf) {
    // The previous leaf's LDS reads are done before the tile is refilled.
    __syncthreads();
    if (!lhas[leaf]) return;
    int o[3];
    leaf_origin(d, leaf, o);
    const int t = threadIdx.x;
    for (int i = t; i < 1000; i += T) {
#pragma unroll
      for (int a = 0; a < 3; ++a) { s_vw[a][i] = 0.0f; s_w[a][i] = 0.0f; }
      s_touch[i] = 0;
    }
    const int64_t sb = sort_begin(d, leaf);
    // The leaf's particle run, the same for every thread.
    LeafRuns R;
    R.first = __builtin_amdgcn_readfirstlane(vstart[sb]);
    const uint32_t last = __builtin_amdgcn_readfirstlane(vstart[sb + 512]);
    const int bytes = int((last - R.first + 3) * 4);
#pragma unroll
    for (int a = 0; a < 6; ++a)
      R.a[a] = __builtin_amdgcn_make_buffer_rsrc(P.a[a] + R.first, 0, bytes, kRawBufferWord3);
    uint32_t beg[NC], end[NC];
#pragma unroll
    for (int h = 0; h < NC; ++h) {
      beg[h] = vstart[sb | (t + h * T)];
      end[h] = vstart[(sb | (t + h * T)) + 1];
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < NC; ++h) {
      const int cell = t + h * T, lx = cell >> 6, ly = (cell >> 3) & 7, lz = cell & 7;
      const int c[3] = {o[0] + lx, o[1] + ly, o[2] + lz};
      p2g_own(R, c, (lx + 1) * 100 + (ly + 1) * 10 + (lz + 1), beg[h], end[h], s_vw, s_w,
              s_touch);
    }
#pragma unroll
    for (int h = 0; h < NC; ++h) {
      const int cell = t + h * T, lx = cell >> 6, ly = (cell >> 3) & 7, lz = cell & 7;
      const int r = (lx + 1) * 100 + (ly + 1) * 10 + (lz + 1);
      const int64_t v = (leaf << 9) | cell;
      vx[v] = s_vw[0][r]; vy[v] = s_vw[1][r]; vz[v] = s_vw[2][r];
      if (!ADJ) {
        wx[v] = s_w[0][r]; wy[v] = s_w[1][r]; wz[v] = s_w[2][r];
        vflag[v] = s_touch[r];
      }
    }
    for (int i = t; i < kRing; i += T) {
      int x, y, z;
      ring_cell(i, x, y, z);
      const int g = x * 100 + y * 10 + z;
      const int64_t e = leaf * kRing + i;
#pragma unroll
      for (int a = 0; a < 3; ++a) ring[a * rstride + e] = s_vw[a][g];
      if (!ADJ) {
#pragma unroll
        for (int a = 0; a < 3; ++a) ring[(3 + a) * rstride + e] = s_w[a][g];
        ring_t[e] = s_touch[g];
      }
    }
  });
}

// Completes the samples of every active leaf: its own partial sums (zero
// when it holds no particles) plus, for samples on its boundary, the ring
// entries of the neighbouring leaves with particles, neighbours in (dx, dy,
// dz) order; then normalises (flip_simulation.cc:900-922) and marks FLUID
// (flip_simulation.cc:852-857). 128 threads per leaf, each a run of four
// voxels along z, so a thread's own values move as 16-byte (float) and
// 4-byte (label) accesses. The neighbours holding particles are found once
// per leaf (one ballot per wavefront), and each voxel's ring entries (up to
// seven neighbours) are loaded together before they are added in (dx, dy,
// dz) order: one round trip per voxel instead of one per neighbour.
constexpr int kCombineThreads = 128;
template <bool ADJ>
__global__ __launch_bounds__(kCombineThreads) void k_p2g_combine(
    Dom d, LeafList L, const uint32_t *lhas, const uint32_t *vstart, float *vx, float *vy,
    float *vz, float *wx, float *wy, float *wz, uint8_t *vflag, uint8_t *marker,
    const float *ring, const uint8_t *ring_t, int64_t rstride, float *wsx, float *wsy,
    float *wsz) {
  for_each_leaf(L.list, L.count, L.n, [&](int64_t leaf) {
    const int t = threadIdx.x;
    const int lx = t >> 4, ly = (t >> 1) & 7, lz0 = (t & 1) * 4;
    const int64_t v0 = (leaf << 9) | (lx << 6) | (ly << 3) | lz0;
    float vw[3][4] = {}, w[3][4] = {};
    uint32_t touch = 0;  // one byte per voxel
    float *const V[3] = {vx, vy, vz};
    float *const W[3] = {wx, wy, wz};
    if (lhas[leaf]) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float4 x = *reinterpret_cast<const float4 *>(V[a] + v0);
        vw[a][0] = x.x; vw[a][1] = x.y; vw[a][2] = x.z; vw[a][3] = x.w;
        if (!ADJ) {
          const float4 y = *reinterpret_cast<const float4 *>(W[a] + v0);
          w[a][0] = y.x; w[a][1] = y.y; w[a][2] = y.z; w[a][3] = y.w;
        }
      }
      if (!ADJ) touch = *reinterpret_cast<const uint32_t *>(vflag + v0);
    }
    const int64_t l2 = d.L[2], l12 = int64_t(d.L[1]) * d.L[2];
    const int L[3] = {int(leaf / l12), int((leaf / l2) % d.L[1]), int(leaf % l2)};
    // Neighbour leaves holding particles: bit (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)
    // (lanes 0-26 of each wavefront test one neighbour each).
    uint32_t nbm;
    {
      const int j = int(threadIdx.x & 63);
      bool has = false;
      if (j < 27 && j != 13) {
        const int dx = j / 9 - 1, dy = (j / 3) % 3 - 1, dz = j % 3 - 1;
        if (L[0] + dx >= 0 && L[0] + dx < d.L[0] && L[1] + dy >= 0 && L[1] + dy < d.L[1] &&
            L[2] + dz >= 0 && L[2] + dz < d.L[2])
          has = lhas[leaf + dx * l12 + dy * l2 + dz] != 0;
      }
      nbm = uint32_t(__ballot(has));
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int lc[3] = {lx, ly, lz0 + u};
      // Per axis the neighbour offsets this voxel reaches, in increasing order:
      // option 0 is -1 on a lower face, else 0; option 1 (faces only) is 0 or +1.
      int o0[3];
      bool face[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        face[a] = lc[a] == 0 || lc[a] == 7;
        o0[a] = lc[a] == 0 ? -1 : 0;
      }
      // The loads of all (up to 7) neighbours first, then the sums in (dx, dy, dz) order.
      float rv[8][6];
      uint32_t rt[8];
      bool ok[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int ix = q >> 2, iy = (q >> 1) & 1, iz = q & 1;
        const int dx = o0[0] + ix, dy = o0[1] + iy, dz = o0[2] + iz;
        ok[q] = (ix == 0 || face[0]) && (iy == 0 || face[1]) && (iz == 0 || face[2]) &&
                (((nbm >> ((dx + 1) * 9 + (dy + 1) * 3 + (dz + 1))) & 1u) != 0);
        if (ok[q]) {
          const int64_t nb = leaf + dx * l12 + dy * l2 + dz;
          const int64_t e =
              nb * kRing + ring_index(lc[0] + 1 - 8 * dx, lc[1] + 1 - 8 * dy, lc[2] + 1 - 8 * dz);
#pragma unroll
          for (int a = 0; a < 3; ++a) rv[q][a] = ring[a * rstride + e];
          if (!ADJ) {
#pragma unroll
            for (int a = 0; a < 3; ++a) rv[q][3 + a] = ring[(3 + a) * rstride + e];
            rt[q] = ring_t[e];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (!ok[q]) continue;
#pragma unroll
        for (int a = 0; a < 3; ++a) vw[a][u] += rv[q][a];
        if (!ADJ) {
#pragma unroll
          for (int a = 0; a < 3; ++a) w[a][u] += rv[q][3 + a];
          touch |= rt[q] << (8 * u);
        }
      }
    }
    if (ADJ) {
#pragma unroll
      for (int a = 0; a < 3; ++a)
        *reinterpret_cast<float4 *>(V[a] + v0) = make_float4(vw[a][0], vw[a][1], vw[a][2], vw[a][3]);
      return;
    }
    if (wsx) {
      float *const S[3] = {wsx, wsy, wsz};
#pragma unroll
      for (int a = 0; a < 3; ++a)
        *reinterpret_cast<float4 *>(S[a] + v0) = make_float4(w[a][0], w[a][1], w[a][2], w[a][3]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float out[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) out[u] = double(w[a][u]) < 1e-5 ? 0.0f : vw[a][u] / w[a][u];
      *reinterpret_cast<float4 *>(V[a] + v0) = make_float4(out[0], out[1], out[2], out[3]);
    }
    uint32_t flag = 0, mark = 0;
    int o[3];
    leaf_origin(d, leaf, o);
    // Sort indices of the four voxels are consecutive: five range bounds.
    const int64_t sv0 = sort_index(d, v0);
    const uint4 s4 = *reinterpret_cast<const uint4 *>(vstart + sv0);
    const uint32_t sb[5] = {s4.x, s4.y, s4.z, s4.w, vstart[sv0 + 4]};
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if ((touch >> (8 * u)) & 0xFF) flag |= 1u << (8 * u);
      if (in_bbox(d, o[0] + lx, o[1] + ly, o[2] + lz0 + u) && sb[u + 1] > sb[u])
        mark |= uint32_t(kFluid) << (8 * u);
    }
    *reinterpret_cast<uint32_t *>(vflag + v0) = flag;
    *reinterpret_cast<uint32_t *>(marker + v0) = mark;
  });
}

// ------------------------------------------------------------------ G2P
// UpdateParticlesFromGrid (flip_simulation.cc:1024-1189): per leaf with
// particles, the 10^3 x 3 samples of v and dv are staged in LDS, then each
// particle is interpolated with the reference's lerp order (k, j, i).
template <int D>
__device__ inline float interp(const float pos[3], const int c[3], const int lc[3],
                               const float *buf) {
#pragma clang fp contract(off)
  float lower[3] = {float(c[0]), float(c[1]), float(c[2])};
  int off[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (a == D) continue;
    lower[a] = lower[a] + 0.5f;
    if (pos[a] < lower[a]) { lower[a] = lower[a] - 1.0f; off[a] = -1; }
  }
  const int i = lc[0] + off[0], j = lc[1] + off[1], k = lc[2] + off[2];
  const float w0 = pos[0] - lower[0], w1 = pos[1] - lower[1], w2 = pos[2] - lower[2];
  auto at = [&](int a, int b, int cc) { return buf[(a * 10 + b) * 10 + cc]; };
  const float v000 = at(i, j, k), v001 = at(i, j, k + 1), v010 = at(i, j + 1, k),
              v011 = at(i, j + 1, k + 1), v100 = at(i + 1, j, k), v101 = at(i + 1, j, k + 1),
              v110 = at(i + 1, j + 1, k), v111 = at(i + 1, j + 1, k + 1);
  float tt;
  tt = w2 * (v001 - v000); const float v00 = v000 + tt;
  tt = w2 * (v011 - v010); const float v01 = v010 + tt;
  tt = w2 * (v101 - v100); const float v10 = v100 + tt;
  tt = w2 * (v111 - v110); const float v11 = v110 + tt;
  tt = w1 * (v01 - v00); const float v0 = v00 + tt;
  tt = w1 * (v11 - v10); const float v1 = v10 + tt;
  tt = w0 * (v1 - v0);
  return v0 + tt;
}

__global__ __launch_bounds__(256) void k_g2p(Dom d, LeafList L,
                                             const uint32_t *lhas, const uint32_t *vstart,
                                             Particles P, const float *vx, const float *vy,
                                             const float *vz, const float *ux, const float *uy,
                                             const float *uz, float flip) {
#pragma clang fp contract(off)
  __shared__ float s_v[3][1000];
  __shared__ float s_u[3][1000];
  for_each_leaf(L.list, L.count, L.n, [&](int64_t leaf) {
#pragma clang fp contract(off)
    // The previous leaf's LDS reads are done before the tile is refilled.
    __syncthreads();
    if (!lhas[leaf]) return;
    int o[3];
    leaf_origin(d, leaf, o);
    for (int idx = threadIdx.x; idx < 1000; idx += blockDim.x) {
      const int i = o[0] - 1 + idx / 100, j = o[1] - 1 + (idx / 10) % 10, k = o[2] - 1 + idx % 10;
      float a[6] = {0, 0, 0, 0, 0, 0};
      if (in_store(d, i, j, k)) {
        const int64_t v = vox_of(d, i, j, k);
        a[0] = vx[v]; a[1] = vy[v]; a[2] = vz[v]; a[3] = ux[v]; a[4] = uy[v]; a[5] = uz[v];
      }
      s_v[0][idx] = a[0]; s_v[1][idx] = a[1]; s_v[2][idx] = a[2];
      s_u[0][idx] = a[3]; s_u[1][idx] = a[4]; s_u[2][idx] = a[5];
    }
    __syncthreads();
    const int64_t sb = sort_begin(d, leaf);
    const uint32_t b = __builtin_amdgcn_readfirstlane(vstart[sb]),
                   e = __builtin_amdgcn_readfirstlane(vstart[sb + kLeaf]);
    // Each thread takes four consecutive particles per trip, read with one
    // 16-byte load per array through descriptors based at the leaf's first
    // particle (the arrays carry kParticlePad spare elements, so the loads
    // past the leaf's last particle stay inside them) and written with one
    // 16-byte store per array when all four are in the leaf and in the box.
    // Particles outside the box (removed by DeleteOutsideParticles) keep
    // their values, so a quad holding one is written particle by particle.
    const int bytes = int((e - b + 3) * 4);
    __amdgpu_buffer_rsrc_t R[9];
#pragma unroll
    for (int a = 0; a < 9; ++a)
      R[a] = __builtin_amdgcn_make_buffer_rsrc(P.a[a] + b, 0, bytes, kRawBufferWord3);
    for (uint32_t q0 = b + 4 * threadIdx.x; q0 < e; q0 += 4 * blockDim.x) {
      const int off = int(q0 - b) * 4;
      float pp[4][3], pv[4][3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const u32x4 x = __builtin_amdgcn_raw_buffer_load_b128(R[a], off, 0, 0);
        const u32x4 y = __builtin_amdgcn_raw_buffer_load_b128(R[3 + a], off, 0, 0);
        pp[0][a] = __uint_as_float(x.x); pp[1][a] = __uint_as_float(x.y);
        pp[2][a] = __uint_as_float(x.z); pp[3][a] = __uint_as_float(x.w);
        pv[0][a] = __uint_as_float(y.x); pv[1][a] = __uint_as_float(y.y);
        pv[2][a] = __uint_as_float(y.z); pv[3][a] = __uint_as_float(y.w);
      }
      float nv[4][3], ng[4][3];
      bool all_in = q0 + 3 < e;
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        const float *pos = pp[u];
        const int c[3] = {int(floorf(pos[0])), int(floorf(pos[1])), int(floorf(pos[2]))};
        const bool in = q0 + u < e && in_bbox(d, c[0], c[1], c[2]);
        all_in = all_in && in;
        const int lc[3] = {c[0] - o[0] + 1, c[1] - o[1] + 1, c[2] - o[2] + 1};
        if (!in) {
#pragma unroll
          for (int a = 0; a < 3; ++a) { nv[u][a] = 0.0f; ng[u][a] = 0.0f; }
          continue;
        }
        const float gv[3] = {interp<0>(pos, c, lc, s_v[0]), interp<1>(pos, c, lc, s_v[1]),
                             interp<2>(pos, c, lc, s_v[2])};
        const float dv[3] = {interp<0>(pos, c, lc, s_u[0]), interp<1>(pos, c, lc, s_u[1]),
                             interp<2>(pos, c, lc, s_u[2])};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float s1 = dv[a] + pv[u][a];
          const float s2 = s1 - gv[a];
          const float s3 = flip * s2;
          nv[u][a] = gv[a] + s3;
          ng[u][a] = gv[a];
        }
      }
      if (all_in) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const u32x4 x = {__float_as_uint(nv[0][a]), __float_as_uint(nv[1][a]),
                           __float_as_uint(nv[2][a]), __float_as_uint(nv[3][a])};
          const u32x4 y = {__float_as_uint(ng[0][a]), __float_as_uint(ng[1][a]),
                           __float_as_uint(ng[2][a]), __float_as_uint(ng[3][a])};
          __builtin_amdgcn_raw_buffer_store_b128(x, R[3 + a], off, 0, 0);
          __builtin_amdgcn_raw_buffer_store_b128(y, R[6 + a], off, 0, 0);
        }
      } else {
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          const uint32_t q = q0 + u;
          if (q >= e) break;
          const int c[3] = {int(floorf(pp[u][0])), int(floorf(pp[u][1])), int(floorf(pp[u][2]))};
          if (!in_bbox(d, c[0], c[1], c[2])) continue;
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            P.a[3 + a][q] = nv[u][a];
            P.a[6 + a][q] = ng[u][a];
          }
        }
      }
    }
  });
}

// Transpose of P2G's normalised transfer: for every particle (outside the
// partition box too, since P2G ran before DeleteOutsideParticles)
//   out_D = keep * in_D + sum_faces w_D(p, face) * g_D(face) / wsum_D(face),
// with g / wsum taken as 0 where wsum < 1e-5 (the faces P2G set to 0). The
// weights are the trilinear ones of the forward transfer, evaluated with the
// G2P interpolation.
__global__ __launch_bounds__(256) void k_p2g_transpose(
    Dom d, LeafList L, const uint32_t *lhas, const uint32_t *vstart,
    Particles P, const float *gx, const float *gy, const float *gz, const float *wx,
    const float *wy, const float *wz, float keep, Particles io) {
#pragma clang fp contract(off)
  __shared__ float s_g[3][1000];
  for_each_leaf(L.list, L.count, L.n, [&](int64_t leaf) {
#pragma clang fp contract(off)
    // The previous leaf's LDS reads are done before the tile is refilled.
    __syncthreads();
    if (!lhas[leaf]) return;
    int o[3];
    leaf_origin(d, leaf, o);
    for (int idx = threadIdx.x; idx < 1000; idx += blockDim.x) {
      const int i = o[0] - 1 + idx / 100, j = o[1] - 1 + (idx / 10) % 10, k = o[2] - 1 + idx % 10;
      float a[3] = {0, 0, 0};
      if (in_store(d, i, j, k)) {
        const int64_t v = vox_of(d, i, j, k);
        const float g[3] = {gx[v], gy[v], gz[v]}, ws[3] = {wx[v], wy[v], wz[v]};
#pragma unroll
        for (int c = 0; c < 3; ++c) a[c] = double(ws[c]) < 1e-5 ? 0.0f : g[c] / ws[c];
      }
      s_g[0][idx] = a[0]; s_g[1][idx] = a[1]; s_g[2][idx] = a[2];
    }
    __syncthreads();
    const uint32_t b = vstart[sort_begin(d, leaf)], e = vstart[sort_begin(d, leaf) + kLeaf];
    for (uint32_t q = b + threadIdx.x; q < e; q += blockDim.x) {
      const float pos[3] = {P.a[0][q], P.a[1][q], P.a[2][q]};
      const int c[3] = {int(floorf(pos[0])), int(floorf(pos[1])), int(floorf(pos[2]))};
      const int lc[3] = {c[0] - o[0] + 1, c[1] - o[1] + 1, c[2] - o[2] + 1};
      const float t[3] = {interp<0>(pos, c, lc, s_g[0]), interp<1>(pos, c, lc, s_g[1]),
                          interp<2>(pos, c, lc, s_g[2])};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float m = keep * io.a[a][q];
        io.a[a][q] = m + t[a];
      }
    }
  });
}

__global__ void k_cell_counts_box(Dom d, const uint32_t *vstart, int32_t *out) {
  const int bx = d.n[0] + 2, by = d.n[1] + 2, bz = d.n[2] + 2;
  const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= int64_t(bx) * by * bz) return;
  const int i = int(q / (int64_t(by) * bz)) - 1, j = int((q / bz) % by) - 1, k = int(q % bz) - 1;
  const int64_t v = sort_index(d, vox_of(d, i, j, k));
  out[q] = int32_t(vstart[v + 1] - vstart[v]);
}

inline unsigned blocks_for(int64_t n, int t) { return unsigned((n + t - 1) / t); }

}  // namespace

hipError_t launch_seed(fsl_ctx *c, const uint8_t *mask_dev, const uint32_t *offs, int ppc,
                       uint32_t seed, const float vel[3]) {
  const int64_t ncell = int64_t(c->dom.n[0]) * c->dom.n[1] * c->dom.n[2];
  k_seed<<<blocks_for(ncell, kThreads), kThreads, 0, c->stream>>>(
      c->dom, mask_dev, offs, ppc, seed, vel[0], vel[1], vel[2], c->P);
  return hipGetLastError();
}

hipError_t launch_mask_count(fsl_ctx *c, const uint8_t *mask_dev, int64_t n, uint32_t *cnt) {
  k_mask_to_count<<<blocks_for(n, kThreads), kThreads, 0, c->stream>>>(mask_dev, n, cnt);
  return hipGetLastError();
}

hipError_t launch_advect(fsl_ctx *c, float dt, const float *dt_dev, bool drop_outside,
                         const KillBoxes &kill) {
  if (c->np == 0) return hipSuccess;
  const uint32_t *np_dev = c->counts_on_device ? c->dcount + kCountParticles : nullptr;
  k_advect<<<blocks_for(c->np, kThreads), kThreads, 0, c->stream>>>(
      c->dom, c->np, np_dev, c->P, dt, dt_dev, c->solid, c->key, c->vcount,
      drop_outside ? 1 : 0, kill);
  return hipGetLastError();
}

hipError_t launch_source_count(fsl_ctx *c, const SourceBox &s, int64_t ncell, uint32_t *cnt) {
  k_source_count<<<blocks_for(ncell, kThreads), kThreads, 0, c->stream>>>(c->dom, s, ncell, cnt);
  return hipGetLastError();
}

hipError_t launch_source_emit(fsl_ctx *c, const SourceBox &s, int64_t ncell, const uint32_t *off,
                              int64_t base) {
  k_source_emit<<<blocks_for(ncell, kThreads), kThreads, 0, c->stream>>>(c->dom, s, ncell, off,
                                                                          c->P, base);
  return hipGetLastError();
}

// After the advection kernel and the scan of vcount into vstart: scatter,
// restore in-cell order, gather into the other buffer and swap. Without
// keep_gv the grid velocities are left behind (stale until G2P writes them).
// With the counts on the device, n_alive is not known to the host: the
// sorted count is read from vstart's last entry by the kernels, the launches
// are sized by the count before the sort (an upper bound, kept in c->np),
// and the new count is stored in dcount.
hipError_t launch_sort_tail(fsl_ctx *c, uint32_t n_alive, bool keep_gv) {
  const bool dev = c->counts_on_device;
  const uint32_t *np_dev = dev ? c->dcount + kCountParticles : nullptr;
  const uint32_t *na_dev = dev ? c->vstart + c->nvox : nullptr;
  uint32_t *out = dev ? c->dcount + kCountParticles : nullptr;
  const int64_t nb = dev ? c->np : int64_t(n_alive);
  if (c->np > 0) {
    k_scatter<<<blocks_for((c->np + 3) / 4, kThreads), kThreads, 0, c->stream>>>(
        c->np, np_dev, c->key, c->vstart, c->vcount, c->perm);
  }
  if (nb > 0) {
    k_cell_sort<<<unsigned(c->nleaves), 512, 0, c->stream>>>(c->vstart, c->perm);
    if (keep_gv)
      k_gather<9><<<blocks_for((nb + 3) / 4, kThreads), kThreads, 0, c->stream>>>(n_alive, na_dev, out,
                                                                         c->perm, c->P, c->Q);
    else
      k_gather<6><<<blocks_for((nb + 3) / 4, kThreads), kThreads, 0, c->stream>>>(n_alive, na_dev, out,
                                                                         c->perm, c->P, c->Q);
  }
  Particles t = c->P;
  c->P = c->Q;
  c->Q = t;
  if (!dev) c->np = n_alive;
  return hipGetLastError();
}

// Active leaves: holds particles or neighbours a leaf that does.
hipError_t launch_leaf_flags(fsl_ctx *c) {
  const int64_t nl = c->nleaves;
  k_leaf_has<<<blocks_for(nl, kThreads), kThreads, 0, c->stream>>>(nl, c->lrank, c->vstart,
                                                                   c->lhas);
  k_leaf_dilate<<<blocks_for(nl, kThreads), kThreads, 0, c->stream>>>(c->dom, c->lhas, c->lact);
  return launch_compact(c, c->lact, c->act_list, c->dcount + kCountActive);
}

// Compacts the flagged leaves into `list` in Morton order; the count is
// left in lscan[nleaves] and stored in *count.
hipError_t launch_compact(fsl_ctx *c, const uint32_t *flag, int32_t *list, uint32_t *count) {
  const int64_t nl = c->nleaves;
  k_gather_flags<<<blocks_for(nl, kThreads), kThreads, 0, c->stream>>>(nl, c->morton, flag,
                                                                        c->lflag_m);
  hipError_t e = exclusive_scan(c->lflag_m, c->lscan, nl, c->scan_tmp, c->stream);
  if (e != hipSuccess) return e;
  k_compact<<<blocks_for(nl, kThreads), kThreads, 0, c->stream>>>(nl, c->morton, c->lflag_m,
                                                                   c->lscan, list, count);
  return hipGetLastError();
}

// Ring blocks: six float planes (vw of x, y, z, then w of x, y, z) of
// nleaves * kRing entries, then the touch bytes.
size_t p2g_ring_bytes(int64_t nleaves) {
  return size_t(nleaves) * kRing * (6 * sizeof(float) + 1);
}

hipError_t launch_p2g(fsl_ctx *c, float *const wsum[3]) {
  if (c->n_act == 0) return hipSuccess;
  const int64_t rs = c->nleaves * kRing;
  float *ring = static_cast<float *>(c->p2g_ring);
  uint8_t *ring_t = reinterpret_cast<uint8_t *>(ring + 6 * rs);
  float *w[3] = {c->p2g_w, c->p2g_w + c->nvox, c->p2g_w + 2 * c->nvox};
  k_p2g_own<false><<<leaf_grid(c->n_act, c->counts_on_device), kP2gThreads, 0, c->stream>>>(
      c->dom, act_leaves(c), c->lhas, c->vstart, c->P, c->v[0], c->v[1], c->v[2], w[0],
      w[1], w[2], c->vflag, ring, ring_t, rs);
  k_p2g_combine<false><<<leaf_grid(c->n_act, c->counts_on_device), kCombineThreads, 0, c->stream>>>(
      c->dom, act_leaves(c), c->lhas, c->vstart, c->v[0], c->v[1], c->v[2], w[0], w[1],
      w[2], c->vflag, c->marker, ring, ring_t, rs, wsum ? wsum[0] : nullptr,
      wsum ? wsum[1] : nullptr, wsum ? wsum[2] : nullptr);
  return hipGetLastError();
}

hipError_t launch_g2p_transpose(fsl_ctx *c, float *const pval[3], float *const out[3]) {
  if (c->n_act == 0) return hipSuccess;
  Particles A = c->P;
  for (int a = 0; a < 3; ++a) A.a[3 + a] = pval[a];
  const int64_t rs = c->nleaves * kRing;
  float *ring = static_cast<float *>(c->p2g_ring);
  k_p2g_own<true><<<leaf_grid(c->n_act, c->counts_on_device), kP2gThreads, 0, c->stream>>>(
      c->dom, act_leaves(c), c->lhas, c->vstart, A, out[0], out[1], out[2], nullptr,
      nullptr, nullptr, nullptr, ring, nullptr, rs);
  k_p2g_combine<true><<<leaf_grid(c->n_act, c->counts_on_device), kCombineThreads, 0, c->stream>>>(
      c->dom, act_leaves(c), c->lhas, c->vstart, out[0], out[1], out[2], nullptr,
      nullptr, nullptr, nullptr, nullptr, ring, nullptr, rs, nullptr, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_g2p(fsl_ctx *c) {
  if (c->n_act == 0) return hipSuccess;
  k_g2p<<<leaf_grid(c->n_act, c->counts_on_device), 256, 0, c->stream>>>(
      c->dom, act_leaves(c), c->lhas, c->vstart, c->P, c->v[0], c->v[1], c->v[2], c->u[0],
      c->u[1], c->u[2], c->flip);
  return hipGetLastError();
}

hipError_t launch_p2g_transpose(fsl_ctx *c, float *const g[3], float *const wsum[3], float keep,
                                float *const io[3]) {
  if (c->n_act == 0) return hipSuccess;
  Particles A{};
  for (int a = 0; a < 3; ++a) A.a[a] = io[a];
  k_p2g_transpose<<<leaf_grid(c->n_act, c->counts_on_device), 256, 0, c->stream>>>(
      c->dom, act_leaves(c), c->lhas, c->vstart, c->P, g[0], g[1], g[2], wsum[0],
      wsum[1], wsum[2], keep, A);
  return hipGetLastError();
}

hipError_t launch_cell_counts_box(fsl_ctx *c, int32_t *out_dev) {
  const int64_t nb = int64_t(c->dom.n[0] + 2) * (c->dom.n[1] + 2) * (c->dom.n[2] + 2);
  k_cell_counts_box<<<blocks_for(nb, kThreads), kThreads, 0, c->stream>>>(c->dom, c->vstart,
                                                                           out_dev);
  return hipGetLastError();
}

}  // namespace fsl
