/*
 * fsl.hpp - C++ host classes over the C ABI in fsl.h.
 *
 * These classes keep the call surface of the reference's FLIP step
 * (schinmayee/fluid-sim-lb):
 *   fsl::FlipSimulation  <- application::FlipSimulation<3,3,float>
 *                           (common/flip_simulation.h:39-187)
 *   fsl::PCGDriver       <- common::PCGDriver<ScalarGrid<3,3,float>>
 *                           (common/solver_driver.h:30-106)
 *   fsl::FlipDriver      <- FlipDriver's step loop (common/flip_app_sim.cc:211-254,
 *                           357-462), run in order in one process: the Canary
 *                           transforms of the reference become plain calls.
 *
 * Differences from the reference, all forced by the device-resident design:
 *  - The grids and particles live in the context on the GPU; the reference's
 *    set_grid_*() / grid_*() pointer plumbing has no counterpart, and the
 *    grids are read and written through Get/Set calls in the box format of
 *    fsl.h.
 *  - The reference aborts on a failed CHECK; these classes throw fsl::Error
 *    with the library's message and status code instead.
 *  - Particle seeding uses Philox4x32-10 keyed by (seed, cell, slot) on cell
 *    centres inside the shape, instead of glibc random() and a level set.
 *  - DeleteParticlesInSolids runs inside MoveParticles (one sort per step).
 */
#ifndef FSL_HPP_
#define FSL_HPP_

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <vector>

#include "fsl.h"

namespace fsl {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

struct Coord {
  int x = 0, y = 0, z = 0;
  Coord() = default;
  Coord(int a, int b, int c) : x(a), y(b), z(c) {}
  int operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

struct Vec3 {
  float x = 0, y = 0, z = 0;
  Vec3() = default;
  Vec3(float a, float b, float c) : x(a), y(b), z(c) {}
};

/* Inclusive integer box, as common::CoordBBox. */
struct CoordBBox {
  Coord lo, hi;
};

/* Cube with inclusive bounds in voxel units (common/primitives.cc:56-64). */
struct Cube {
  Vec3 start, end;
};

/* common::PreconditionerType (common/definitions.h:7). */
enum PreconditionerType { IC, IC0, JACOBI, NUM_PRECONDITIONER_TYPES };

/* application::Source (common/sources.h:11-76, sources.cc): a water source,
 * static, timed or oscillating. Behaviour kept as the reference has it:
 *  - get_box returns box1 with its end reduced by 1 whether or not the box
 *    oscillates (the oscillating branch of sources.cc:64-82 fills a local
 *    that shadows the returned box);
 *  - an oscillating velocity uses the box period (sources.cc:88).
 * Boxes are in normalised coordinates until ScaleBox. */
class Source {
 public:
  Source() = default;
  Source(const Cube &box, const Vec3 &velocity) : box1_(box), velocity1_(velocity) {}

  void set_box(const Cube &box) {
    box1_ = box;
    box_oscillating_ = false;
  }
  void set_velocity(const Vec3 &velocity) {
    velocity1_ = velocity;
    velocity_oscillating_ = false;
  }
  void set_box(const Cube &box1, const Cube &box2, float period) {
    box1_ = box1;
    box2_ = box2;
    box_period_ = period;
    box_oscillating_ = true;
  }
  void set_velocity(const Vec3 &velocity1, const Vec3 &velocity2, float period) {
    velocity1_ = velocity1;
    velocity2_ = velocity2;
    velocity_period_ = period;
    velocity_oscillating_ = true;
  }
  void set_timed(float start, float end) {
    start_time_ = start;
    end_time_ = end;
    timed_ = true;
  }

  Cube get_box(float t = 0) const {
    (void)t;
    Cube box = box1_;
    box.end = Vec3(box.end.x - 1, box.end.y - 1, box.end.z - 1);
    return box;
  }
  Vec3 get_velocity(float t = 0) const {
    if (!velocity_oscillating_) return velocity1_;
    const float w = float(M_PI) * t / box_period_;
    const float c = std::cos(w) + 1;
    return Vec3(velocity1_.x + 0.5f * (velocity2_.x - velocity1_.x) * c,
                velocity1_.y + 0.5f * (velocity2_.y - velocity1_.y) * c,
                velocity1_.z + 0.5f * (velocity2_.z - velocity1_.z) * c);
  }
  bool is_active(float t = 0) const {
    return !timed_ || (t >= start_time_ && t <= end_time_);
  }

  /* Cube::Scale (primitives.cc:76-83) of both boxes. */
  void ScaleBox(const Coord scale) {
    for (Cube *b : {&box1_, &box2_}) {
      b->start = Vec3(b->start.x * scale.x, b->start.y * scale.y, b->start.z * scale.z);
      b->end = Vec3(b->end.x * scale.x, b->end.y * scale.y, b->end.z * scale.z);
    }
  }
  void ScaleVelocity(float scale) {
    velocity1_ = Vec3(velocity1_.x * scale, velocity1_.y * scale, velocity1_.z * scale);
    velocity2_ = Vec3(velocity2_.x * scale, velocity2_.y * scale, velocity2_.z * scale);
  }

 private:
  Cube box1_, box2_;
  Vec3 velocity1_, velocity2_;
  float box_period_ = 0, velocity_period_ = 0;
  float start_time_ = 0, end_time_ = 0;
  bool box_oscillating_ = false, velocity_oscillating_ = false, timed_ = false;
};

class FlipSimulation {
 public:
  FlipSimulation() = default;
  FlipSimulation(const FlipSimulation &) = delete;
  FlipSimulation &operator=(const FlipSimulation &) = delete;
  ~FlipSimulation() {
    if (ctx_) fsl_destroy(ctx_);
  }

  /* FlipSimulation::InitializeMetadata (flip_simulation.cc:83-108). rank is
   * kept for the interface; one context holds every partition as a tile.
   * gravity is the scaled gravity (flip_app_sim.cc:74). */
  void InitializeMetadata(Coord global_dims, Coord partitions, int rank = 0,
                          float voxel_len = 1, int num_particles_per_cell = 8,
                          float gravity = 9.8f, float flip_factor = 1,
                          std::string output_dir = "output", bool debug_flag = false) {
    output_dir_ = output_dir;
    debug_ = debug_flag;
    frame_ = 0;
    debug_step_ = 0;
    if (std::fabs(voxel_len - 1.0f) > 1e-2f)
      throw Error(FSL_ERR_ARG, "voxel_len must be 1 (flip_simulation.cc:99)");
    dims_ = global_dims;
    partitions_ = partitions;
    rank_ = rank;
    ppc_ = num_particles_per_cell;
    gravity_ = gravity;
    flip_ = flip_factor;
    force_ = Vec3(0, 1, 0);
  }

  /* FlipSimulation::InitializeVariables (flip_simulation.cc:110-122): creates
   * the device context. capacity = most particles the context may hold
   * (0: ppc x cells / 4, at least 1M). */
  void InitializeVariables(int device = 0, int64_t capacity = 0) {
    const int d[3] = {dims_.x, dims_.y, dims_.z};
    const int t[3] = {partitions_.x, partitions_.y, partitions_.z};
    if (capacity <= 0) {
      capacity = int64_t(ppc_) * d[0] * d[1] * d[2] / 4;
      if (capacity < (1 << 20)) capacity = 1 << 20;
    }
    check(fsl_create(&ctx_, d, t, device, capacity));
    mask_.assign(size_t(d[0]) * d[1] * d[2], 0);
    push_params();
  }

  /* InitializeWaterCuboid / InitializeWaterSphere / InitializeReservoir /
   * InitializeMultipleWaterDrops (flip_simulation.cc:124-188): each call adds
   * the shape's cells (cell-centre test) to the fluid mask and reseeds all
   * particles with `velocity`. */
  void InitializeWaterCuboid(const Cube &box, const Vec3 velocity) {
    add_shape([&](float x, float y, float z) {
      return x >= box.start.x && x <= box.end.x && y >= box.start.y && y <= box.end.y &&
             z >= box.start.z && z <= box.end.z;
    }, velocity);
  }
  void InitializeWaterSphere(const Vec3 &center, const float radius, const Vec3 velocity) {
    add_shape([&](float x, float y, float z) {
      const float dx = x - center.x, dy = y - center.y, dz = z - center.z;
      return dx * dx + dy * dy + dz * dz < radius * radius;
    }, velocity);
  }
  void InitializeReservoir(float height) {
    add_shape([&](float, float y, float) { return y < height; }, Vec3());
  }
  void InitializeMultipleWaterDrops(float height, const std::vector<Vec3> &centers,
                                    const std::vector<float> &radii) {
    InitializeReservoir(height);
    for (size_t i = 0; i < centers.size() && i < radii.size(); ++i)
      InitializeWaterSphere(centers[i], radii[i], Vec3());
  }
  /* Seed for the Philox particle placement (no reference counterpart). */
  void set_seed(uint32_t seed) { seed_ = seed; }

  /* SetForces (flip_simulation.cc:248-253): force = f * gravity. */
  void SetForces(const Vec3 &f) {
    force_ = f;
    push_params();
  }
  void SetBoxBoundary() { check(fsl_set_box_boundary(ctx_, 0)); }
  void SetTankBoundary() { check(fsl_set_box_boundary(ctx_, 1)); }
  void AddSolidCuboid(const CoordBBox &box) {
    const int lo[3] = {box.lo.x, box.lo.y, box.lo.z}, hi[3] = {box.hi.x, box.hi.y, box.hi.z};
    check(fsl_add_solid_box(ctx_, lo, hi));
  }

  /* AddWaterSource (flip_simulation.h): a source, already scaled
   * (flip_app_sim.cc:98-104). */
  void AddWaterSource(const Source &source) { sources_.push_back(source); }
  /* AddContributionFromSources (flip_simulation.cc:1206-1266): particles of
   * every source active at time t. Each call draws a new Philox stream. */
  void AddContributionFromSources(float t) {
    for (const Source &src : sources_) {
      if (!src.is_active(t)) continue;
      const Cube box = src.get_box(t);
      const Vec3 v = src.get_velocity(t);
      const float lo[3] = {box.start.x, box.start.y, box.start.z};
      const float hi[3] = {box.end.x, box.end.y, box.end.z};
      const float vel[3] = {v.x, v.y, v.z};
      check(fsl_add_source(ctx_, lo, hi, vel, ppc_, seed_, source_stream_++));
    }
  }
  /* DeleteFluidInSources (flip_simulation.cc:1268-1277): the particles in
   * every source's box (active or not). */
  void DeleteFluidInSources(float t) {
    std::vector<float> boxes;
    for (const Source &src : sources_) {
      const Cube b = src.get_box(t);
      boxes.insert(boxes.end(), {b.start.x, b.start.y, b.start.z, b.end.x, b.end.y, b.end.z});
    }
    if (!boxes.empty())
      check(fsl_delete_particles_in_boxes(ctx_, int(boxes.size() / 6), boxes.data()));
  }
  bool HasSources() const { return !sources_.empty(); }

  void ClearGridData() { check(fsl_clear(ctx_)); }
  size_t FluidCellCount() {
    std::vector<int32_t> m(box_cells());
    check(fsl_get_int(ctx_, FSL_GRID_MARKER, m.data()));
    size_t n = 0;
    for (int32_t v : m) n += v == 1;
    return n;
  }
  size_t ParticleCount() {
    int64_t n = 0;
    check(fsl_num_particles(ctx_, &n));
    return size_t(n);
  }

  /* CFL (flip_simulation.cc:459-488). */
  float CFL() {
    float s = 0;
    check(fsl_cfl(ctx_, &s));
    return s;
  }
  void SaveGridVelocities() { check(fsl_save_velocities(ctx_)); }
  /* MoveParticles (flip_simulation.cc:498-503) + DeleteParticlesInSolids. */
  void MoveParticles(float dt, bool defragment = true) {
    (void)defragment;
    check(fsl_move_particles(ctx_, dt));
  }
  void AddGravity(float dt) { check(fsl_add_gravity(ctx_, dt)); }
  void ApplyBoundaryConditions() { check(fsl_apply_bc(ctx_)); }
  void ComputeDivergence() { check(fsl_divergence(ctx_)); }
  void UpdateVelocityFromPressure() { check(fsl_pressure_update(ctx_)); }
  void GetVelocityUpdate() { check(fsl_velocity_update(ctx_)); }
  void ExtendGridVelocityIntoBoundaryRegion() { check(fsl_extrapolate(ctx_)); }
  void TransferParticlesToGrid() { check(fsl_p2g(ctx_)); }
  /* UpdateParticlesFromGrid (flip_simulation.cc:1024-1189); extrapolates first. */
  void UpdateParticlesFromGrid() { check(fsl_g2p(ctx_)); }
  /* Runs inside MoveParticles here; kept for the call sequence. */
  void DeleteParticlesInSolids() {}
  void DeleteOutsideParticles() { check(fsl_delete_outside(ctx_)); }

  /* Grid access (box format of fsl.h). */
  std::vector<float> GetVelocity(int which, std::vector<uint8_t> *active = nullptr) {
    std::vector<float> v(box_cells() * 3);
    if (active) active->resize(box_cells());
    check(fsl_get_vec(ctx_, which, v.data(), active ? active->data() : nullptr));
    return v;
  }
  void SetVelocity(int which, const std::vector<float> &v, const std::vector<uint8_t> &active) {
    if (v.size() != box_cells() * 3 || active.size() != box_cells())
      throw Error(FSL_ERR_ARG, "velocity grid size");
    check(fsl_set_vec(ctx_, which, v.data(), active.data()));
  }
  std::vector<float> GetScalar(int which) {
    std::vector<float> s(box_cells());
    check(fsl_get_scalar(ctx_, which, s.data()));
    return s;
  }
  std::vector<int32_t> GetInt(int which) {
    std::vector<int32_t> s(box_cells());
    check(fsl_get_int(ctx_, which, s.data()));
    return s;
  }
  void GetParticles(std::vector<float> *pos, std::vector<float> *vel, std::vector<float> *gv) {
    const size_t n = ParticleCount();
    if (pos) pos->resize(n * 3);
    if (vel) vel->resize(n * 3);
    if (gv) gv->resize(n * 3);
    check(fsl_get_particles(ctx_, pos ? pos->data() : nullptr, vel ? vel->data() : nullptr,
                            gv ? gv->data() : nullptr));
  }

  /* Particles::Write (particles.cc:166-200), one partition covering the
   * domain: start, end (inclusive), count, then per particle its running
   * index, position and velocity; binary (int, float[3], float[3]) or text
   * ("%d : %f %f %f, %f %f %f"). Particles are written in device order
   * (sorted by cell). */
  void WriteParticles(const std::string &file_name, bool text = false) {
    std::vector<float> pos, vel;
    GetParticles(&pos, &vel, nullptr);
    FILE *fp = std::fopen(file_name.c_str(), text ? "wt" : "wb");
    if (!fp) throw Error(FSL_ERR_ARG, "could not create " + file_name);
    const int start[3] = {0, 0, 0}, end[3] = {dims_.x - 1, dims_.y - 1, dims_.z - 1};
    const int count = int(pos.size() / 3);
    if (text) {
      std::fprintf(fp, "%d %d %d\n%d %d %d\n%d\n", start[0], start[1], start[2], end[0], end[1],
                   end[2], count);
      for (int p = 0; p < count; ++p)
        std::fprintf(fp, "%d : %f %f %f, %f %f %f\n", p, pos[3 * p], pos[3 * p + 1],
                     pos[3 * p + 2], vel[3 * p], vel[3 * p + 1], vel[3 * p + 2]);
    } else {
      std::fwrite(start, sizeof(int), 3, fp);
      std::fwrite(end, sizeof(int), 3, fp);
      std::fwrite(&count, sizeof(int), 1, fp);
      for (int p = 0; p < count; ++p) {
        std::fwrite(&p, sizeof(int), 1, fp);
        std::fwrite(&pos[3 * p], sizeof(float), 3, fp);
        std::fwrite(&vel[3 * p], sizeof(float), 3, fp);
      }
    }
    std::fclose(fp);
  }

  /* VectorGrid::Write / ScalarGrid::Write, text mode (vector_grid.cc:133-160,
   * scalar_grid.cc:134-160): start and end of the partition, then
   * "i j k : values" for every active voxel in OpenVDB iteration order
   * (64^3 internal nodes by origin, then 8^3 leaves, then voxels, each
   * x-major), restricted to [start, end] (to end + 1 for the staggered
   * velocity grids).
   *   which = FSL_GRID_VELOCITY / FSL_GRID_VELOCITY_UPDATE: vector grids;
   *   which = 10 + FSL_GRID_DIVERGENCE / FSL_GRID_PRESSURE / FSL_GRID_MARKER:
   *   scalar grids (active where the reference sets them: FLUID cells for
   *   divergence and pressure, FLUID or SOLID cells for the marker). */
  void WriteGrid(const std::string &file_name, int which) {
    const int by = dims_.y + 2, bz = dims_.z + 2;
    const bool vec = which < 10;
    std::vector<float> vals;
    std::vector<uint8_t> act;
    std::vector<int32_t> marker = GetInt(FSL_GRID_MARKER);
    if (vec) {
      vals = GetVelocity(which, &act);
    } else if (which - 10 == FSL_GRID_MARKER) {
      vals.assign(marker.begin(), marker.end());
    } else {
      vals = GetScalar(which - 10);
    }
    const int ext = vec ? 1 : 0;
    std::vector<std::pair<uint64_t, size_t>> order;
    for (int i = 0; i <= dims_.x - 1 + ext; ++i)
      for (int j = 0; j <= dims_.y - 1 + ext; ++j)
        for (int k = 0; k <= dims_.z - 1 + ext; ++k) {
          const size_t q = (size_t(i + 1) * by + (j + 1)) * bz + (k + 1);
          bool on;
          if (vec) on = act[q] != 0;
          else if (which - 10 == FSL_GRID_MARKER) on = marker[q] == 1 || marker[q] == 2;
          else on = marker[q] == 1;
          if (on) order.push_back({vdb_key(i, j, k), q});
        }
    std::sort(order.begin(), order.end());
    FILE *fp = std::fopen(file_name.c_str(), "wt");
    if (!fp) throw Error(FSL_ERR_ARG, "could not create " + file_name);
    std::fprintf(fp, "%d %d %d\n%d %d %d\n", 0, 0, 0, dims_.x - 1, dims_.y - 1, dims_.z - 1);
    for (const auto &e : order) {
      const size_t q = e.second;
      const int i = int(q / (size_t(by) * bz)) - 1, j = int((q / bz) % by) - 1,
                k = int(q % bz) - 1;
      if (vec)
        std::fprintf(fp, "%d %d %d : %f %f %f\n", i, j, k, vals[3 * q], vals[3 * q + 1],
                     vals[3 * q + 2]);
      else
        std::fprintf(fp, "%d %d %d : %f\n", i, j, k, vals[q]);
    }
    std::fclose(fp);
  }

  /* OpenVDB Tree3<T,3,3> value-iteration key of voxel (i, j, k). */
  static uint64_t vdb_key(int i, int j, int k) {
    const uint64_t ri = uint64_t((i >> 6) + 1024), rj = uint64_t((j >> 6) + 1024),
                   rk = uint64_t((k >> 6) + 1024);
    const uint64_t leaf = uint64_t(((i >> 3) & 7) << 6 | ((j >> 3) & 7) << 3 | ((k >> 3) & 7));
    const uint64_t vox = uint64_t((i & 7) << 6 | (j & 7) << 3 | (k & 7));
    return (((ri * 2048 + rj) * 2048 + rk) << 18) | (leaf << 9) | vox;
  }

  /* SaveDataToFile (flip_simulation.cc:1284-1292): the binary particle
   * frame <output_dir>/particles_<frame, %08d> (the reference's prefix,
   * flip_simulation.cc:61-66; the directory is created on first use). The
   * velocity and phi grids are not written there (commented out in the
   * reference). */
  void SaveDataToFile() {
    std::filesystem::create_directories(output_dir_);
    WriteParticles(output_dir_ + "/particles_" + num8(frame_), false);
  }
  /* SaveDebugDataToFile (flip_simulation.cc:1295-1317), debug_flag set: in
   * <output_dir>/debug, particles_<step> (the text particle format, the
   * default of Particles::Write), grid_velocity_<step> and
   * grid_velocity_<step>.txt (both the text dump, as the reference's
   * ScalarGrid/VectorGrid::Write print text in either mode), and the
   * pressure grid as grid_phi_<step>.txt (the reference writes it under the
   * phi prefix). The phi grid itself (level set, not built) is not written.
   * Advances the debug step. */
  void SaveDebugDataToFile() {
    if (!debug_) throw Error(FSL_ERR_STATE, "SaveDebugDataToFile needs debug_flag");
    const std::string dir = output_dir_ + "/debug", step = num8(debug_step_);
    std::filesystem::create_directories(dir);
    WriteParticles(dir + "/particles_" + step, true);
    WriteGrid(dir + "/grid_velocity_" + step, FSL_GRID_VELOCITY);
    WriteGrid(dir + "/grid_velocity_" + step + ".txt", FSL_GRID_VELOCITY);
    WriteGrid(dir + "/grid_phi_" + step + ".txt", 10 + FSL_GRID_PRESSURE);
    ++debug_step_;
  }
  /* NextFrame (flip_simulation.cc:1320-1322): the frame number of the next
   * SaveDataToFile. */
  void NextFrame() { ++frame_; }
  int frame() const { return frame_; }
  /* Empty in the reference (flip_simulation.cc:518-521, 642-644, 1192-1194). */
  void AdvectPhi() {}
  void ExtendVelocity() {}
  void ResampleParticles() {}

  Coord global_dims() const { return dims_; }
  Coord partitions() const { return partitions_; }
  float flip_factor() const { return flip_; }
  fsl_ctx *context() { return ctx_; }

  void check(int rc) const {
    if (rc != FSL_OK) throw Error(rc, ctx_ ? fsl_last_error(ctx_) : "fsl: no context");
  }

 private:
  static std::string num8(int n) {
    char buf[16];
    std::snprintf(buf, sizeof(buf), "%08d", n);
    return buf;
  }
  size_t box_cells() const {
    return size_t(dims_.x + 2) * (dims_.y + 2) * (dims_.z + 2);
  }
  void push_params() {
    if (!ctx_) return;
    const float f[3] = {force_.x * gravity_, force_.y * gravity_, force_.z * gravity_};
    check(fsl_set_params(ctx_, gravity_, f, flip_));
  }
  template <typename Inside>
  void add_shape(Inside inside, const Vec3 &velocity) {
    if (!ctx_) throw Error(FSL_ERR_STATE, "InitializeVariables first");
    for (int i = 0; i < dims_.x; ++i)
      for (int j = 0; j < dims_.y; ++j)
        for (int k = 0; k < dims_.z; ++k)
          if (inside(i + 0.5f, j + 0.5f, k + 0.5f))
            mask_[(size_t(i) * dims_.y + j) * dims_.z + k] = 1;
    const float v[3] = {velocity.x, velocity.y, velocity.z};
    check(fsl_seed_particles(ctx_, mask_.data(), ppc_, seed_, v));
  }

  fsl_ctx *ctx_ = nullptr;
  Coord dims_, partitions_{1, 1, 1};
  int rank_ = 0, ppc_ = 8;
  float gravity_ = 9.8f, flip_ = 1.0f;
  Vec3 force_{0, 1, 0};
  uint32_t seed_ = 1;
  uint32_t source_stream_ = 0;  // Philox key word of the next source call
  std::vector<Source> sources_;
  std::vector<uint8_t> mask_;
  std::string output_dir_ = "output";
  bool debug_ = false;
  int frame_ = 0, debug_step_ = 0;
};

/* PCGDriver (solver_driver.h:30-106): Jacobi PCG on the FLUID cells of the
 * simulation's marker grid, b = divergence grid, x = pressure grid. As in the
 * reference, the constructor's preconditioner type is stored but the solve
 * runs Jacobi (solver_driver.cc:250 default-constructs the solver);
 * set_preconditioner(IC) selects the IC branch of BuildPreconditioner
 * (solver.cc:103-111): block IC(0) per tile. */
class PCGDriver {
 public:
  PCGDriver(std::string name, int max_iterations, float threshold, bool fixed_iterations,
            PreconditionerType preconditioner_type, Coord partitions, Coord global_dims)
      : name_(std::move(name)), max_iterations_(max_iterations), threshold_(threshold),
        fixed_iterations_(fixed_iterations), preconditioner_type_(preconditioner_type),
        partitions_(partitions), global_dims_(global_dims) {
    params_.precision = FSL_PCG_FP32;
    params_.stop_mode = FSL_STOP_ABS_RR;
  }

  /* SetVariables(b_grid, marker, x_grid) (solver_driver.h:89-91): the three
   * grids belong to the simulation's context. */
  void SetVariables(FlipSimulation *simulation) { sim_ = simulation; }

  /* DeclareVariables, AssignWorkerForGlobals, AssignWorkersForPartitions and
   * SetMigratableVariables (solver_driver.h:96-102) declare the solver's
   * Canary variables and place them on workers. The simulation's context
   * holds every solver vector on its device, so there is nothing to declare
   * or place; the calls are kept so reference call sites compile. */
  void DeclareVariables() {}
  void AssignWorkerForGlobals(int idx) { (void)idx; }
  void AssignWorkersForPartitions(const std::vector<int> &idxs) { (void)idxs; }
  void SetMigratableVariables() {}

  /* Solve (solver_driver.cc:13-69): the reference's stopping rule is
   * r.r < threshold (FP32); fixed_iterations runs exactly max_iterations. */
  void Solve() {
    if (!sim_) throw Error(FSL_ERR_STATE, name_ + ": SetVariables first");
    const fsl_solve_params p = step_params();
    sim_->check(fsl_pcg_solve(sim_->context(), &p, &stats_));
  }
  /* The parameters Solve runs with (the stopping rule and iteration cap of
   * the constructor, or of set_solve_params). */
  fsl_solve_params step_params() const {
    fsl_solve_params p = params_;
    p.threshold = fixed_iterations_ ? 0.0 : double(threshold_);
    p.max_iterations = max_iterations_;
    return p;
  }
  /* The statistics of the last solve a device-resident frame ran. */
  void set_stats(const fsl_solve_stats &s) { stats_ = s; }

  /* The preconditioner the solve builds: IC gives block IC(0), every other
   * type Jacobi (BuildPreconditioner's branches, solver.cc:103-111). */
  void set_preconditioner(PreconditionerType t) {
    params_.preconditioner = t == IC ? FSL_PRECOND_IC : FSL_PRECOND_JACOBI;
  }
  PreconditionerType preconditioner_type() const { return preconditioner_type_; }

  /* Precision / stopping rule beyond the reference's (FSL_PCG_*, FSL_STOP_*). */
  void set_solve_params(const fsl_solve_params &p) {
    params_ = p;
    threshold_ = float(p.threshold);
  }
  const fsl_solve_stats &stats() const { return stats_; }
  int num_iterations() const { return stats_.iterations; }
  const std::string &name() const { return name_; }

 private:
  std::string name_;
  int max_iterations_;
  float threshold_;
  bool fixed_iterations_;
  PreconditionerType preconditioner_type_;
  Coord partitions_, global_dims_;
  FlipSimulation *sim_ = nullptr;
  fsl_solve_params params_{};
  fsl_solve_stats stats_{};
};

/* PCGSolver (solver.h:143-205): the solver's stages, one call each, on the
 * simulation's grids (b = divergence, x = pressure, the FLUID rows of the
 * marker). The reference driver's While loop (solver_driver.cc:318-343) is
 *   PreparePCGSolver(); while (!(it >= max || rr() < threshold)) {
 *     ComputeLocal1(); ComputeLocal2(); ComputeLocal3(); ++it; }
 *   CopySolution();
 * The scalars are the global (reduced) values. The stages run the two-pass
 * iteration; params.dot_order = FSL_DOT_ROW_ORDER sums the dot products in
 * the row order of CreateIndexAndMaskTree (solver.cc:114-151). */
class PCGSolver {
 public:
  PCGSolver(FlipSimulation *simulation, const fsl_solve_params &params)
      : sim_(simulation), params_(params) {
    params_.variant = FSL_PCG_TWO_PASS;
  }
  void PreparePCGSolver() {
    double rr = 0, rz = 0;
    sim_->check(fsl_pcg_prepare(sim_->context(), &params_, &rr, &rz));
    rr_ = rr;
    rz_ = rz;
  }
  void ComputeLocal1() { sim_->check(fsl_pcg_local1(sim_->context(), &pAp_)); }
  void ComputeLocal2() {
    rz_old_ = rz_;
    sim_->check(fsl_pcg_local2(sim_->context(), &alpha_, &rr_, &rz_));
  }
  void ComputeLocal3() { sim_->check(fsl_pcg_local3(sim_->context(), &beta_)); }
  /* CopySolution (solver.cc:93-99) into the pressure grid. */
  void CopySolution() { sim_->check(fsl_pcg_finish(sim_->context(), &stats_)); }

  double rz() const { return rz_; }
  double rz_old() const { return rz_old_; }
  double rr() const { return rr_; }
  double pAp() const { return pAp_; }
  double alpha() const { return alpha_; }
  double beta() const { return beta_; }
  const fsl_solve_stats &stats() const { return stats_; }

 private:
  FlipSimulation *sim_;
  fsl_solve_params params_;
  double rz_ = 0, rz_old_ = 0, rr_ = 0, pAp_ = 0, alpha_ = 0, beta_ = 0;
  fsl_solve_stats stats_{};
};

/* FlipDriver's time loop (flip_app_sim.cc:211-254, 357-462) as plain calls. */
class FlipDriver {
 public:
  struct Params {
    int frames = 1;
    int frame_rate = 30;
    int cfl = 1;
    int max_iterations = 1500;  // flip_app.h:52
  };

  FlipDriver(FlipSimulation *simulation, PCGDriver *solver, Params params)
      : sim_(simulation), solver_(solver), params_(params) {
    frame_step_ = 1.0f / float(params_.frame_rate);
    solver_->SetVariables(sim_);
  }

  /* WhileAndComputeDt's dt for the current step (flip_app_sim.cc:371-410). */
  float ComputeDt() {
    float dt = 0;
    sim_->check(fsl_compute_dt(sim_->context(), params_.cfl, frame_step_, &step_, &dt));
    return dt;
  }

  /* RunOneStep (flip_app_sim.cc:211-254); the level-set stages (phi) are
   * not part of this core. Sources use the time at the start of the step
   * (advanced time, flip_app_sim.cc:438-462). */
  void RunOneStep(float dt) {
    if (sim_->HasSources()) sim_->AddContributionFromSources(advanced_time_);
    sim_->MoveParticles(dt);
    sim_->DeleteParticlesInSolids();
    if (sim_->HasSources()) sim_->DeleteFluidInSources(advanced_time_);
    sim_->ClearGridData();
    sim_->TransferParticlesToGrid();
    sim_->ApplyBoundaryConditions();
    sim_->SaveGridVelocities();
    sim_->DeleteOutsideParticles();
    sim_->AddGravity(dt);
    sim_->ApplyBoundaryConditions();
    sim_->ComputeDivergence();
    solver_->Solve();
    sim_->UpdateVelocityFromPressure();
    sim_->GetVelocityUpdate();
    sim_->UpdateParticlesFromGrid();
    advanced_time_ += dt;
    ++steps_;
  }

  /* One frame: steps until the frame's time is used up (the While loop of
   * flip_app_sim.cc:357-369), then ResetDt and AdvanceTime. */
  void NextFrame() {
    while (double(step_) + 1e-5 < double(frame_step_)) RunOneStep(ComputeDt());
    step_ = 0;
    time_ += frame_step_;
    ++frame_;
  }
  /* NextFrame with the frame on the device (fsl_advance_frame): dt from the
   * CFL maximum there, the counts kept there, and with graph set the front
   * half of each step replayed from a HIP graph and each solve enqueued at
   * the previous solve's iteration count. Same results as NextFrame. Water
   * sources are evaluated on the host at each step's start, so a scene with
   * sources runs NextFrame. */
  void NextFrameOnDevice(bool graph = true) {
    if (sim_->HasSources()) {
      NextFrame();
      return;
    }
    sim_->check(fsl_set_step_graph(sim_->context(), graph ? 1 : 0));
    const fsl_solve_params p = solver_->step_params();
    fsl_frame_stats fs{};
    sim_->check(fsl_advance_frame(sim_->context(), params_.cfl, frame_step_, &p, &fs));
    solver_->set_stats(fs.last_solve);
    steps_ += fs.steps;
    // The frame's steps end exactly at its end (the last dt is cut there).
    advanced_time_ += frame_step_;
    step_ = 0;
    time_ += frame_step_;
    ++frame_;
  }
  void Run() {
    for (int f = 0; f < params_.frames; ++f) NextFrame();
  }

  int steps() const { return steps_; }
  int frame() const { return frame_; }
  float time() const { return time_; }

 private:
  FlipSimulation *sim_;
  PCGDriver *solver_;
  Params params_;
  float frame_step_ = 1.0f / 30;
  float step_ = 0;
  float time_ = 0;
  float advanced_time_ = 0;  // sum of the steps' dt (AdvanceTime)
  int steps_ = 0, frame_ = 0;
};

}  // namespace fsl

#endif  // FSL_HPP_
