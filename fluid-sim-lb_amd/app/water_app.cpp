// water_app - runs FLIP frames of a scene through fsl::FlipDriver, in one
// process on one GPU (the reference's Canary water application,
// projects/flip-current-lb/water_app.cc, with its transforms run in order).
//
// Usage: fsl_water_app key=value ...
//   scene=sphere_drop|dam_break   n=64 (cells per axis)   frames=1
//   partition_x=1 partition_y=1 partition_z=1   seed=1   ppc=8
//   gravity=9.8 flip=0.95 frame_rate=30 cfl=1
//   solver_max_iterations=1500 solver_threshold=5e-9 (times n^3)
//   preconditioner=jacobi|ic   (ic: block IC(0) per partition)
//   device_loop=1     each frame on the device (FlipDriver::NextFrameOnDevice:
//                     dt there, step graph on; scenes without sources)
//   output=<prefix>   writes <prefix>NNNNNNNN binary particle frames
//   output_dir=<dir>  SaveDataToFile as the reference: <dir>/particles_NNNNNNNN,
//                     the initial state as frame 0, then every frame
//   debug=1           SaveDebugDataToFile into <output_dir>/debug at the same
//                     points (the reference also writes inside each step)
//   grid_output=<prefix>  also writes text grid dumps <prefix>{velocity,
//                     pressure,marker}_NNNNNNNN each frame
//   init=<file>       scene from an init file in the reference's grammar
//                     (flip_app.cc:159-251) instead of scene=: lines
//                     fluid N, boundary N, gravity g, force x y z,
//                     fluid-cube x0 y0 z0 x1 y1 z1 vx vy vz,
//                     fluid-sphere cx cy cz r vx vy vz,
//                     source-static x0 y0 z0 x1 y1 z1 vx vy vz,
//                     source-timed <static fields> t0 t1,
//                     source-oscillating <box1> <box2> period <v1> <v2> period,
//                     solid x0 y0 z0 x1 y1 z1; coordinates normalised to
//                     [0, 1], velocities and forces in units of the scaled
//                     gravity; separators ", :".
// Prints one JSON object per frame.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "fsl.hpp"

namespace {

std::map<std::string, std::string> parse(int argc, char **argv) {
  std::map<std::string, std::string> kv;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    const size_t eq = a.find('=');
    if (eq == std::string::npos) throw fsl::Error(FSL_ERR_ARG, "expected key=value: " + a);
    kv[a.substr(0, eq)] = a.substr(eq + 1);
  }
  return kv;
}

std::string get(const std::map<std::string, std::string> &kv, const char *k, const char *def) {
  auto it = kv.find(k);
  return it == kv.end() ? def : it->second;
}

// FlipApp::ParseSimulationConfiguration (flip_app.cc:159-251): the scene of
// an init file, still in normalised units.
struct Config {
  int fluid = 0, boundary = 0;
  float gravity = 9.8f;
  fsl::Vec3 force{0, 1, 0};
  struct Fluid {
    bool cube;
    fsl::Cube box;
    fsl::Vec3 center;
    float radius;
    fsl::Vec3 velocity;
  };
  std::vector<Fluid> fluids;
  std::vector<fsl::Source> sources;
  std::vector<fsl::Cube> solids;
};

std::vector<std::string> tokens_of(const std::string &line) {
  std::vector<std::string> t;
  std::string cur;
  for (char ch : line) {
    if (ch == ',' || ch == ' ' || ch == ':' || ch == '\t' || ch == '\r') {
      if (!cur.empty()) t.push_back(cur);
      cur.clear();
    } else {
      cur += ch;
    }
  }
  if (!cur.empty()) t.push_back(cur);
  return t;
}

Config parse_init(const std::string &file) {
  std::ifstream in(file);
  if (!in) throw fsl::Error(FSL_ERR_ARG, "cannot open init file " + file);
  Config c;
  std::string line;
  while (std::getline(in, line)) {
    const std::vector<std::string> t = tokens_of(line);
    if (t.empty()) continue;
    size_t id = 1;
    auto num = [&]() {
      if (id >= t.size()) throw fsl::Error(FSL_ERR_ARG, "init file: short line: " + line);
      return std::strtof(t[id++].c_str(), nullptr);
    };
    auto vec = [&]() {
      const float x = num(), y = num(), z = num();
      return fsl::Vec3(x, y, z);
    };
    auto cube = [&]() {
      const fsl::Vec3 a = vec(), b = vec();
      return fsl::Cube{a, b};
    };
    const std::string &key = t[0];
    if (key == "fluid") {
      c.fluid = int(num());
    } else if (key == "boundary") {
      c.boundary = int(num());
    } else if (key == "gravity") {
      c.gravity = num();
    } else if (key == "force") {
      c.force = vec();
    } else if (key == "fluid-cube") {
      Config::Fluid f{true, cube(), {}, 0, {}};
      f.velocity = vec();
      c.fluids.push_back(f);
    } else if (key == "fluid-sphere") {
      Config::Fluid f{false, {}, vec(), 0, {}};
      f.radius = num();
      f.velocity = vec();
      c.fluids.push_back(f);
    } else if (key == "source-static" || key == "source-timed") {
      const fsl::Cube b = cube();
      const fsl::Vec3 v = vec();
      c.sources.emplace_back(b, v);
      if (key == "source-timed") {
        const float t0 = num(), t1 = num();
        c.sources.back().set_timed(t0, t1);
      }
    } else if (key == "source-oscillating") {
      const fsl::Cube b1 = cube(), b2 = cube();
      const float bp = num();
      const fsl::Vec3 v1 = vec(), v2 = vec();
      const float vp = num();
      c.sources.emplace_back();
      c.sources.back().set_box(b1, b2, bp);
      c.sources.back().set_velocity(v1, v2, vp);
    } else if (key == "solid") {
      c.solids.push_back(cube());
    }
  }
  return c;
}

fsl::Cube scaled(fsl::Cube b, const fsl::Coord &d) {
  b.start = fsl::Vec3(b.start.x * d.x, b.start.y * d.y, b.start.z * d.z);
  b.end = fsl::Vec3(b.end.x * d.x, b.end.y * d.y, b.end.z * d.z);
  return b;
}

// FlipDriver::Initialize's scene set-up (flip_app_sim.cc:72-176): scaled
// gravity and forces, sources, boundary, solids (Cube::get_coord_bbox,
// primitives.cc:37-54), the default fluid, then the additional fluids.
float scaled_gravity(const Config &c, const fsl::Coord &dims) {
  return c.gravity * (float(dims.y) / 128.0f);
}

void init_from_config(const Config &c, const fsl::Coord &dims, fsl::FlipSimulation &sim) {
  const float g = scaled_gravity(c, dims);
  sim.SetForces(c.force);
  for (fsl::Source s : c.sources) {
    s.ScaleBox(dims);
    s.ScaleVelocity(g);
    sim.AddWaterSource(s);
  }
  if (c.boundary == 0) sim.SetBoxBoundary();
  else sim.SetTankBoundary();
  for (const fsl::Cube &sc : c.solids) {
    const fsl::Cube b = scaled(sc, dims);
    fsl::CoordBBox box;
    box.lo = fsl::Coord(int(b.start.x), int(b.start.y), int(b.start.z));
    box.hi = fsl::Coord(int(b.end.x - 1), int(b.end.y - 1), int(b.end.z - 1));
    if (box.hi.x < box.lo.x) box.hi.x = box.lo.x;
    if (box.hi.y < box.lo.y) box.hi.y = box.lo.y;
    if (box.hi.z < box.lo.z) box.hi.z = box.lo.z;
    sim.AddSolidCuboid(box);
  }
  if (c.fluid == 1) {
    const float h = float(dims.y / 4);
    const fsl::Vec3 ctr(float(dims.x / 2), float(dims.y * 3 / 4), float(dims.z * 5 / 8));
    sim.InitializeMultipleWaterDrops(h, {ctr}, {float(dims.x / 4) * 0.8f});
  } else if (c.fluid == 2) {
    sim.InitializeWaterCuboid(
        fsl::Cube{fsl::Vec3(0, 0, 0),
                  fsl::Vec3(float(dims.x / 4) - 1, float(dims.y * 3 / 4) - 1, float(dims.z) - 1)},
        fsl::Vec3());
  } else if (c.fluid != 0) {
    throw fsl::Error(FSL_ERR_ARG, "fluid must be 0, 1 or 2");
  }
  for (const Config::Fluid &f : c.fluids) {
    const fsl::Vec3 v(f.velocity.x * g, f.velocity.y * g, f.velocity.z * g);
    if (f.cube) {
      fsl::Cube b = scaled(f.box, dims);
      b.end = fsl::Vec3(b.end.x - 1, b.end.y - 1, b.end.z - 1);
      sim.InitializeWaterCuboid(b, v);
    } else {
      const fsl::Vec3 ctr(f.center.x * dims.x, f.center.y * dims.y, f.center.z * dims.z);
      sim.InitializeWaterSphere(ctr, f.radius * dims.x, v);
    }
  }
}

}  // namespace

int main(int argc, char **argv) {
  try {
    const auto kv = parse(argc, argv);
    const int n = std::atoi(get(kv, "n", "64").c_str());
    const std::string scene = get(kv, "scene", "sphere_drop");
    const fsl::Coord dims(n, n, n);
    const fsl::Coord parts(std::atoi(get(kv, "partition_x", "1").c_str()),
                           std::atoi(get(kv, "partition_y", "1").c_str()),
                           std::atoi(get(kv, "partition_z", "1").c_str()));
    const std::string init = get(kv, "init", "");
    Config cfg;
    if (!init.empty()) cfg = parse_init(init);
    const float gravity = init.empty() ? std::strtof(get(kv, "gravity", "9.8").c_str(), nullptr)
                                       : scaled_gravity(cfg, dims);
    const float flip = std::strtof(get(kv, "flip", "0.95").c_str(), nullptr);
    const int ppc = std::atoi(get(kv, "ppc", "8").c_str());

    fsl::FlipSimulation sim;
    const std::string out_dir = get(kv, "output_dir", "");
    const bool debug = get(kv, "debug", "0") == "1";
    sim.InitializeMetadata(dims, parts, 0, 1.0f, ppc, gravity, flip,
                           out_dir.empty() ? "output" : out_dir, debug);
    sim.InitializeVariables();
    sim.set_seed(uint32_t(std::strtoul(get(kv, "seed", "1").c_str(), nullptr, 10)));
    if (!init.empty()) {
      init_from_config(cfg, dims, sim);
    } else if (sim.SetBoxBoundary(), scene == "sphere_drop") {
      // Pool in the bottom quarter, sphere of radius n/6.4 at (n/2, 0.72n, n/2):
      // the 64^3 sphere drop is pool j < 16, r = 10 at (32, 46, 32).
      const float h = n / 4.0f, r = n * 10.0f / 64.0f;
      sim.InitializeMultipleWaterDrops(h, {fsl::Vec3(n / 2.0f, n * 46.0f / 64.0f, n / 2.0f)},
                                       {r});
    } else if (scene == "dam_break") {
      // Fluid x < 3n/8, y < 5n/8 (the 256^3 dam break is x < 96, y < 160).
      // Cell centres are tested against the inclusive cube.
      fsl::Cube box{fsl::Vec3(0, 0, 0), fsl::Vec3(3.0f * n / 8, 5.0f * n / 8, float(n))};
      sim.InitializeWaterCuboid(box, fsl::Vec3());
    } else {
      throw fsl::Error(FSL_ERR_ARG, "unknown scene " + scene);
    }

    const float threshold =
        float(double(n) * n * n) * std::strtof(get(kv, "solver_threshold", "5e-9").c_str(), nullptr);
    fsl::PCGDriver solver("pressure", std::atoi(get(kv, "solver_max_iterations", "1500").c_str()),
                          threshold, false, fsl::IC, parts, dims);
    // The reference passes IC and runs Jacobi (flip_app.cc:329-332);
    // preconditioner=ic runs block IC(0).
    if (get(kv, "preconditioner", "jacobi") == "ic") solver.set_preconditioner(fsl::IC);
    fsl::FlipDriver::Params prm;
    prm.frames = std::atoi(get(kv, "frames", "1").c_str());
    prm.frame_rate = std::atoi(get(kv, "frame_rate", "30").c_str());
    prm.cfl = std::atoi(get(kv, "cfl", "1").c_str());
    fsl::FlipDriver driver(&sim, &solver, prm);
    // device_loop=1: each frame on the device (dt there, step graph on).
    const bool device_loop = get(kv, "device_loop", "0") == "1";
    const std::string out = get(kv, "output", "");
    // FlipDriver::SaveData (flip_app_sim.cc:204, 259-274): the state after
    // initialisation is frame 0, then one frame after every frame.
    auto save = [&] {
      if (debug) sim.SaveDebugDataToFile();
      if (out_dir.empty()) return;
      sim.SaveDataToFile();
      sim.NextFrame();
    };
    save();

    for (int f = 0; f < prm.frames; ++f) {
      const int s0 = driver.steps();
      const auto t0 = std::chrono::steady_clock::now();
      if (device_loop) driver.NextFrameOnDevice(true);
      else driver.NextFrame();
      const double sec =
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (!out.empty()) {
        char num[16];
        std::snprintf(num, sizeof(num), "%08d", driver.frame());
        sim.WriteParticles(out + num, false);
      }
      const std::string gout = get(kv, "grid_output", "");
      if (!gout.empty()) {
        char num[16];
        std::snprintf(num, sizeof(num), "%08d", driver.frame());
        sim.WriteGrid(gout + "velocity_" + num, FSL_GRID_VELOCITY);
        sim.WriteGrid(gout + "pressure_" + num, 10 + FSL_GRID_PRESSURE);
        sim.WriteGrid(gout + "marker_" + num, 10 + FSL_GRID_MARKER);
      }
      save();
      std::printf("{\"frame\": %d, \"steps\": %d, \"particles\": %zu, \"last_iterations\": %d, "
                  "\"seconds\": %.6f}\n",
                  driver.frame(), driver.steps() - s0, sim.ParticleCount(),
                  solver.num_iterations(), sec);
      std::fflush(stdout);
    }
    return 0;
  } catch (const fsl::Error &e) {
    std::fprintf(stderr, "fsl_water_app: error %d: %s\n", e.code(), e.what());
    return 1;
  }
}
