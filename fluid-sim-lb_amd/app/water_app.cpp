// water_app - runs FLIP frames of a scene through fsl::FlipDriver, in one
// process on one GPU (the reference's Canary water application,
// projects/flip-current-lb/water_app.cc, with its transforms run in order).
//
// Usage: fsl_water_app key=value ...
//   scene=sphere_drop|dam_break   n=64 (cells per axis)   frames=1
//   partition_x=1 partition_y=1 partition_z=1   seed=1   ppc=8
//   gravity=9.8 flip=0.95 frame_rate=30 cfl=1
//   solver_max_iterations=1500 solver_threshold=5e-9 (times n^3)
//   output=<prefix>   writes <prefix>NNNNNNNN binary particle frames
//   grid_output=<prefix>  also writes text grid dumps <prefix>{velocity,
//                     pressure,marker}_NNNNNNNN each frame
// Prints one JSON object per frame.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>

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
    const float gravity = std::strtof(get(kv, "gravity", "9.8").c_str(), nullptr);
    const float flip = std::strtof(get(kv, "flip", "0.95").c_str(), nullptr);
    const int ppc = std::atoi(get(kv, "ppc", "8").c_str());

    fsl::FlipSimulation sim;
    sim.InitializeMetadata(dims, parts, 0, 1.0f, ppc, gravity, flip);
    sim.InitializeVariables();
    sim.set_seed(uint32_t(std::strtoul(get(kv, "seed", "1").c_str(), nullptr, 10)));
    sim.SetBoxBoundary();
    if (scene == "sphere_drop") {
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
                          threshold, false, fsl::JACOBI, parts, dims);
    fsl::FlipDriver::Params prm;
    prm.frames = std::atoi(get(kv, "frames", "1").c_str());
    prm.frame_rate = std::atoi(get(kv, "frame_rate", "30").c_str());
    prm.cfl = std::atoi(get(kv, "cfl", "1").c_str());
    fsl::FlipDriver driver(&sim, &solver, prm);
    const std::string out = get(kv, "output", "");

    for (int f = 0; f < prm.frames; ++f) {
      const int s0 = driver.steps();
      const auto t0 = std::chrono::steady_clock::now();
      driver.NextFrame();
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
