// psx_compat.cpp -- runSimulateHIP, the reference's plugin signature (psx_compat.hpp) implemented on
// the C ABI. Mirrors runSimulateCUDA (PhysicsEngineCUDA.cpp:86-175): build the model and the
// container, run maxSteps steps with the periodic statistics log, copy the particles out by index.
#include "psx_compat.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "psx.h"

namespace {

void log_info(const std::string& s) { std::fprintf(stdout, "[psx] %s\n", s.c_str()); }

[[noreturn]] void critical(const std::string& what) {
  std::fprintf(stderr, "[psx] CRITICAL: %s: %s\n", what.c_str(), psx_last_error());
  std::exit(EXIT_FAILURE);
}

int dtype_from_env() {
  const char* e = std::getenv("PSX_DTYPE");
  if (e && (std::strcmp(e, "f32") == 0 || std::strcmp(e, "float32") == 0)) return PSX_F32;
  return PSX_F64;  // the reference's scalar type
}

// PSX_SEMANTICS=cuda: the reference CUDA engine's step (PSX_FLAG_SEMANTICS_CUDA); anything else,
// or unset, its CPU engine's.
int flags_from_env() {
  const char* e = std::getenv("PSX_SEMANTICS");
  return (e && std::strcmp(e, "cuda") == 0) ? PSX_FLAG_SEMANTICS_CUDA : 0;
}

}  // namespace

extern "C" void runSimulateHIP(const std::vector<psxcompat::ParticlePtr>& particles,
                               const psxcompat::dvec3 boxMin, const psxcompat::dvec3 boxMax,
                               const std::string containerPath, const psxcompat::dvec3 gravity,
                               const psxcompat::CoefficientOfRestitutionWall walls,
                               const double coefficientOfSpring,
                               const double coefficientOfRestitutionSphere, const double dt,
                               const size_t maxSteps, const bool toVisualize,
                               std::vector<psxcompat::Particle>& resultParticles) {
  const auto t_start = std::chrono::high_resolution_clock::now();
  const size_t n = particles.size();
  if (toVisualize) log_info("visualization is not available in the HIP backend; running headless");

  std::vector<double> pos(3 * n), vel(3 * n), rad(n), mas(n);
  for (size_t i = 0; i < n; ++i) {
    const psxcompat::Particle& p = *particles[i];
    pos[3 * i] = p.position.x;
    pos[3 * i + 1] = p.position.y;
    pos[3 * i + 2] = p.position.z;
    vel[3 * i] = p.velocity.x;
    vel[3 * i + 1] = p.velocity.y;
    vel[3 * i + 2] = p.velocity.z;
    rad[i] = p.radius;
    mas[i] = p.mass;
  }

  psx_params prm;
  std::memset(&prm, 0, sizeof(prm));
  prm.gravity[0] = gravity.x;
  prm.gravity[1] = gravity.y;
  prm.gravity[2] = gravity.z;
  const double w[6] = {walls.zMin, walls.zMax, walls.xMin, walls.xMax, walls.yMin, walls.yMax};
  std::memcpy(prm.walls, w, sizeof(w));
  prm.spring = coefficientOfSpring;
  prm.restitution = coefficientOfRestitutionSphere;
  prm.dt = dt;
  prm.box_min[0] = boxMin.x;
  prm.box_min[1] = boxMin.y;
  prm.box_min[2] = boxMin.z;
  prm.box_max[0] = boxMax.x;
  prm.box_max[1] = boxMax.y;
  prm.box_max[2] = boxMax.z;
  prm.dtype = dtype_from_env();
  prm.flags = flags_from_env();

  psx_mesh mesh;
  std::memset(&mesh, 0, sizeof(mesh));
  double* tri = nullptr;
  if (!containerPath.empty()) {
    if (psx_read_obj(containerPath.c_str(), &mesh.n_triangles, &tri) != PSX_OK)
      critical("cannot read container " + containerPath);
    mesh.vertices = tri;
    log_info("nPolygons = " + std::to_string(mesh.n_triangles));
  }

  psx_sim* sim = nullptr;
  if (psx_create(&prm, (int64_t)n, pos.data(), vel.data(), rad.data(), mas.data(),
                 containerPath.empty() ? nullptr : &mesh, &sim) != PSX_OK)
    critical("psx_create");
  psx_free(tri);
  {
    // steps replayed from hipGraphs and the incremental bucketing: the same results as every other
    // launch shape, faster on every workload measured (the shapes themselves keep their defaults)
    psx_tuning t;
    if (psx_get_tuning(sim, &t) != PSX_OK) critical("psx_get_tuning");
    t.use_graph = 1;
    t.sort_variant = 3;
    if (psx_set_tuning(sim, &t) != PSX_OK) critical("psx_set_tuning");
  }

  {
    double mn[3], h;
    int64_t n3[3], base;
    psx_get_grid(sim, 0, mn, &h, n3, &base);
    log_info("interval  = " + std::to_string(h));
    log_info("nBuckets  = (" + std::to_string(n3[0]) + ", " + std::to_string(n3[1]) + ", " +
             std::to_string(n3[2]) + ")");
  }

  // the reference logs progress when more than 10 s have passed (PhysicsEngine.cpp:50-67)
  auto t_prev = std::chrono::high_resolution_clock::now();
  size_t done = 0;
  const size_t chunk = 1000;
  while (done < maxSteps) {
    const size_t k = std::min(chunk, maxSteps - done);
    if (psx_step(sim, (int64_t)k) != PSX_OK) critical("psx_step");
    done += k;
    const auto now = std::chrono::high_resolution_clock::now();
    if (std::chrono::duration<double>(now - t_prev).count() > 10.0 || done == maxSteps) {
      // ParticleStatistics::calculateStatistics (ParticleStatistics.cpp:3-43), computed on the
      // GPU over the particle grid, logged as PhysicsEngine.cpp:59-64 does
      psx_stats st;
      if (psx_statistics(sim, &st) != PSX_OK) critical("psx_statistics");
      auto num = [](double v) { return std::to_string(v); };
      log_info("######################## Step " + std::to_string(done) + ", Elapsed time: " +
               num(std::chrono::duration<double>(now - t_start).count()) +
               " [sec] ########################");
      log_info("Total kinetic energy         : " + num(st.kinetic_energy));
      log_info("Max position                 : (" + num(st.max[0]) + ", " + num(st.max[1]) + ", " +
               num(st.max[2]) + ")");
      log_info("Min position                 : (" + num(st.min[0]) + ", " + num(st.min[1]) + ", " +
               num(st.min[2]) + ")");
      log_info("Number of collided particles : " + std::to_string(st.n_collided) + " / " +
               std::to_string(n));
      log_info("Max collision distance       : " + num(st.max_overlap));
      t_prev = now;
    }
  }
  if (psx_synchronize(sim) != PSX_OK) critical("psx_synchronize");
  if (psx_get_state(sim, pos.data(), vel.data()) != PSX_OK) critical("psx_get_state");
  psx_destroy(sim);

  if (resultParticles.size() < n) resultParticles.resize(n);
  for (size_t i = 0; i < n; ++i)
    resultParticles[i] = psxcompat::Particle{{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]},
                                             {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]},
                                             rad[i],
                                             mas[i]};
  const double total =
      std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t_start).count();
  log_info("Total elapsed time: " + std::to_string(total) + " [sec]");
}
