// particle_stabilizer_main -- the reference's command line (src/App/ParticleStabilizerMain.cpp:40-301)
// with the HIP backend as its simulator. Flags, defaults and the CSV format are the reference's;
// the backend is resolved the way the reference resolves its CUDA plugin: dlopen of the shared
// library, dlsym of the runSimulate* symbol, call through the same function-pointer type.
//
// Differences, all deliberate: --hip selects this backend and is the default, running the
// reference CPU engine's step (its CPU engine is not part of this build); --cuda runs this backend
// with the reference CUDA engine's step (PSX_FLAG_SEMANTICS_CUDA, passed to the plugin as
// PSX_SEMANTICS=cuda, as --dtype is passed as PSX_DTYPE); --help works
// like -h; --coeff-of-restitution-wall is read in its documented order Z_MIN Z_MAX X_MIN X_MAX
// Y_MIN Y_MAX (the reference's parenthesised constructor call leaves the order unspecified and
// g++ reverses it); a flag missing its value is an error instead of reading past argv;
// --dtype f32|f64 picks the device arithmetic (default f64, the reference's scalar type); an
// --resolve-overlap-axis other than 0, 1, 2 (negative: off) is an error instead of an out-of-range
// vector index.
#include <dlfcn.h>
#include <limits.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "psx.h"
#include "psx_compat.hpp"

using psxcompat::CoefficientOfRestitutionWall;
using psxcompat::dvec3;
using psxcompat::Particle;
using psxcompat::ParticlePtr;

// the reference's runSimulateCUDA_t (ParticleStabilizerMain.cpp:24-35)
using runSimulate_t = void (*)(const std::vector<ParticlePtr>&, const dvec3, const dvec3,
                               const std::string, const dvec3, const CoefficientOfRestitutionWall,
                               const double, const double, const double, const size_t, const bool,
                               std::vector<Particle>&);

namespace {

struct Argument {  // defaults: ParticleStabilizerMain.cpp:41-57
  std::string particleFilePath, objOutPath, particleOutPath;
  dvec3 initialVelocity{0.0, 0.0, -10.0};
  double density = 1.0;
  CoefficientOfRestitutionWall wall{0.5, 0.5, 0.5, 0.5, 0.5, 0.5};
  double coefficientOfSpring = 1.0;
  double coefficientOfRestitutionSphere = 0.1;
  dvec3 gravity{0.0, 0.0, -10.0};
  dvec3 boxMin{0.0, 0.0, 0.0};
  dvec3 boxMax{10.0, 20.0, 30.0};
  std::string containerPath;
  double dt = 0.00001;
  bool toVisualize = false;
  long long maxSteps = 100000;
  int resolveOverlapAxis = -1;
  bool toUseCUDA = false;
  std::string dtype = "f64";
};

void usage() {
  std::cout
      << "######################## Particle Stabilizer (HIP backend) ########################\n"
         "\n"
         "A simulator for particle dynamics in a container.\n"
         "\n"
         "usage: ./particle_stabilizer_main [Options] particleFilePath\n"
         "\n"
         "[Options]\n"
         "  General\n"
         "    -h, --help                                                          Show this help message\n"
         "    --hip                                                               HIP backend, CPU engine's step (default)\n"
         "    --cuda                                                              HIP backend, CUDA engine's step\n"
         "    --dtype                        f32|f64                              Device arithmetic (default f64)\n"
         "\n"
         "  Particle Constants\n"
         "    --initial-velocity             X Y Z                                Initial velocity of the particles\n"
         "    --density                      MASS                                 Density of the particles\n"
         "\n"
         "  Simulation Constants\n"
         "    --coeff-of-restitution-wall    Z_MIN Z_MAX X_MIN X_MAX Y_MIN Y_MAX  Coefficient of restitution for the wall\n"
         "    --coeff-of-spring              SPRING                               Coefficient of spring\n"
         "    --coeff-of-restitution-sphere  RESTITUTION                          Coefficient of restitution for the sphere\n"
         "    --gravity                      X Y Z                                Gravity\n"
         "    --dt                           DT                                   Time step\n"
         "    --max-steps                    MAX_STEPS                            Maximum number of steps\n"
         "\n"
         "  Box Container\n"
         "    --box-min                      X Y Z                                Minimum coordinates of the box container\n"
         "    --box-max                      X Y Z                                Maximum coordinates of the box container\n"
         "    --container                    CONTAINER_PATH                       Path to the container file (.obj).\n"
         "                                                                        If this option is specified, 0th value of\n"
         "                                                                        '--coeff-of-restitution-wall' will be used for\n"
         "                                                                        the polygon restitution coefficient.\n"
         "\n"
         "  Postprocess\n"
         "    --resolve-overlap-axis         AXIS                                 Resolve overlaps along the specified axis\n"
         "\n"
         "  Save\n"
         "    --out-obj-path                 OBJ_PATH                             Path to save the simulation as .obj file\n"
         "    --out-csv-path                 CSV_PATH                             Path to save the simulation as .csv file\n";
}

[[noreturn]] void die(const std::string& msg) {
  std::cerr << "[psx] CRITICAL: " << msg << "\n";
  std::exit(EXIT_FAILURE);
}

Argument parse_args(int argc, char** argv) {
  Argument a;
  bool help = argc < 2;
  auto need = [&](int& i, int k) -> char** {
    if (i + k >= argc) die(std::string("option ") + argv[i] + " needs " + std::to_string(k) + " value(s)");
    char** v = argv + i + 1;
    i += k;
    return v;
  };
  auto num = [](const char* s) {
    try {
      return std::stod(s);
    } catch (const std::exception&) {
      die(std::string("not a number: '") + s + "'");
    }
  };
  for (int i = 1; i < argc; ++i) {
    const std::string arg = argv[i];
    if (arg == "-h" || arg == "--help") {
      help = true;
      break;
    } else if (arg == "--out-obj-path") {
      a.objOutPath = need(i, 1)[0];
    } else if (arg == "--out-csv-path") {
      a.particleOutPath = need(i, 1)[0];
    } else if (arg == "--initial-velocity") {
      char** v = need(i, 3);
      a.initialVelocity = {num(v[0]), num(v[1]), num(v[2])};
    } else if (arg == "--density") {
      a.density = num(need(i, 1)[0]);
    } else if (arg == "--coeff-of-restitution-wall") {
      char** v = need(i, 6);
      a.wall = {num(v[0]), num(v[1]), num(v[2]), num(v[3]), num(v[4]), num(v[5])};
    } else if (arg == "--coeff-of-spring") {
      a.coefficientOfSpring = num(need(i, 1)[0]);
    } else if (arg == "--coeff-of-restitution-sphere") {
      a.coefficientOfRestitutionSphere = num(need(i, 1)[0]);
    } else if (arg == "--gravity") {
      char** v = need(i, 3);
      a.gravity = {num(v[0]), num(v[1]), num(v[2])};
    } else if (arg == "--box-min") {
      char** v = need(i, 3);
      a.boxMin = {num(v[0]), num(v[1]), num(v[2])};
    } else if (arg == "--box-max") {
      char** v = need(i, 3);
      a.boxMax = {num(v[0]), num(v[1]), num(v[2])};
    } else if (arg == "--container") {
      a.containerPath = need(i, 1)[0];
    } else if (arg == "--dt") {
      a.dt = num(need(i, 1)[0]);
    } else if (arg == "--visualize") {
      a.toVisualize = true;
    } else if (arg == "--max-steps") {
      a.maxSteps = std::atoll(need(i, 1)[0]);
    } else if (arg == "--resolve-overlap-axis") {
      a.resolveOverlapAxis = std::atoi(need(i, 1)[0]);
    } else if (arg == "--cuda") {
      a.toUseCUDA = true;
    } else if (arg == "--hip") {
      a.toUseCUDA = false;
    } else if (arg == "--dtype") {
      a.dtype = need(i, 1)[0];
      if (a.dtype != "f32" && a.dtype != "f64") die("--dtype must be f32 or f64");
    } else {
      a.particleFilePath = arg;  // ParticleStabilizerMain.cpp:150-152
    }
  }
  if (a.particleFilePath.empty()) help = true;
  if (help) {
    usage();
    std::exit(EXIT_SUCCESS);
  }
  return a;
}

std::string library_path() {
  if (const char* e = std::getenv("PSX_LIBRARY")) return e;
  char buf[PATH_MAX];
  const ssize_t k = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
  if (k > 0) {
    buf[k] = 0;
    std::string dir(buf);
    dir = dir.substr(0, dir.find_last_of('/'));
    return dir + "/../lib/libparticle_stabilizer_hip.so";
  }
  return "./libparticle_stabilizer_hip.so";
}

}  // namespace

int main(int argc, char** argv) {
  const Argument args = parse_args(argc, argv);
  std::cout << "[psx] particleFilePath = " << args.particleFilePath << "\n"
            << "[psx] coefficientOfRestitutionWall = zMin: " << args.wall.zMin
            << ", zMax: " << args.wall.zMax << ", xMin: " << args.wall.xMin
            << ", xMax: " << args.wall.xMax << ", yMin: " << args.wall.yMin
            << ", yMax: " << args.wall.yMax << "\n"
            << "[psx] spring = " << args.coefficientOfSpring
            << ", restitutionSphere = " << args.coefficientOfRestitutionSphere
            << ", dt = " << args.dt << ", maxSteps = " << args.maxSteps << ", dtype = " << args.dtype
            << "\n";
  if (args.maxSteps < 0) die("--max-steps must be >= 0");

  // backend: dlopen + dlsym, as the reference loads its CUDA plugin (DynamicLoader.cpp:5-46)
  const std::string lib = library_path();
  void* handle = dlopen(lib.c_str(), RTLD_LAZY);
  if (!handle) die(std::string("Cannot open library: ") + dlerror());
  auto run = reinterpret_cast<runSimulate_t>(dlsym(handle, "runSimulateHIP"));
  if (!run) die("Cannot load symbol: 'runSimulateHIP' from the shared library.");
  using read_csv_t = int (*)(const char*, double, const double*, int64_t*, double**, double**,
                             double**, double**, int64_t*);
  using write_csv_t = int (*)(const char*, int64_t, const double*, const double*);
  using free_t = void (*)(void*);
  using err_t = const char* (*)();
  auto read_csv = reinterpret_cast<read_csv_t>(dlsym(handle, "psx_read_csv"));
  auto write_csv = reinterpret_cast<write_csv_t>(dlsym(handle, "psx_write_csv"));
  auto pfree = reinterpret_cast<free_t>(dlsym(handle, "psx_free"));
  auto last_error = reinterpret_cast<err_t>(dlsym(handle, "psx_last_error"));
  if (!read_csv || !write_csv || !pfree || !last_error) die("library lacks the psx I/O entry points");

  const double v0[3] = {args.initialVelocity.x, args.initialVelocity.y, args.initialVelocity.z};
  int64_t n = 0, skipped = 0;
  double *pos = nullptr, *vel = nullptr, *rad = nullptr, *mas = nullptr;
  if (read_csv(args.particleFilePath.c_str(), args.density, v0, &n, &pos, &vel, &rad, &mas, &skipped) != 0)
    die(last_error());
  std::cout << "[psx] Read " << n << " particles from file: " << args.particleFilePath << "\n";
  if (n == 0) die("no particles");

  std::vector<ParticlePtr> particles((size_t)n);
  for (int64_t i = 0; i < n; ++i)
    particles[(size_t)i] = std::make_shared<Particle>(
        Particle{{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]}, rad[i], mas[i]});
  pfree(pos);
  pfree(vel);
  pfree(rad);
  pfree(mas);

  std::vector<Particle> result((size_t)n);
  setenv("PSX_DTYPE", args.dtype.c_str(), 1);
  setenv("PSX_SEMANTICS", args.toUseCUDA ? "cuda" : "cpu", 1);
  std::cout << "[psx] Using HIP for the simulation.\n";
  run(particles, args.boxMin, args.boxMax, args.containerPath, args.gravity, args.wall,
      args.coefficientOfSpring, args.coefficientOfRestitutionSphere, args.dt,
      (size_t)args.maxSteps, args.toVisualize, result);

  std::vector<double> p(3 * (size_t)n), r((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    p[3 * i] = result[(size_t)i].position.x;
    p[3 * i + 1] = result[(size_t)i].position.y;
    p[3 * i + 2] = result[(size_t)i].position.z;
    r[(size_t)i] = result[(size_t)i].radius;
  }

  // ParticleStabilizerMain.cpp:275-296: overlap post-process, then the CSV, then the OBJ
  if (args.resolveOverlapAxis >= 0) {
    using resolve_t = int (*)(int64_t, double*, const double*, int32_t, int64_t*, double*);
    auto resolve = reinterpret_cast<resolve_t>(dlsym(handle, "psx_resolve_overlaps"));
    if (!resolve) die("library lacks psx_resolve_overlaps");
    std::cout << "[psx] ### Resolving overlaps\n";
    int64_t collided = 0;
    double worst = 0.0;
    if (resolve(n, p.data(), r.data(), args.resolveOverlapAxis, &collided, &worst) != 0)
      die(last_error());
    char line[128];
    std::snprintf(line, sizeof line, "%.15f", worst);
    std::cout << "[psx] Number of collided particles: " << collided << " / " << n << "\n"
              << "[psx] Max collision distance: " << line << "\n"
              << "[psx] ### Finished resolving overlaps\n";
  }

  if (!args.particleOutPath.empty()) {
    if (write_csv(args.particleOutPath.c_str(), n, p.data(), r.data()) != 0) die(last_error());
    std::cout << "[psx] Finished writing to file " << args.particleOutPath << "\n";
  }
  if (!args.objOutPath.empty()) {
    using write_obj_t = int (*)(const char*, int64_t, const double*, const double*);
    auto write_obj = reinterpret_cast<write_obj_t>(dlsym(handle, "psx_write_obj_spheres"));
    if (!write_obj) die("library lacks psx_write_obj_spheres");
    std::cout << "[psx] Writing to file " << args.objOutPath << "\n";
    // ObjData::write logs a file that cannot be opened and carries on
    if (write_obj(args.objOutPath.c_str(), n, p.data(), r.data()) != 0)
      std::cerr << "[psx] " << last_error() << "\n";
    else
      std::cout << "[psx] Finished writing to file " << args.objOutPath << "\n";
  }
  std::cout << "[psx] Simulation completed successfully.\n";
  return EXIT_SUCCESS;
}
