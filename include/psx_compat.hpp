// psx_compat.hpp -- the reference's plugin entry point, kept signature-for-signature.
//
// runSimulateHIP replaces runSimulateCUDA (include/ParticleStabilizerCUDA/PhysicsEngineCUDA.hpp:43-54),
// which the reference's CLI resolves with dlopen("./libparticle_stabilizer_cuda.so") + dlsym
// (src/App/ParticleStabilizerMain.cpp:24-35, 235-256). The argument types below are layout- and
// calling-convention-compatible stand-ins for glm::dvec3, Particle / ParticlePtr
// (include/ParticleStabilizer/ParticleModel.hpp:7-28) and CoefficientOfRestitutionWall
// (include/ParticleStabilizer/ContainerBase.hpp:10-34), so a binary compiled against the reference
// headers with the same C++ standard library can call it through the same function-pointer type.
//
// Conventions kept from the reference: synchronous, device 0, errors are logged and end the process
// with exit(EXIT_FAILURE) (CudaCommon.cuh:18-28). New code should call the C ABI in psx.h instead,
// which never exits. The signature has no room for options, so two come from the environment:
// PSX_DTYPE=f32 (device arithmetic; binary64 otherwise) and PSX_SEMANTICS=cuda (the reference
// CUDA engine's step, PSX_FLAG_SEMANTICS_CUDA; its CPU engine's step otherwise).
#pragma once

#include <cstddef>
#include <memory>
#include <string>
#include <vector>

namespace psxcompat {

struct dvec3 {  // glm::dvec3: three doubles, trivially copyable, 24 bytes
  double x, y, z;
};

struct Particle {  // ParticlePrimitive<glm::dvec3>
  dvec3 position;
  dvec3 velocity;
  double radius;
  double mass;
};
using ParticlePtr = std::shared_ptr<Particle>;

struct CoefficientOfRestitutionWall {  // field order of ContainerBase.hpp:28-33
  double zMin, zMax, xMin, xMax, yMin, yMax;
};

static_assert(sizeof(dvec3) == 24, "glm::dvec3 layout");
static_assert(sizeof(Particle) == 64, "Particle layout");
static_assert(sizeof(CoefficientOfRestitutionWall) == 48, "CoefficientOfRestitutionWall layout");

}  // namespace psxcompat

extern "C" void runSimulateHIP(const std::vector<psxcompat::ParticlePtr>& particles,
                               const psxcompat::dvec3 boxMin, const psxcompat::dvec3 boxMax,
                               const std::string containerPath, const psxcompat::dvec3 gravity,
                               const psxcompat::CoefficientOfRestitutionWall coefficientOfRestitutionWall,
                               const double coefficientOfSpring,
                               const double coefficientOfRestitutionSphere, const double dt,
                               const size_t maxSteps, const bool toVisualize,
                               std::vector<psxcompat::Particle>& resultParticles);
