// psx_host.hpp -- host-side types of the HIP backend (scene description, the simulation interface
// shared by the float and double implementations, per-kernel event timing).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "psx.h"

namespace psx {

constexpr int kNumKernels = 5;

int set_error(int code, const std::string& msg);

// Grid of ParticleBucket (ParticleBucket.cpp:17-27).
struct HostGrid {
  double min[3] = {0, 0, 0}, max[3] = {0, 0, 0};
  double h = 0;
  int64_t n[3] = {0, 0, 0};
};
HostGrid make_grid(double rmax, const double* mn, const double* mx, bool cuda = false);
struct MeshHost;
void build_tri_grid(MeshHost& m, const HostGrid& g, double rmax, bool fp32);

struct MeshHost {
  int64_t nt = 0;
  std::vector<double> v;  // 9 * nt de-indexed vertices
  double min[3], max[3];  // container bounds (PolygonContainer.cpp:288-318)
  // triangle lists per cell of the scene's particle grid
  std::vector<int32_t> off, list;
};

struct SceneHost {
  psx_params prm;
  int64_t n = 0, p_begin = 0;
  int mesh = -1;
  double rmax = 0;
  HostGrid grid;
  int64_t cell_base = 0;
};

inline psx_tuning default_tuning() {
  psx_tuning t;
  t.block_integrate = 256;
  t.block_sort = 256;
  t.block_pairs = 256;
  t.use_graph = 0;
  t.pair_variant = 0;
  t.sort_variant = 0;
  t.pair_chunk = 0;
  return t;
}

class Profiler {
 public:
  Profiler(int nk, int64_t nsteps);  // nsteps: number of steps that will be timed
  ~Profiler();
  void next_step();
  hipEvent_t begin(int k);
  hipEvent_t end(int k);
  // ms[k]: mean duration of kernel k's launches; per_step[k]: its launches per step
  void collect(double* ms, double* per_step);

 private:
  int nk_;
  int64_t step_ = 0;
  std::vector<hipEvent_t> ev_;
  std::vector<char> rec_;  // (step, kernel) pairs that were launched
};

struct SimBase {
  virtual ~SimBase() = default;
  int64_t n = 0;
  int64_t ncells = 0, cpad = 0;  // cells, and cells padded to whole scan tiles
  int64_t id_stride = 0;          // cells per id-slot plane (slot-major id slots)
  int32_t ntiles = 0;
  int64_t tri_count = 0, tcell_entries = 0;
  std::vector<SceneHost> scenes;
  std::vector<MeshHost> meshes;
  bool cuda = false;  // PSX_FLAG_SEMANTICS_CUDA: the reference CUDA backend's step
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  psx_tuning tuning = default_tuning();

  virtual int32_t dtype() const = 0;
  virtual void setup(const std::vector<double>& pos, const std::vector<double>& vel,
                     const std::vector<double>& radius, const std::vector<double>& mass) = 0;
  virtual void step(int64_t nsteps) = 0;
  // Captures the graphs that step(nsteps) would replay from the current state (no step is run).
  virtual void prepare(int64_t nsteps) = 0;
  virtual void profile(int64_t nsteps, double* ms, double* per_step) = 0;
  virtual void set_stream(hipStream_t s) = 0;
  virtual void retune(const psx_tuning& t) = 0;
  virtual void get_state(double* pos, double* vel) = 0;
  virtual void set_state(const double* pos, const double* vel) = 0;
  virtual void get_state_device(void* pos, void* vel) = 0;
  virtual void set_state_device(const void* pos, const void* vel) = 0;
  virtual void debug_bucketing(int32_t* cell, uint32_t* count, uint32_t* offset, int32_t* perm) = 0;
  virtual uint32_t read_flags() = 0;
  virtual void set_params(int32_t scene, const psx_params& p) = 0;
  // returns the window's id (psx_window_forward)
  virtual int64_t window_forward(int64_t nsteps, int64_t every) = 0;
  virtual void window_backward(int64_t window_id, const void* dlx, const void* dlv, void* dlx0,
                               void* dlv0, double* gparams) = 0;
  virtual void window_stats(double* replay_ms, double* adjoint_ms) = 0;
  virtual void statistics(psx_stats* out) = 0;
  // statistics() with its per-workgroup partials and per-scene totals in the caller's device
  // buffer `scratch` of `bytes` bytes (no allocation); the totals are copied to out.
  virtual void statistics_into(psx_stats* out, void* scratch, size_t bytes) = 0;
};

std::unique_ptr<SimBase> make_sim(int dtype);
const char* kernel_name(int k);

}  // namespace psx
