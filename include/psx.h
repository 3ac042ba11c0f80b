/*
 * psx.h -- C ABI of libparticle_stabilizer_hip.so, the MI355X (gfx950) backend for the per-step
 * sphere-settling hot path of particle-stabilizer.
 *
 * Every entry point takes plain pointers and sizes, returns an int status (PSX_OK == 0) and never
 * calls exit(). The text of the last error on the calling thread is returned by psx_last_error().
 *
 * Reference interface each group replaces (paths relative to the reference repository):
 *   - runSimulateCUDA(...)   include/ParticleStabilizerCUDA/PhysicsEngineCUDA.hpp:43-54, the dlopen'd
 *     plugin entry (src/App/ParticleStabilizerMain.cpp:24-35,235-256). Its std::/glm signature is
 *     kept by runSimulateHIP in psx_compat.hpp; the POD entry points below are what that shim and
 *     every other caller (ctypes, the CLI, the tests) bind.
 *   - runSimulate(...)       include/ParticleStabilizer/PhysicsEngine.hpp:31-42 (CPU backend, same
 *     signature). Its step order (PhysicsEngine.cpp:38-48) is the semantics implemented here by
 *     default; PSX_FLAG_SEMANTICS_CUDA selects the reference CUDA backend's step instead.
 *
 * Threading: a psx_sim handle is not thread-safe; separate handles are independent. All device work
 * of a handle is ordered on one HIP stream (its own, or one injected with psx_set_stream).
 */
#ifndef PSX_H_
#define PSX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSX_ABI_VERSION 5

enum psx_status {
  PSX_OK = 0,
  PSX_ERR_INVALID = 1,  /* bad argument (null pointer, n <= 0, dt <= 0, min >= max, ...) */
  PSX_ERR_HIP = 2,      /* a HIP runtime call failed */
  PSX_ERR_ALLOC = 3,    /* device or host allocation failed */
  PSX_ERR_STATE = 4,    /* call not valid in the handle's current state */
  PSX_ERR_NUMERIC = 5,  /* non-finite particle state detected */
  PSX_ERR_IO = 6        /* file could not be read or written */
};

enum psx_dtype { PSX_F32 = 0, PSX_F64 = 1 };

/* Simulation constants. Mirrors the arguments of runSimulate / runSimulateCUDA
 * (PhysicsEngine.hpp:31-42) and the CLI defaults of ParticleStabilizerMain.cpp:41-57. */
typedef struct psx_params {
  double gravity[3];
  /* CoefficientOfRestitutionWall order (ContainerBase.hpp:10-34): zMin, zMax, xMin, xMax, yMin, yMax.
   * With a mesh container walls[0] (zMin) is the mesh restitution (PhysicsEngine.cpp:107-109). */
  double walls[6];
  double spring;       /* coefficientOfSpring */
  double restitution;  /* coefficientOfRestitutionSphere */
  double dt;
  double box_min[3];   /* box container; ignored when a mesh is given */
  double box_max[3];
  int32_t dtype;       /* enum psx_dtype: arithmetic type of the particle state on the device */
  int32_t flags;       /* 0 or PSX_FLAG_SEMANTICS_CUDA; other bits must be 0 */
} psx_params;

/* psx_params.flags: the step of the reference's CUDA backend (PhysicsEngineCUDA.cpp:36-43) instead of
 * its CPU backend: container before bucketing (membership and home cell from the post-container
 * position), cell width r_max * 2.0 * 1.05 (ParticleModelCUDA.cpp:84), the pair kick of
 * ParticleModel_kernels.cu:163-199 (spring term divided by m_i), the mesh contact of
 * PolygonContainerCUDA_kernels.cu:69-294 (hits d^2 < r^2 from the edge picked by dot(center, q),
 * position corrections from the incoming position summed, reflections averaged), normals
 * cross / length (PolygonContainerCUDA.cpp:64-82). The reference adds the kicks and the mesh terms
 * with atomics in no fixed order; here they are added in one of those orders: cells x-outer, y,
 * z-inner, members by ascending id, triangles by ascending index. Fixed at creation: every scene of
 * a batch must carry the same flag, and psx_set_params cannot change it. The differentiable window
 * implements the CPU semantics only. */
#define PSX_FLAG_SEMANTICS_CUDA 1

/* A triangle-soup container (PolygonContainer.cpp:8-90): n_triangles de-indexed triangles,
 * vertices[9*t + 3*k + a] = coordinate a of corner k of triangle t. Normals follow the winding,
 * normalize(cross(b - a, c - a)); particles are kept on the -normal side. */
typedef struct psx_mesh {
  int64_t n_triangles;
  const double* vertices;
} psx_mesh;

/* One scene of a batch: its own constants, container and particle count. */
typedef struct psx_scene {
  psx_params params;
  psx_mesh mesh;      /* n_triangles == 0: box container from params.box_min/max */
  int64_t n;          /* number of particles of this scene (its slice of the batch arrays) */
} psx_scene;

/* Launch shapes, chosen by the autotuner (psx/tune.py) or left at 0 for the built-in defaults. */
typedef struct psx_tuning {
  int32_t block_integrate;  /* workgroup size of the integrate / hash / count kernel */
  int32_t block_sort;       /* workgroup size of the rank / container / gather kernel
                               (at most 512 for PSX_F64) */
  int32_t block_pairs;      /* workgroup size of the pair-contact kernel */
  int32_t use_graph;        /* 1: replay steps from a captured hipGraph */
  int32_t pair_variant;     /* pair-contact kernel: lanes per particle 0 or 1, 2, 4, 8, 16; or 32: one
                               lane per particle with the workgroup's neighbour particles staged in
                               LDS (block_pairs up to 512, 256 in binary64); or 64: one lane per
                               particle with each wave's neighbour columns staged in LDS
                               (block_pairs up to 512 in binary64) */
  int32_t sort_variant;     /* cell lists of the bucketing: 1 per-cell id slots (three kernels per
                               step in a run: scan, rank/gather, pairs), 2 dense lists filled by a
                               scatter kernel (four), 0 automatic (id slots), 3 incremental id
                               slots: counts and per-tile totals kept from step to step, only the
                               spheres that change cell update them, and a scan from the kept tile
                               totals (with PSX_FLAG_SEMANTICS_CUDA as 1). Results are the same for
                               every value. */
  int32_t pair_chunk;       /* candidates per round of the one-lane pair kernel (pair_variant 0 or 1):
                               0 or 4, or 2 (fewer wasted tests where column slices hold few spheres) */
} psx_tuning;

typedef struct psx_sim psx_sim;

/* ---- lifetime ------------------------------------------------------------------------------ */

/* Builds a single-scene simulation (runSimulateCUDA's ParticleModelCUDA + container construction,
 * PhysicsEngineCUDA.cpp:152-171). pos/vel are n x 3 (row-major), radius/mass are n; all host memory,
 * copied. mesh may be NULL (box container). The grid follows ParticleBucket::initBucket
 * (ParticleBucket.cpp:17-38) over the container bounds. */
int psx_create(const psx_params* params, int64_t n, const double* pos, const double* vel,
               const double* radius, const double* mass, const psx_mesh* mesh, psx_sim** out);

/* Builds a batch of independent scenes advanced together, one launch per kernel per step.
 * Particle arrays are the concatenation of the scenes' particles in scene order. All scenes share
 * one dtype (scenes[0].params.dtype). No counterpart in the reference: the reference runs one
 * scene per process. */
int psx_batch_create(int32_t n_scenes, const psx_scene* scenes, const double* pos,
                     const double* vel, const double* radius, const double* mass, psx_sim** out);

int psx_destroy(psx_sim* sim);

/* ---- stepping ------------------------------------------------------------------------------ */

/* Enqueues nsteps steps (PhysicsEngine::startSimulationLoop's body, PhysicsEngine.cpp:38-48) on the
 * handle's stream and returns without waiting. Statistics are not computed on this path. */
int psx_step(psx_sim* sim, int64_t nsteps);

/* Captures the hipGraphs that psx_step(sim, nsteps) would replay from the current state, without
 * running a step (the current state's bucketing hash pass is enqueued, as psx_step would): 64- and
 * 16-step graphs, and one of the 1-15 steps left after them, which psx_step replays only where it
 * was prepared (otherwise those steps are launched one by one). A run whose graphs are prepared
 * spends no time capturing or instantiating them and launches no step on its own. No counterpart in
 * the reference (its loop launches every kernel from the host, PhysicsEngineCUDA.cpp:25-81). */
int psx_prepare(psx_sim* sim, int64_t nsteps);

int psx_synchronize(psx_sim* sim);

/* Injects a caller-owned HIP stream (hipStream_t passed as void*); NULL restores the own stream. */
int psx_set_stream(psx_sim* sim, void* hip_stream);

int psx_set_tuning(psx_sim* sim, const psx_tuning* tuning);
int psx_get_tuning(const psx_sim* sim, psx_tuning* tuning);

/* ---- state --------------------------------------------------------------------------------- */

/* Copies the state out in the caller's original particle order (runSimulate's copy-out,
 * PhysicsEngine.cpp:148-159). Waits for the stream. Any of the pointers may be NULL. */
int psx_get_state(psx_sim* sim, double* pos, double* vel);
int psx_set_state(psx_sim* sim, const double* pos, const double* vel);

/* Device-memory variants for PyTorch interop: n x 3 arrays of the handle's dtype, original order,
 * stream-ordered on the handle's stream (no host wait). */
int psx_get_state_device(psx_sim* sim, void* pos_dev, void* vel_dev);
int psx_set_state_device(psx_sim* sim, const void* pos_dev, const void* vel_dev);

int64_t psx_num_particles(const psx_sim* sim);
int32_t psx_num_scenes(const psx_sim* sim);
int64_t psx_num_cells(const psx_sim* sim);     /* total grid cells over all scenes */
int32_t psx_dtype_of(const psx_sim* sim);

/* Grid geometry of one scene: min corner, cell size h, cell counts (x, y, z), first global cell. */
int psx_get_grid(const psx_sim* sim, int32_t scene, double* min3, double* h, int64_t* n3,
                 int64_t* cell_base);

/* ---- bucketing (ParticleModel::updateBuckets, ParticleModel.cpp:49-62) ------------------------- */

/* Buckets the current positions without stepping and returns, in the caller's original order:
 * cell[i] (linear cell id (ix*ny + iy)*nz + iz plus the scene's cell_base), count[c] and the
 * exclusive scan offset[c] (C+1 entries, offset[C] == n), and perm[k] = particle id at sorted slot k
 * (stable: ascending id inside a cell). Any output may be NULL. */
int psx_debug_bucketing(psx_sim* sim, int32_t* cell, uint32_t* count, uint32_t* offset,
                        int32_t* perm);

/* ---- statistics (ParticleStatistics::calculateStatistics, ParticleStatistics.cpp:3-43) --------- */

typedef struct psx_stats {
  double kinetic_energy;  /* sum of 0.5 m (v . v) */
  double min[3];          /* AABB of the spheres: min of (x - r), max of (x + r) per axis */
  double max[3];
  int64_t n_collided;     /* particles i with a partner j > i (original index) such that
                             r_i + r_j - |x_j - x_i| > 0 */
  double max_overlap;     /* the largest such r_i + r_j - |x_j - x_i|; 0 when there is none */
} psx_stats;

/* Statistics of the current state of every scene into out[0 .. num_scenes). The pair part runs on a
 * bucketing of the current positions (O(N), the reference's loop is O(N^2)); everything is evaluated
 * in binary64 and combined in a fixed order (the kinetic-energy sum is therefore deterministic but
 * not the reference's sequential rounding). Waits for the stream. */
int psx_statistics(psx_sim* sim, psx_stats* out);

/* psx_statistics without any allocation, for callers that cannot allocate device memory (tight
 * serving loops, harnesses that forbid hipMalloc). No counterpart in the reference. Its
 * per-workgroup partials and per-scene totals live in a fixed 4 MiB device array that each GPU
 * running the library holds until the process ends, so one call takes at most about 14 million
 * particles (PSX_ERR_INVALID beyond); the totals are combined on the device in psx_statistics'
 * order and equal its results bit for bit. Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other threads, other handles) race. Waits for the stream. */
int psx_statistics_noalloc(psx_sim* sim, psx_stats* out);

/* ---- differentiable windows ---------------------------------------------------------------- */
/* No counterpart in the reference. Gradients of a loss on the state after a window of steps with
 * respect to the initial state and to spring, sphere restitution and gravity_z, holding each step's
 * contact sets (pair contacts, wall and triangle hits) fixed. */

/* Replaces the coefficients of one scene (gravity, walls, spring, restitution, dt); the container
 * bounds, dtype and flags must be unchanged. */
int psx_set_params(psx_sim* sim, int32_t scene, const psx_params* params);

/* Runs nsteps steps from the current state, keeping a copy of the state every checkpoint_every
 * steps in device memory for psx_window_backward. *window_id (may be NULL) receives the window's
 * id; a handle holds the checkpoints of its latest window only. PSX_ERR_STATE for a handle created
 * with PSX_FLAG_SEMANTICS_CUDA. */
int psx_window_forward(psx_sim* sim, int64_t nsteps, int64_t checkpoint_every,
                       int64_t* window_id);

/* Adjoint of the forward window window_id, which must be the handle's latest window with no
 * psx_set_params since (otherwise PSX_ERR_STATE: its checkpoints or coefficients are gone).
 * dl_dpos/dl_dvel: dL/d(final positions, velocities), n x 3 of the handle's dtype in device memory,
 * original order. Writes dL/d(initial positions, velocities) to dl_dpos0/dl_dvel0 (same layout)
 * and, for each scene s, dL/d(spring, restitution, gravity_z) to dl_dparams[3s .. 3s+2] (host).
 * Each segment between checkpoints is recomputed forward from its checkpoint; the handle's current
 * state is left as it is. */
int psx_window_backward(psx_sim* sim, int64_t window_id, const void* dl_dpos, const void* dl_dvel,
                        void* dl_dpos0, void* dl_dvel0, double* dl_dparams);

/* Device time of the last psx_window_backward spent recomputing segments and in the adjoint steps. */
int psx_window_stats(const psx_sim* sim, double* replay_ms, double* adjoint_ms);

/* ---- measurement --------------------------------------------------------------------------- */

/* Runs nsteps steps with a HIP event pair around every kernel launch on the handle's stream and
 * returns the mean device time per launch of each kernel (ms) in kernel_ms[0..*n_kernels). Kernel
 * k's name is psx_kernel_name(k). Launches per step of each kernel go to launches_per_step. */
int psx_profile(psx_sim* sim, int64_t nsteps, double* kernel_ms, double* launches_per_step,
                int32_t max_kernels, int32_t* n_kernels);
const char* psx_kernel_name(int32_t k);

/* ---- host file formats (no GPU involved) --------------------------------------------------- */

/* ParticleData::readFromCSV (ParticleIO.cpp:90-144): header line skipped, tokens split on ',',
 * empty and "null" tokens dropped, lines without exactly 4 tokens reported on stderr and skipped
 * (counted in *n_skipped), x,y,z,r parsed with std::stod; velocity3 is every particle's initial
 * velocity, mass = density*4/3*pi*r^3. Arrays are malloc'd (release with psx_free). A missing file
 * returns PSX_ERR_IO (the reference exits). */
int psx_read_csv(const char* path, double density, const double* velocity3, int64_t* n,
                 double** pos, double** vel, double** radius, double** mass, int64_t* n_skipped);

/* ParticleData::writeAsCSV (ParticleIO.cpp:58-88): header "x,y,z,radius", each value in fixed
 * notation with 10 - len(str(int(|v|))) - [v < 0] - 1 decimals; parent directories created. */
int psx_write_csv(const char* path, int64_t n, const double* pos, const double* radius);

/* Triangle container from a Wavefront OBJ file (v / f records; polygons split as a fan from their
 * first corner), returned de-indexed in face order as PolygonContainer::readFromFile builds it
 * (PolygonContainer.cpp:381-417). *vertices holds 9 * *n_triangles doubles (psx_free). */
int psx_read_obj(const char* path, int64_t* n_triangles, double** vertices);

void psx_free(void* p);

/* ---- post-run outputs (host, after a simulation; ParticleStabilizerMain.cpp:275-296) ----------- */

/* ParticleModel::resolveOverlaps (ParticleModel.cpp:154-252): lifts overlapping particles along
 * `axis` (0, 1, 2) in the order of their coordinate on that axis; pos (n x 3) is updated in place.
 * Afterwards *n_collided / *max_overlap (either may be NULL) receive the overlap check the reference
 * logs: particles with a partner of higher index closer than r_i + r_j, and the largest overlap. */
int psx_resolve_overlaps(int64_t n, double* pos, const double* radius, int32_t axis,
                         int64_t* n_collided, double* max_overlap);

/* ObjData::write (ObjData.cpp:8-140): every particle as a triangulated UV sphere in a Wavefront OBJ
 * file (shared vertex normals; resolution ndivs = max((int)sqrt(1e6 / (6 + n)) + 1, 6)); parent
 * directories are created. */
int psx_write_obj_spheres(const char* path, int64_t n, const double* pos, const double* radius);

/* ---- errors / version ---------------------------------------------------------------------- */

const char* psx_last_error(void);
int psx_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PSX_H_ */
