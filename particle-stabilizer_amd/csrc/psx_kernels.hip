// psx_kernels.hip -- the per-step kernels of the HIP backend (gfx950, wave64).
//
// One step (PhysicsEngine.cpp:38-48, CPU order: integrate, bucket from the pre-container positions,
// container, pairs) is these launches on one stream (in a run of steps, k_integrate's work for the
// next step is done by k_pairs at its end, so k_integrate only starts a run):
//   k_integrate    x += v*dt; cell hash of the new (pre-container) x; per-cell count
//                  (atomic, returns the particle's arrival slot), and the particle's id written to
//                  that slot of its cell's id slots (kCellIds slot-major arrays; arrivals past
//                  them go to an overflow list)
//   k_scan         exclusive scan of the per-cell counts into offsets, one pass, wave64 scans
//                  inside a tile; each tile's prefix is summed from the per-tile counts of the
//                  hash pass (k_scan_direct, grids up to 2^20 cells) or found by decoupled
//                  look-back (k_scan); clears the counts (id blocks; with the dense lists
//                  k_scatter_ids counts them back down to zero)
//   k_rank_gather  rank of each particle among its cell mates by id (stable order); x += v*dt;
//                  v += g*dt (the same operations again); box walls or mesh contact; gather of the
//                  result into cell-sorted order
//   k_pairs        27-cell spring/restitution velocity kick, Jacobi over the sorted snapshot
// Particle state lives in cell-sorted order between steps; an id array carries original indices.
//
// Workgroups of the per-particle kernels are numbered XCD-major: the GPU deals workgroups to its 8
// XCDs round-robin, so workgroup b is given tile (b % 8) * ceil(tiles / 8) + b / 8, and each XCD
// (with its own L2) sweeps one contiguous eighth of the cell-sorted particles, where neighbour
// columns are close in memory.
#include <hip/hip_ext.h>

#include <cstring>

#include "psx_bucket.hpp"
#include "psx_kernels.hpp"

namespace psx {

// Timing events for the next launch (set by the profiler): the launch is then issued with
// hipExtLaunchKernelGGL, whose start/stop events take the dispatch's own begin/end timestamps.
static thread_local hipEvent_t g_ev_start = nullptr, g_ev_stop = nullptr;

void set_launch_events(hipEvent_t start, hipEvent_t stop) {
  g_ev_start = start;
  g_ev_stop = stop;
}

#define PSX_LAUNCH(kern, grid, block, shm, st, ...)                                           \
  do {                                                                                       \
    if (g_ev_start) {                                                                        \
      hipExtLaunchKernelGGL(kern, grid, block, shm, st, g_ev_start, g_ev_stop, 0, __VA_ARGS__); \
      g_ev_start = g_ev_stop = nullptr;                                                      \
    } else {                                                                                 \
      hipLaunchKernelGGL(kern, grid, block, shm, st, __VA_ARGS__);                           \
    }                                                                                        \
  } while (0)

// ---------------------------------------------------------------------------------------------
// wave64 and launch helpers
// ---------------------------------------------------------------------------------------------
constexpr int kXcds = 8;

__host__ __device__ inline int32_t xcd_grid(int32_t tiles) {
  return ((tiles + kXcds - 1) / kXcds) * kXcds;
}

__device__ inline int32_t xcd_tile(int32_t b, int32_t tiles) {
  const int32_t per = (tiles + kXcds - 1) / kXcds;
  return (b % kXcds) * per + b / kXcds;
}


__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// How a workgroup finds its particles and its scene's constants. A batch of scenes has one table
// entry per workgroup (scene, particle slice). A single scene needs no table: its constants travel
// by value in the kernel arguments and a workgroup's slice follows from its index, which takes a
// dependent load (table entry, then scene record) off the start of every workgroup.
template <class T>
struct Locator {
  const BlockInfo* btab;       // batches: one entry per workgroup
  const SceneDev<T>* scenes;   // batches: the scene records
  int32_t ntab;                // workgroups with work
  int32_t per_block;           // single scene: particles per workgroup
  SceneDev<T> one;             // single scene: its constants
};

// Runs body(scene, p0, p1) for the workgroup's scene and particle slice [p0, p1); padding
// workgroups return. The scene is passed by reference straight from the table or from the kernel
// arguments, so no pointer has to merge the two address spaces.
template <class T, bool kTable, class Body>
__device__ __forceinline__ void with_slice(const Locator<T>& loc, Body&& body) {
  const int32_t tb = xcd_tile(blockIdx.x, loc.ntab);
  if (tb >= loc.ntab) return;
  if constexpr (kTable) {
    const BlockInfo bi = loc.btab[tb];
    body(loc.scenes[bi.scene], bi.p_start, bi.p_end);
  } else {
    const int32_t p0 = tb * loc.per_block;
    body(loc.one, p0, min(p0 + loc.per_block, loc.one.p_end));
  }
}

// ---------------------------------------------------------------------------------------------
// container contact (per particle)
// ---------------------------------------------------------------------------------------------

// BoxContainer::resolveCollisions (BoxContainer.cpp:30-65): lo/hi from the incoming position, then
// xmin, ymin, zmin, xmax, ymax, zmax.
template <class T>
__device__ __forceinline__ void box_walls(const SceneDev<T>* __restrict__ s, Vec4<T>& x, Vec4<T>& v) {
  const T r = x.w;
  const T lox = x.x - r, loy = x.y - r, loz = x.z - r;
  const T hix = x.x + r, hiy = x.y + r, hiz = x.z + r;
  if (lox < s->bmin[0]) {
    x.x = s->bmin[0] + r;
    v.x = -v.x * s->emin[0];
  }
  if (loy < s->bmin[1]) {
    x.y = s->bmin[1] + r;
    v.y = -v.y * s->emin[1];
  }
  if (loz < s->bmin[2]) {
    x.z = s->bmin[2] + r;
    v.z = -v.z * s->emin[2];
  }
  if (hix > s->bmax[0]) {
    x.x = s->bmax[0] - r;
    v.x = -v.x * s->emax[0];
  }
  if (hiy > s->bmax[1]) {
    x.y = s->bmax[1] - r;
    v.y = -v.y * s->emax[1];
  }
  if (hiz > s->bmax[2]) {
    x.z = s->bmax[2] - r;
    v.z = -v.z * s->emax[2];
  }
}

// PolygonContainer::resolveCollisions (PolygonContainer.cpp:178-286) for one particle. Instead of
// all T triangles, only the list of the particle's grid cell is visited; it holds, in ascending
// order, every triangle a particle hashed to that cell can touch, and the triangles it omits are
// exactly ones whose contact test fails (they change nothing), so the sequential position update
// and the velocity sum see the same triangles in the same order as the full loop.
//
// Contact state of one particle: the running position y, the velocity sum and the hit count.
template <class T>
struct MeshAcc {
  Vec3<T> y;
  T ax, ay, az;
  int32_t cnt;
};

// One hit: acc += e * (v - 2 (v.n) n); y -= (r - dot(-n, y - a)) n (PolygonContainer.cpp:195-211).
template <class T>
__device__ __forceinline__ void mesh_hit(MeshAcc<T>& m, T e, T r, T vx, T vy, T vz, T nx, T ny,
                                         T nz, T a0, T a1, T a2) {
  const T s2 = T(2) * dot3(vx, vy, vz, nx, ny, nz);
  m.ax = m.ax + e * (vx - s2 * nx);
  m.ay = m.ay + e * (vy - s2 * ny);
  m.az = m.az + e * (vz - s2 * nz);
  const T dtp = dot3(-nx, -ny, -nz, m.y.x - a0, m.y.y - a1, m.y.z - a2);
  const T mod = r - dtp;
  m.y.x = m.y.x - mod * nx;
  m.y.y = m.y.y - mod * ny;
  m.y.z = m.y.z - mod * nz;
  ++m.cnt;
}

// AABB reject of one candidate (PolygonContainer.cpp:251-259).
template <class T>
__device__ __forceinline__ bool mesh_aabb(const T (&lo)[3], const T (&hi)[3], const Vec3<T>& P,
                                          T r) {
  return !(P.x - r > hi[0] || P.x + r < lo[0] || P.y - r > hi[1] || P.y + r < lo[1] ||
           P.z - r > hi[2] || P.z + r < lo[2]);
}

// Closest point on the triangle and the contact test |q - x| < r (PolygonContainer.cpp:260-263).
template <class T>
__device__ __forceinline__ bool mesh_contact(const TriGeo<T>& g, const Vec3<T>& P, T r) {
  const Vec3<T> q = closest_on_triangle(g, P);
  return dist_to(P, q) < r;
}

template <class T>
__device__ __forceinline__ void mesh_finish(const MeshAcc<T>& m, Vec4<T>& x, Vec4<T>& v) {
  if (m.cnt > 0) {  // PolygonContainer.cpp:280-285
    const T cf = (T)m.cnt;
    x.x = m.y.x;
    x.y = m.y.y;
    x.z = m.y.z;
    v.x = m.ax / cf;
    v.y = m.ay / cf;
    v.z = m.az / cf;
  }
}

// Lists longer than this are walked by the whole wave: 64 lanes test 64 candidates at once and the
// hits are then applied in ascending order, one after the other, exactly as a single lane would.
// Shorter lists are walked by their own lane four candidates at a time: the four AABBs are loaded
// together (the entry array carries kTriPad spare entries for the last chunk) and the geometry of
// each candidate that passes is read in candidate order. The contact test uses the particle's
// original position (PolygonContainer.cpp:260-263), so only the hits' effect is sequential.
constexpr int32_t kHeavyList = 12;

// k0, k1: the particle's list [k0, k1) in tent (any values when !active).
template <class T>
__device__ __forceinline__ void mesh_contact_wave(const SceneDev<T>* __restrict__ s, bool active,
                                                  int32_t k0, int32_t k1,
                                                  const TriEntry<T>* __restrict__ tent,
                                                  Vec4<T>& x, Vec4<T>& v) {
  const T e = s->ew;
  const bool heavy = active && (k1 - k0) > kHeavyList;
  if (active && !heavy && k1 > k0) {
    const Vec3<T> P = {x.x, x.y, x.z};
    const T r = x.w;
    MeshAcc<T> m = {P, T(0), T(0), T(0), 0};
    for (int32_t kb = k0; kb < k1; kb += 4) {
      const TriEntry<T>* __restrict__ c4 = tent + kb;
      uint32_t pass = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        T lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          lo[a] = c4[j].lo[a];
          hi[a] = c4[j].hi[a];
        }
        if (kb + j < k1 && mesh_aabb(lo, hi, P, r)) pass |= 1u << j;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (pass & (1u << j)) {
          const TriGeo<T> g = c4[j].g;
          if (mesh_contact(g, P, r))
            mesh_hit(m, e, r, v.x, v.y, v.z, g.n[0], g.n[1], g.n[2], g.a[0], g.a[1], g.a[2]);
        }
      }
    }
    mesh_finish(m, x, v);
  }
  uint64_t hm = __ballot(heavy);
  while (hm) {
    const int owner = __ffsll((unsigned long long)hm) - 1;
    hm &= hm - 1;
    const Vec3<T> P = {__shfl(x.x, owner), __shfl(x.y, owner), __shfl(x.z, owner)};
    const T r = __shfl(x.w, owner);
    const T vx = __shfl(v.x, owner), vy = __shfl(v.y, owner), vz = __shfl(v.z, owner);
    const int32_t b0 = __shfl(k0, owner), b1 = __shfl(k1, owner);
    MeshAcc<T> m = {P, T(0), T(0), T(0), 0};
    for (int32_t base = b0; base < b1; base += 64) {
      const int32_t k = base + lane_id();
      T n0 = T(0), n1 = T(0), n2 = T(0), a0 = T(0), a1 = T(0), a2 = T(0);
      bool hit = false;
      if (k < b1) {  // only the normal and the first corner outlive the test
        const TriEntry<T>* __restrict__ ep = tent + k;
        T lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          lo[a] = ep->lo[a];
          hi[a] = ep->hi[a];
        }
        if (mesh_aabb(lo, hi, P, r)) {
          const TriGeo<T> g = ep->g;
          hit = mesh_contact(g, P, r);
          n0 = g.n[0];
          n1 = g.n[1];
          n2 = g.n[2];
          a0 = g.a[0];
          a1 = g.a[1];
          a2 = g.a[2];
        }
      }
      uint64_t hits = __ballot(hit);
      while (hits) {  // ascending lane = ascending candidate
        const int src = __ffsll((unsigned long long)hits) - 1;
        hits &= hits - 1;
        mesh_hit(m, e, r, vx, vy, vz, __shfl(n0, src), __shfl(n1, src), __shfl(n2, src),
                 __shfl(a0, src), __shfl(a1, src), __shfl(a2, src));
      }
    }
    if (lane_id() == owner) mesh_finish(m, x, v);
  }
}

// PhysicsEngine::updatePositionAndVelocity (PhysicsEngine.cpp:22-25): x += v dt with the old v,
// then v += g dt. The hash kernel needs the new positions, the rank/gather kernel the whole state;
// both evaluate the same operations, so they agree bit for bit.
template <class T>
__device__ __forceinline__ void integrate_position(const SceneDev<T>& s, Vec4<T>& x,
                                                   const Vec4<T>& v) {
  const T dt = s.dt;
  x.x = x.x + v.x * dt;
  x.y = x.y + v.y * dt;
  x.z = x.z + v.z * dt;
}

template <class T>
__device__ __forceinline__ void integrate_velocity(const SceneDev<T>& s, Vec4<T>& v) {
  v.x = v.x + s.gdt[0];
  v.y = v.y + s.gdt[1];
  v.z = v.z + s.gdt[2];
}

// The pair kernel's hash pass for the next step (NextBucket): x + v_new dt from the snapshot
// position and the new velocity, then the bucket entry, exactly as the hash kernel computes them
// from the state the next step starts from. Returns the cell (tile counts are added per workgroup
// by WgTiles).
// With the incremental bucketing only a particle whose next cell differs from its bucket cell
// counts and files itself (bucket_move); every particle writes its next cell. hw: the particle's
// home word (nb.home[p]), which the pair kernels have loaded for their walk.
template <class T>
__device__ __forceinline__ int32_t next_bucket(const SceneDev<T>& s, const NextBucket& nb,
                                               Vec4<T> x, const Vec4<T>& v, int32_t p,
                                               uint32_t hw) {
  integrate_position(s, x, v);
  if (nb.arr) {
    const int32_t c = grid_cell(s, x);
    nb.cell_of[p] = c;
    const int32_t c_old = (hw & kHomeMoved) ? nb.cell_cur[p] : (int32_t)(hw & kHomeCell);
    if (c != c_old)
      bucket_move(nb.id[p], c, c_old, nb.ovf_ctr[2] + 1u, nb.count, nb.arr, nb.mark, nb.tt,
                  nb.cell_ids, nb.id_stride, nb.ovf, nb.ovf_ctr);
    return c;
  }
  return bucket_entry(s, x, nb.id[p], p, nb.blocks != 0, nb.cell_of, nb.count, nb.cell_ids,
                      nb.id_stride, nb.ovf, nb.ovf_ctr);
}

// Tile counts of a hash-pass workgroup (TileCounts). Its particles are consecutive in cell order
// and move little in a step, so their cells lie almost all in the tiles around the first
// particle's cell: they are summed in LDS slots for the kWgTiles tiles from one below that tile
// and added to the global counts once per slot; cells outside the slots (particles crossing into
// a distant x or y column) go to the global counts directly. Every thread calls begin(), add()
// (with c < 0 where it holds no particle) and end().
constexpr int kWgTiles = 4;
struct WgTiles {
  uint32_t cnt[kWgTiles];
  int32_t base;

  // cell0: a cell of the workgroup's first particle (thread 0's value is used)
  __device__ void begin(int32_t cell0) {
    if (threadIdx.x < kWgTiles) cnt[threadIdx.x] = 0u;
    if (threadIdx.x == 0) base = (cell0 >> kDirectShift) - 1;
    __syncthreads();
  }
  __device__ void put(uint32_t* __restrict__ tc, int32_t t, uint32_t k) {
    const int32_t slot = t - base;
    if (slot >= 0 && slot < kWgTiles) atomicAdd(&cnt[slot], k);
    else atomicAdd(tc + (size_t)t * kTileStride, k);
  }
  // lanes with c >= 0 take part; one addition per wave when they share a tile
  __device__ void add(const TileCounts& tiles, int32_t c) {
    uint32_t* __restrict__ tc =
        tiles.tc + (size_t)(*tiles.gen & 1u) * (size_t)tiles.ntiles * kTileStride;
    const bool has = c >= 0;
    const uint64_t act = __ballot(has);
    if (!has) return;
    const int leader = __ffsll((unsigned long long)act) - 1;
    const int32_t t = c >> kDirectShift;
    const int32_t t0 = __shfl(t, leader, 64);
    if (__ballot(t == t0) == act) {
      if (lane_id() == leader) put(tc, t0, (uint32_t)__popcll(act));
    } else {
      put(tc, t, 1u);
    }
  }
  __device__ void end(const TileCounts& tiles) {
    __syncthreads();
    if (threadIdx.x < kWgTiles && cnt[threadIdx.x] != 0u) {
      uint32_t* __restrict__ tc =
          tiles.tc + (size_t)(*tiles.gen & 1u) * (size_t)tiles.ntiles * kTileStride;
      atomicAdd(tc + (size_t)(base + (int32_t)threadIdx.x) * kTileStride, cnt[threadIdx.x]);
    }
  }
};

// ---------------------------------------------------------------------------------------------
// K1: integrate + cell hash + counts
// ---------------------------------------------------------------------------------------------
template <class T, bool kTable, bool kBlocks>
__global__ void __launch_bounds__(1024)
    k_integrate(const Locator<T> loc, const Vec4<T>* __restrict__ pos,
                const Vec4<T>* __restrict__ vel, const int32_t* __restrict__ id,
                int32_t* __restrict__ cell_of, uint32_t* __restrict__ count,
                int32_t* __restrict__ cell_ids, uint32_t id_stride, int2* __restrict__ ovf,
                uint32_t* __restrict__ ovf_ctr, const TileCounts tiles, int mode,
                unsigned long long* __restrict__ arr, uint8_t* __restrict__ mark) {
  __shared__ WgTiles wt;
  with_slice<T, kTable>(loc, [&](const SceneDev<T>& s, int32_t p0, int32_t p1) {
    const int32_t p = p0 + (int32_t)threadIdx.x;
    int32_t c = -1;
    if (p < p1) {
      Vec4<T> x = pos[p];
      const int32_t me = id[p];
      if (mode & kModeIntegrate) {  // PhysicsEngine.cpp:22-25 (positions only: the hash needs them)
        const Vec4<T> v = vel[p];
        integrate_position(s, x, v);
      }
      if (kBlocks && arr) {  // first pass of the incremental bucketing: every particle arrives
        c = grid_cell(s, x);
        cell_of[p] = c;
        bucket_move(me, c, -1, ovf_ctr[2] + 1u, count, arr, mark, TileTotals{}, cell_ids, id_stride,
                    ovf, ovf_ctr);
      } else {
        c = bucket_entry(s, x, me, p, kBlocks, cell_of, count, cell_ids, id_stride, ovf, ovf_ctr);
      }
    }
    if (tiles.tc) {
      wt.begin(c);
      wt.add(tiles, c);
      wt.end(tiles);
    }
  });
}

// The mesh container of the reference's CUDA backend for one particle at its integrated position
// x (PolygonContainerCUDA::resolveCollisions, PolygonContainerCUDA.cpp:130-214, with the kernels of
// PolygonContainerCUDA_kernels.cu:159-294): a triangle is hit when its AABB overlaps the particle's
// (<= and >=) and the point q that getClosestPointOnPolygonKernelFunc returns (the edge point
// picked by dot(x, q)) has |q - x|^2 < r^2. The velocity becomes the mean of e (v - 2 (v.n) n)
// over the hits, the position x + the sum of -(r - dot(-n, x - a)) n, both from the incoming x and
// v. Only the list of x's grid cell is visited: it holds every triangle within reach of the cell,
// which includes every triangle the test can hit. The reference adds the terms with atomics in no
// fixed order; here they are added in ascending triangle order. modifyPositionKernel adds the sum
// to every particle, hits or not, and so does this.
template <class T>
__device__ __forceinline__ void mesh_contact_cuda(const SceneDev<T>& s,
                                                  const TriEntry<T>* __restrict__ tent,
                                                  const int32_t* __restrict__ tcell_off,
                                                  const uint8_t* __restrict__ cflags, Vec4<T>& x,
                                                  Vec4<T>& v) {
  const GridDev& g = s.grid;
  const int32_t ix = axis_index((double)x.x, g.min[0], g.h, g.n[0]);
  const int32_t iy = axis_index((double)x.y, g.min[1], g.h, g.n[1]);
  const int32_t iz = axis_index((double)x.z, g.min[2], g.h, g.n[2]);
  const int32_t c = g.cell_base + (ix * g.n[1] + iy) * g.n[2] + iz;
  int32_t k0 = 0, k1 = 0;
  if (cflags[c] & kCellTris) {
    const int32_t tc = s.tcell_base + (c - g.cell_base);
    k0 = tcell_off[tc];
    k1 = tcell_off[tc + 1];
  }
  const Vec3<T> P = {x.x, x.y, x.z};
  const T r = x.w;
  const T lox = P.x - r, loy = P.y - r, loz = P.z - r;
  const T hix = P.x + r, hiy = P.y + r, hiz = P.z + r;
  T vbx = T(0), vby = T(0), vbz = T(0), mx = T(0), my = T(0), mz = T(0);
  int32_t cnt = 0;
  for (int32_t k = k0; k < k1; ++k) {  // isIntersectedKernel (PolygonContainerCUDA_kernels.cu:179-211)
    const TriEntry<T>* __restrict__ te = tent + k;
    if (!(lox <= te->hi[0] && hix >= te->lo[0] && loy <= te->hi[1] && hiy >= te->lo[1] &&
          loz <= te->hi[2] && hiz >= te->lo[2]))
      continue;
    const TriGeo<T> tg = te->g;
    const Vec3<T> q = closest_on_triangle<T, true>(tg, P);
    const T rx = q.x - P.x, ry = q.y - P.y, rz = q.z - P.z;
    if (!(dot3(rx, ry, rz, rx, ry, rz) < r * r)) continue;
    // calcModifiedVelocitySumKernel (:214-236): e (v - (2 v.n) n)
    const T tvn = T(2) * dot3(v.x, v.y, v.z, tg.n[0], tg.n[1], tg.n[2]);
    vbx = vbx + s.ew * (v.x - tvn * tg.n[0]);
    vby = vby + s.ew * (v.y - tvn * tg.n[1]);
    vbz = vbz + s.ew * (v.z - tvn * tg.n[2]);
    // calcPosModificationAmountKernel (:254-276): -(r - dot(-n, x - a)) n
    const T sc = r - dot3(-tg.n[0], -tg.n[1], -tg.n[2], P.x - tg.a[0], P.y - tg.a[1], P.z - tg.a[2]);
    mx = mx + (-sc) * tg.n[0];
    my = my + (-sc) * tg.n[1];
    mz = mz + (-sc) * tg.n[2];
    ++cnt;
  }
  if (cnt > 0) {  // calcAveragedVelocityKernel (:238-252)
    const T kc = (T)cnt;
    v.x = vbx / kc;
    v.y = vby / kc;
    v.z = vbz / kc;
  }
  x.x = x.x + mx;  // modifyPositionKernel (:278-286)
  x.y = x.y + my;
  x.z = x.z + mz;
}

// K1 of the reference CUDA backend's step (PSX_FLAG_SEMANTICS_CUDA; PhysicsEngineCUDA.cpp:38-42):
// integration (PhysicsEngine_kernels.cu:3-18, the same operations as the CPU backend), the container
// (box walls as BoxContainer_kernels.cu:3-56, which match BoxContainer.cpp:22-65, or the CUDA mesh
// contact), the state written back in place, then the cell and count of the post-container
// position: membership and home cell both follow it (ParticleModelCUDA.cpp:122-177,
// ParticleModel_kernels.cu:74-121). The reference's count kernel does not clamp the cell
// (ParticleModel_kernels.cu:49-56), so a position outside the grid indexes outside its counters;
// here it is clamped into the grid as the neighbour search clamps it (:61-76).
template <class T, bool kTable, bool kBlocks>
__global__ void __launch_bounds__(1024)
    k_front(const Locator<T> loc, Vec4<T>* __restrict__ pos, Vec4<T>* __restrict__ vel,
            const int32_t* __restrict__ id, const TriEntry<T>* __restrict__ tent,
            const int32_t* __restrict__ tcell_off, const uint8_t* __restrict__ cflags,
            int32_t* __restrict__ cell_of, uint32_t* __restrict__ count,
            int32_t* __restrict__ cell_ids, uint32_t id_stride, int2* __restrict__ ovf,
            uint32_t* __restrict__ ovf_ctr, const TileCounts tiles) {
  __shared__ WgTiles wt;
  with_slice<T, kTable>(loc, [&](const SceneDev<T>& s, int32_t p0, int32_t p1) {
    const int32_t p = p0 + (int32_t)threadIdx.x;
    int32_t c = -1;
    if (p < p1) {
      Vec4<T> x = pos[p], v = vel[p];
      const int32_t me = id[p];
      integrate_position(s, x, v);
      integrate_velocity(s, v);
      if (s.mesh >= 0) mesh_contact_cuda(s, tent, tcell_off, cflags, x, v);
      else box_walls(&s, x, v);
      pos[p] = x;
      vel[p] = v;
      c = bucket_entry(s, x, me, p, kBlocks, cell_of, count, cell_ids, id_stride, ovf, ovf_ctr);
    }
    if (tiles.tc) {
      wt.begin(c);
      wt.add(tiles, c);
      wt.end(tiles);
    }
  });
}

// Dense cell lists (the scatter form of the bucketing): each particle takes a slot of its cell by
// counting the cell's count back down (count - 1, ..., 0), and its id goes to offset[cell] + slot:
// an unordered list per cell that the rank/gather kernel orders by id. The counts end at zero,
// ready for the next hash pass.
__global__ void __launch_bounds__(1024)
    k_scatter_ids(int32_t n, int32_t tiles, const int32_t* __restrict__ cell_of,
                  uint32_t* __restrict__ count, const uint32_t* __restrict__ offset,
                  const int32_t* __restrict__ id, int32_t* __restrict__ list) {
  const int32_t p = xcd_tile(blockIdx.x, tiles) * (int32_t)blockDim.x + (int32_t)threadIdx.x;
  if (p >= n) return;
  const int32_t c = cell_of[p];
  const uint32_t slot = atomicSub(count + c, 1u) - 1u;
  list[offset[c] + slot] = id[p];
}

// ---------------------------------------------------------------------------------------------
// K3: exclusive scan of the per-cell counts in one pass (decoupled look-back). Each workgroup
// scans one tile of kScanTile cells as four sub-tiles of 1024 (coalesced uint4 loads and stores,
// a thread's four cells summed, wave64 inclusive scans, LDS wave totals of all four sub-tiles
// behind one barrier). It publishes the tile total, looks back over the
// published totals of the tiles before it (64 at a time, one per lane of wave 0) until it meets a
// published inclusive prefix, publishes its own inclusive prefix, writes the offsets and clears
// the counts for the next step.
//
// Forward progress: workgroup b scans tile b and waits only for tiles below b. Workgroups are
// dealt to the XCDs round-robin and every XCD dispatches its share in increasing index order, so
// the lowest unfinished tile has always been dispatched (every workgroup resident on its XCD has a
// lower index and has finished, or it would not be the lowest) and waits for finished tiles only;
// tile 0 waits for nobody. Status words carry an epoch: each tile's word holds the epoch of the
// last launch that scanned it, and a launch uses that epoch + 1 (every launch scans every tile, so
// all words agree between launches). The words left by the previous launch are never mistaken for
// current ones and nothing has to be cleared between launches. Every wait is bounded: past
// kScanSpinLimit polls the predecessor is taken as a prefix of 0, the stall flag is raised for the
// host, and all offsets stay at or below their true values (so every index derived from them stays
// inside its array).
// ---------------------------------------------------------------------------------------------
constexpr uint64_t kTileAggregate = 1, kTilePrefix = 2;
constexpr uint32_t kEpochMask = (1u << 30) - 1;
constexpr int kScanSpinLimit = 1 << 16;

__device__ inline uint64_t tile_word(uint32_t value, uint64_t status, uint32_t epoch) {
  return (uint64_t)value | (status << 32) | ((uint64_t)epoch << 34);
}

__global__ void __launch_bounds__(kScanThreads)
    k_scan(uint32_t* __restrict__ count, uint32_t* __restrict__ offset,
           uint64_t* __restrict__ tile_state, int32_t ntiles, uint32_t* __restrict__ ovf_ctr,
           uint32_t* __restrict__ flags, int32_t clear, uint32_t limit) {
  // A tile is kScanRounds sub-tiles of kScanThreads uint4 (4 cells each): thread t holds uint4
  // t of every sub-tile, so every load and store of a wave is contiguous.
  constexpr int kScanRounds = kScanPerThread / 4;
  constexpr int kWaves = kScanThreads / 64;
  __shared__ uint32_t wsum[kScanRounds][kWaves];
  __shared__ uint32_t s_prefix;
  const int32_t tile = (int32_t)blockIdx.x;
  // this tile's word from the previous launch: its epoch + 1 is this launch's (read alongside the
  // counts; only this workgroup writes the word)
  const uint32_t epoch =
      ((uint32_t)(__hip_atomic_load(tile_state + tile, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >>
                  34) + 1u) & kEpochMask;
  if (tile == 0 && threadIdx.x == 0) {
    // the overflow list of the cell id blocks is complete (written by the previous kernel):
    // publish its length for the rank/gather kernel and restart the counter for the next pass
    ovf_ctr[1] = ovf_ctr[0];
    ovf_ctr[0] = 0u;
    ovf_ctr[2] += 1u;  // the generation of this bucketing (incremental bucketing)
  }
  uint4* __restrict__ src = reinterpret_cast<uint4*>(count + (size_t)tile * kScanTile);
  // Only the counts stay in registers across the look-back; the wave scans of the rounds are
  // redone for the output (a few instructions each) rather than held (two more registers per
  // round, fewer resident waves).
  uint4 u[kScanRounds];
  const int w = threadIdx.x >> 6;
  const int lane = lane_id();
#pragma unroll
  for (int q = 0; q < kScanRounds; ++q) u[q] = src[q * kScanThreads + threadIdx.x];
#pragma unroll
  for (int q = 0; q < kScanRounds; ++q) {
    const uint32_t wi = wave_inclusive_scan((u[q].x + u[q].y) + (u[q].z + u[q].w));
    if (lane == 63) wsum[q][w] = wi;
  }
  __syncthreads();
  uint32_t agg = 0;
#pragma unroll
  for (int q = 0; q < kScanRounds; ++q)
#pragma unroll
    for (int k = 0; k < kWaves; ++k) agg += wsum[q][k];

  if (w == 0) {
    if (lane == 0)
      __hip_atomic_store(tile_state + tile,
                         tile_word(agg, tile == 0 ? kTilePrefix : kTileAggregate, epoch),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t prefix = 0;
    int32_t end = tile;  // look back over tiles [end - 64, end)
    bool done = (tile == 0);
    while (!done) {
      const int32_t j = end - 1 - lane;
      uint64_t word = tile_word(0, kTilePrefix, epoch);  // lanes before tile 0 end the walk
      if (j >= 0) {
        for (int spins = 0;; ++spins) {
          word = __hip_atomic_load(tile_state + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if ((uint32_t)(word >> 34) == epoch && ((word >> 32) & 3u) != 0) break;
          if (spins >= kScanSpinLimit) {
            atomicOr(flags, kFlagScanStall);
            word = tile_word(0, kTilePrefix, epoch);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      const uint64_t pm = __ballot(((word >> 32) & 3u) == kTilePrefix);
      const int stop = pm ? __ffsll((unsigned long long)pm) - 1 : 63;  // nearest prefix
      uint32_t val = (lane <= stop) ? (uint32_t)word : 0u;
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) val += __shfl_xor(val, d, 64);
      prefix += val;
      done = pm != 0;
      end -= 64;
    }
    if (lane == 0) {
      if (tile > 0)
        __hip_atomic_store(tile_state + tile, tile_word(prefix + agg, kTilePrefix, epoch),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_prefix = prefix;
    }
  }
  __syncthreads();
  uint4* __restrict__ dst = reinterpret_cast<uint4*>(offset + (size_t)tile * kScanTile);
  const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
  uint32_t round_base = s_prefix;  // cells of the sub-tiles before round q
#pragma unroll
  for (int q = 0; q < kScanRounds; ++q) {
    const uint32_t tot = (u[q].x + u[q].y) + (u[q].z + u[q].w);
    uint32_t run = round_base + wave_inclusive_scan(tot) - tot;
    uint32_t rtot = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
      run += (k < w) ? wsum[q][k] : 0u;
      rtot += wsum[q][k];
    }
    uint4 o;
    o.x = min(run, limit);
    run += u[q].x;
    o.y = min(run, limit);
    run += u[q].y;
    o.z = min(run, limit);
    run += u[q].z;
    o.w = min(run, limit);
    dst[q * kScanThreads + threadIdx.x] = o;
    if (clear && (u[q].x | u[q].y | u[q].z | u[q].w) != 0u) src[q * kScanThreads + threadIdx.x] = zero;
    round_base += rtot;
  }
  if (tile == ntiles - 1 && threadIdx.x == 0) {
    // total: the offset one past the last cell
    offset[(size_t)ntiles * kScanTile] = min(round_base, limit);
    if (round_base != limit) atomicOr(flags, kFlagCountDrift);
  }
}

// The kept tile totals after the incremental bucketing's first pass (TileTotals): each tile's
// counts summed into the current generation's entry, one workgroup per tile of kRounds sub-tiles of
// kScanThreads uint4 (kScanTile cells for 16 rounds, kDirectTile for one).
template <int kRounds>
__global__ void __launch_bounds__(kScanThreads)
    k_tile_totals(const uint32_t* __restrict__ count, const TileTotals tt) {
  constexpr int kScanRounds = kRounds;
  constexpr int kWaves = kScanThreads / 64;
  constexpr size_t kTile = (size_t)kScanThreads * 4 * kRounds;
  __shared__ uint32_t wsum[kWaves];
  const int32_t tile = (int32_t)blockIdx.x;
  const uint4* __restrict__ src = reinterpret_cast<const uint4*>(count + (size_t)tile * kTile);
  uint4 u[kScanRounds];
#pragma unroll
  for (int q = 0; q < kScanRounds; ++q) u[q] = src[q * kScanThreads + threadIdx.x];
  uint32_t sum = 0;
#pragma unroll
  for (int q = 0; q < kScanRounds; ++q) sum += (u[q].x + u[q].y) + (u[q].z + u[q].w);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
  if (lane_id() == 0) wsum[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) total += wsum[k];
    tt.tot[(size_t)(*tt.gen & 1u) * (size_t)tt.ntiles + (size_t)tile] = total;
  }
}

// The incremental bucketing's scan (TileTotals): tile b's prefix is the sum of the kept totals of
// tiles [0, b) (the generation's entries, read in parallel by the workgroup's threads), so no tile
// waits for another. Each tile writes its own total, counted from its cells, into the other
// generation's entry for the next bucketing and raises kFlagCountDrift where the kept total
// differs from it. Offsets are clamped to `limit` as in k_scan; the counts are left in place.
// Tiles of kRounds sub-tiles as in k_tile_totals: kDirectTile cells (one round) on grids of at most
// kDirectTiles of them, 4,096 (four rounds) on grids of at most 4,096 of those, kScanTile cells (16)
// above: more workgroups where a workgroup's sum of the totals before its tile stays short.
template <int kRounds>
__global__ void __launch_bounds__(kScanThreads)
    k_scan_tiles(const uint32_t* __restrict__ count, uint32_t* __restrict__ offset,
                 const TileTotals tt, int32_t ntiles, uint32_t* __restrict__ ovf_ctr,
                 uint32_t* __restrict__ flags, uint32_t limit) {
  constexpr int kScanRounds = kRounds;
  constexpr int kWaves = kScanThreads / 64;
  constexpr size_t kTile = (size_t)kScanThreads * 4 * kRounds;
  __shared__ uint32_t wsum[kScanRounds][kWaves];
  __shared__ uint32_t wpre[kWaves];
  const int32_t tile = (int32_t)blockIdx.x;
  const uint32_t par = *tt.gen & 1u;
  const uint32_t* __restrict__ tin = tt.tot + (size_t)par * (size_t)ntiles;
  uint32_t* __restrict__ tout = tt.tot + (size_t)(par ^ 1u) * (size_t)ntiles;
  if (tile == 0 && threadIdx.x == 0) {
    ovf_ctr[1] = ovf_ctr[0];
    ovf_ctr[0] = 0u;
    ovf_ctr[2] += 1u;
  }
  const uint4* __restrict__ src = reinterpret_cast<const uint4*>(count + (size_t)tile * kTile);
  uint4 u[kScanRounds];
  const int w = threadIdx.x >> 6;
  const int lane = lane_id();
#pragma unroll
  for (int q = 0; q < kScanRounds; ++q) u[q] = src[q * kScanThreads + threadIdx.x];
  uint32_t pre = 0;
  for (int32_t k = (int32_t)threadIdx.x; k < tile; k += kScanThreads) pre += tin[k];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) pre += __shfl_xor(pre, d, 64);
  if (lane == 0) wpre[w] = pre;
#pragma unroll
  for (int q = 0; q < kScanRounds; ++q) {
    const uint32_t wi = wave_inclusive_scan((u[q].x + u[q].y) + (u[q].z + u[q].w));
    if (lane == 63) wsum[q][w] = wi;
  }
  __syncthreads();
  uint32_t prefix = 0;
#pragma unroll
  for (int k = 0; k < kWaves; ++k) prefix += wpre[k];
  uint4* __restrict__ dst = reinterpret_cast<uint4*>(offset + (size_t)tile * kTile);
  uint32_t round_base = prefix;
#pragma unroll
  for (int q = 0; q < kScanRounds; ++q) {
    const uint32_t tot = (u[q].x + u[q].y) + (u[q].z + u[q].w);
    uint32_t run = round_base + wave_inclusive_scan(tot) - tot;
    uint32_t rtot = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
      run += (k < w) ? wsum[q][k] : 0u;
      rtot += wsum[q][k];
    }
    uint4 o;
    o.x = min(run, limit);
    run += u[q].x;
    o.y = min(run, limit);
    run += u[q].y;
    o.z = min(run, limit);
    run += u[q].z;
    o.w = min(run, limit);
    dst[q * kScanThreads + threadIdx.x] = o;
    round_base += rtot;
  }
  if (threadIdx.x == 0) {
    const uint32_t agg = round_base - prefix;
    tout[tile] = agg;
    if (agg != tin[tile]) atomicOr(flags, kFlagCountDrift);
    if (tile == ntiles - 1) {
      offset[(size_t)ntiles * kTile] = min(round_base, limit);
      if (round_base != limit) atomicOr(flags, kFlagCountDrift);
    }
  }
}

// The direct scan (TileCounts): tile b's prefix is the sum of the per-tile counts of tiles
// [0, b) (read in parallel by the workgroup's 256 threads), so the tiles run independently.
constexpr int kDirectThreads = kDirectTile / 4;  // one uint4 of counts per thread

__global__ void __launch_bounds__(kDirectThreads)
    k_scan_direct(uint32_t* __restrict__ count, uint32_t* __restrict__ offset,
                  const TileCounts tiles, uint32_t* __restrict__ ovf_ctr, int32_t clear) {
  constexpr int kWaves = kDirectThreads / 64;
  __shared__ uint32_t wsum[kWaves], wpre[kWaves];
  const int32_t tile = (int32_t)blockIdx.x;
  const int w = threadIdx.x >> 6;
  const int lane = lane_id();
  const uint32_t par = *tiles.gen & 1u;  // every tile reads the same generation
  const uint32_t* __restrict__ tc = tiles.tc + (size_t)par * (size_t)tiles.ntiles * kTileStride;
  if (tile == 0 && threadIdx.x == 0) {  // as in the look-back scan
    ovf_ctr[1] = ovf_ctr[0];
    ovf_ctr[0] = 0u;
    ovf_ctr[2] += 1u;  // the generation of this bucketing (incremental bucketing)
  }
  uint4* __restrict__ src = reinterpret_cast<uint4*>(count + (size_t)tile * kDirectTile);
  const uint4 u = src[threadIdx.x];
  // this thread's share of the earlier tiles' counts: its (at most four) loads are issued
  // together and summed after (a rolled loop waits for each load before the next)
  constexpr int kPerThread = kDirectTiles / kDirectThreads;
  uint32_t share[kPerThread];
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int32_t j = (int32_t)threadIdx.x + k * kDirectThreads;
    share[k] = j < tile ? tc[(size_t)j * kTileStride] : 0u;
  }
  uint32_t before = 0;
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) before += share[k];
  const uint32_t tot = (u.x + u.y) + (u.z + u.w);
  const uint32_t wi = wave_inclusive_scan(tot);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) before += __shfl_xor(before, d, 64);
  if (lane == 63) wsum[w] = wi;
  if (lane == 0) wpre[w] = before;
  __syncthreads();
  uint32_t run = wi - tot, total = 0;
#pragma unroll
  for (int k = 0; k < kWaves; ++k) {
    run += wpre[k] + ((k < w) ? wsum[k] : 0u);
    total += wsum[k];
  }
  uint4 o;
  o.x = run;
  run += u.x;
  o.y = run;
  run += u.y;
  o.z = run;
  run += u.z;
  o.w = run;
  reinterpret_cast<uint4*>(offset + (size_t)tile * kDirectTile)[threadIdx.x] = o;
  if (clear && (u.x | u.y | u.z | u.w) != 0u) src[threadIdx.x] = make_uint4(0u, 0u, 0u, 0u);
  if (threadIdx.x == 0) {
    // this tile's count of the other generation, for the hash pass that follows this bucketing
    tiles.tc[((size_t)(par ^ 1u) * (size_t)tiles.ntiles + tile) * kTileStride] = 0u;
    if (tile == tiles.ntiles - 1) {
      uint32_t pre = 0;
#pragma unroll
      for (int k = 0; k < kWaves; ++k) pre += wpre[k];
      offset[(size_t)tiles.ntiles * kDirectTile] = pre + total;  // one past the last cell
    }
  }
}

// ---------------------------------------------------------------------------------------------
// K4: stable rank + gather. The reference appends ids to a cell in ascending order
// (ParticleModel.cpp:55-61), so a particle's sorted slot is offset[cell] + the number of cell mates
// with a smaller id. The mates are read from the cell's first min(count, kCellIds) id slots
// (slot-major, none for a particle alone in its cell) and, for a cell with more than kCellIds
// members, from the cell's chain of overflow entries (count - kCellIds of them), so the cost of a
// particle is bounded by its own cell's occupancy.
// ---------------------------------------------------------------------------------------------
// The rank a sphere keeps in a cell nothing entered or left: the byte the last ranking stored for
// its slot (r_old), or its place in the cell's old range where that byte is saturated.
__device__ __forceinline__ uint32_t inc_kept_rank(const IncRank& inc, int32_t c, uint32_t p,
                                                  uint32_t r_old) {
  return r_old < 255u ? r_old : p - inc.off_old[c];
}

// The rank of particle me (old sorted slot p, bucket cell c) in the incremental bucketing (IncRank):
// the rank it keeps (inc_kept_rank) when the cell's mark (mk) or, where the mark matches, its arrival
// word is of an earlier generation (nothing entered or left the cell), else the members of the
// cell's old range that stay in c and the cell's arrivals (id slots, then the overflow chain),
// counted where their id is smaller. The range is clamped to the particles, so stale offsets cannot
// send the loop past the arrays.
__device__ __forceinline__ uint32_t inc_rank(const IncRank& inc, uint32_t gen, int32_t c, int32_t me,
                                             uint32_t p, uint32_t r_old, uint8_t mk, uint32_t n,
                                             const int32_t* __restrict__ cell_of,
                                             const int32_t* __restrict__ id,
                                             const int32_t* __restrict__ cell_ids,
                                             uint32_t id_stride, const int2* __restrict__ ovf) {
  if (mk != (uint8_t)gen) return inc_kept_rank(inc, c, p, r_old);
  const uint64_t w = inc.arr[c];
  if ((uint32_t)(w >> 32) != gen) return inc_kept_rank(inc, c, p, r_old);
  const uint32_t lo = inc.off_old[c];
  const uint32_t hi = min(inc.off_old[c + 1], n);
  uint32_t rank = 0;
  for (uint32_t q = min(lo, hi); q < hi; ++q)
    rank += (cell_of[q] == c && id[q] < me) ? 1u : 0u;
  const uint32_t na = (uint32_t)w;
#pragma unroll
  for (int k = 0; k < kCellIds; ++k)
    if ((uint32_t)k < na) rank += (cell_ids[(size_t)k * id_stride + (uint32_t)c] < me) ? 1u : 0u;
  if (na > (uint32_t)kCellIds) {  // the arrivals past the slots, newest first
    int32_t o = cell_ids[(size_t)kCellIds * id_stride + (uint32_t)c];
    for (uint32_t k = kCellIds; k < na; ++k) {
      const int2 q = ovf[o];
      rank += (q.x < me) ? 1u : 0u;
      o = q.y;
    }
  }
  return rank;
}

// The binary32 form is held to seven waves per SIMD (at most 72 VGPRs, no spills; 79 unbounded, six
// waves): its waves wait on memory most of their cycles, and one more resident wave took the hopper's
// rank kernel from 78.0 to 77.0 us (same box; held to 64 VGPRs, eight waves, it spills in the mesh
// path and runs slower).
template <class T, bool kTable, bool kBlocks>
__global__ void __launch_bounds__(kMaxBlockSort<T>, sizeof(T) == 4 ? 7 : 1)
    k_rank_gather(const Locator<T> loc, const int32_t* __restrict__ cell_of,
                  const uint32_t* __restrict__ offset, const int32_t* __restrict__ cell_ids,
                  uint32_t id_stride,
                  const int2* __restrict__ ovf, const uint32_t* __restrict__ ovf_ctr,
                  const int32_t* __restrict__ list, const Vec4<T>* __restrict__ pos,
                  const Vec4<T>* __restrict__ vel,
                  const int32_t* __restrict__ id, const TriEntry<T>* __restrict__ tent,
                  const int32_t* __restrict__ tcell_off, uint32_t* __restrict__ flags, int mode,
                  const uint8_t* __restrict__ cull, int32_t n, Vec4<T>* __restrict__ pos_o,
                  Vec4<T>* __restrict__ vel_o, int32_t* __restrict__ id_o,
                  int32_t* __restrict__ cell_o, uint32_t* __restrict__ home_o,
                  Vec4<T>* __restrict__ xpre_o, uint32_t* __restrict__ gen, const IncRank inc) {
  // the scan of this bucketing is complete: the next hash pass counts tiles for the next one
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(gen, 1u);
  const uint32_t inc_gen = inc.arr ? ovf_ctr[2] : 0u;  // generation of this bucketing
  with_slice<T, kTable>(loc, [&](const SceneDev<T>& s, int32_t p0, int32_t p1) {
    const int32_t p = p0 + (int32_t)threadIdx.x;
    const bool active = p < p1;  // inactive lanes stay for the wave-wide mesh path
    Vec4<T> x = {}, v = {};
    int32_t c = 0;
    uint32_t dst = 0;
    int32_t me = 0;
    int32_t k0 = 0, k1 = 0;  // triangle list of the particle's cell
    uint8_t cull_c = 0;      // flag byte of the bucket cell
    const bool mesh = (mode & kModeContainer) && s.mesh >= 0;
    if (active) {
      c = cell_of[p];
      me = id[p];
      // the slot's kept rank, loaded with the slot's other words (a cell's mark decides its use)
      const uint32_t r_old = (kBlocks && inc.arr) ? inc.rank_old[p] : 0u;
      x = pos[p];
      v = vel[p];
      // the cell's words, all issued once the cell is known and before anything waits for the
      // particle's state (the integration) or for the flag byte (the triangle-list branch)
      cull_c = cull[c];
      const uint32_t b = offset[c], e = offset[c + 1];
      const uint8_t mk = (kBlocks && inc.arr) ? inc.mark[c] : (uint8_t)0;
      if (mode & kModeIntegrate) {
        integrate_position(s, x, v);
        integrate_velocity(s, v);
      }
      if (mesh && (cull_c & kCellTris)) {  // most cells list no triangle: their bounds go unread
        const int32_t tc = s.tcell_base + (c - s.grid.cell_base);
        k0 = tcell_off[tc];
        k1 = tcell_off[tc + 1];
      }
      uint32_t rank = 0;
      if constexpr (kBlocks) {
        if (inc.arr) {
          rank = inc_rank(inc, inc_gen, c, me, (uint32_t)p, r_old, mk, (uint32_t)n, cell_of, id,
                          cell_ids, id_stride, ovf);
        } else {
        const uint32_t cnt = e - b;
        if (cnt > 1u) {  // a particle alone in its cell has rank 0
#pragma unroll
          for (int k = 0; k < kCellIds; ++k)
            if ((uint32_t)k < cnt) rank += (cell_ids[(size_t)k * id_stride + (uint32_t)c] < me) ? 1u : 0u;
        }
        if (cnt > (uint32_t)kCellIds) {  // the cell's own overflow members, newest first
          int32_t o = cell_ids[(size_t)kCellIds * id_stride + (uint32_t)c];
          for (uint32_t k = kCellIds; k < cnt; ++k) {
            const int2 q = ovf[o];
            rank += (q.x < me) ? 1u : 0u;
            o = q.y;
          }
        }
        }
      } else {  // the first four members in one go (the list carries kListPad spare entries)
        const int32_t l0 = list[b], l1 = list[b + 1], l2 = list[b + 2], l3 = list[b + 3];
        rank = ((b < e && l0 < me) ? 1u : 0u) + ((b + 1 < e && l1 < me) ? 1u : 0u) +
               ((b + 2 < e && l2 < me) ? 1u : 0u) + ((b + 3 < e && l3 < me) ? 1u : 0u);
        for (uint32_t k = b + 4; k < e; ++k) rank += (list[k] < me) ? 1u : 0u;
      }
      // b + rank < n whenever the offsets are the exact scan; the bound keeps the store inside the
      // arrays even if a stalled scan (reported through the flags) left them short
      dst = min(b + rank, (uint32_t)(n - 1));
      if constexpr (kBlocks) {
        if (inc.arr) inc.rank_o[dst] = (uint8_t)min(rank, 255u);
      }
    }
    // The container acts on the pre-container state the bucketing hashed; its triangle-list loads
    // run alongside the rank's offset and id-block loads.
    const Vec4<T> xpre = x;
    if (mesh) {
      mesh_contact_wave(&s, active, k0, k1, tent, x, v);
    } else if (mode & kModeContainer) {
      if (active) box_walls(&s, x, v);
    }
    if (!active) return;
    // home cell of the post-container position for the pair kernel (the bucket cell unless the
    // container moved the particle), with the cell's cull flag in bit 31
    int32_t home = c;
    if (!(x.x == xpre.x && x.y == xpre.y && x.z == xpre.z)) {
      const GridDev& g = s.grid;
      const int32_t ix = axis_index((double)x.x, g.min[0], g.h, g.n[0]);
      const int32_t iy = axis_index((double)x.y, g.min[1], g.h, g.n[1]);
      const int32_t iz = axis_index((double)x.z, g.min[2], g.h, g.n[2]);
      home = g.cell_base + (ix * g.n[1] + iy) * g.n[2] + iz;
      cull_c = cull[home];
    }
    const bool moved = inc.cell_cur_o && home != c;  // incremental bucketing: see kHomeMoved
    home_o[dst] = (uint32_t)home | ((cull_c & kCellCull) ? kHomeCull : 0u) | (moved ? kHomeMoved : 0u);
    if (moved) inc.cell_cur_o[dst] = c;
    if (xpre_o) xpre_o[dst] = xpre;
    if (mode & kModeCheck) {
      const bool ok = isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(v.x) &&
                      isfinite(v.y) && isfinite(v.z);
      if (!ok) atomicOr(flags, kFlagNonFinite);
    }
    pos_o[dst] = x;
    vel_o[dst] = v;
    id_o[dst] = me;
    if (cell_o) cell_o[dst] = c;
  });
}

// ---------------------------------------------------------------------------------------------
// K5: pair contact (ParticleModel::resolveCollisions / reboundWithOtherParticle,
// ParticleModel.cpp:64-144). One lane per particle in sorted order; the 27 neighbour cells are
// walked x-outer, y, z-inner. Cells are numbered z-fastest, so the three z-neighbours of an (x, y)
// column are one contiguous slice [offset[c(zlo)], offset[c(zhi)+1]) of the sorted arrays, and
// walking it visits cells in order and members by ascending id -- the reference's summation order.
// ---------------------------------------------------------------------------------------------

// Per-lane constants of particle i.
template <class T>
struct PairSelf {
  Vec4<T> x, v;
  T thr;  // (r_i + r_max)^2 (1 + margin): no candidate beyond it can be in contact
  T ks, es;
  int32_t cuda;  // the scene's semantics flag (uniform): the CUDA backend's kick
};

// The prefilter of candidate j (fused multiply-adds, conservative by the margin): false only for
// candidates the exact test would reject. NaN coordinates pass it and fail the exact test.
template <class T, class P>
__device__ __forceinline__ bool pair_near(const PairSelf<T>& me, const P& xj) {
  const T ex = xj.x - me.x.x, ey = xj.y - me.x.y, ez = xj.z - me.x.z;
  return !(fma(ex, ex, fma(ey, ey, ez * ez)) > me.thr);
}

// The contact geometry of candidate j with particle i: separation, squared distance, radius sum
// and distance.
template <class T>
struct PairGeo {
  T ex, ey, ez, dd, R, len;
};

// The exact contact test of candidate j on particle i (isHit: distance(x_i, x_j) = length(x_j -
// x_i) < r_i + r_j); false when they do not touch. The squared-distance pretest is conservative by
// the margin, so only non-contacts skip the square root.
template <class T, class P>
__device__ __forceinline__ bool pair_touch(const PairSelf<T>& me, const P& xj, T rj,
                                           PairGeo<T>& g) {
  g.ex = xj.x - me.x.x;
  g.ey = xj.y - me.x.y;
  g.ez = xj.z - me.x.z;
  g.dd = dot3(g.ex, g.ey, g.ez, g.ex, g.ey, g.ez);
  g.R = me.x.w + rj;
  if (g.dd > (g.R * g.R) * PairMargin<T>::value) return false;
  g.len = sqrt(g.dd);
  return g.len < g.R;
}

// The kick of a touching candidate j (velocity vj) on particle i, in the reference's expression
// order.
template <class T>
__device__ __forceinline__ void pair_force(const PairSelf<T>& me, const PairGeo<T>& g,
                                           const Vec4<T>& vj, T& fx, T& fy, T& fz) {
  // d = x_i - x_j; dot(d, d) and length(d) equal dd and len bit for bit.
  const T dx = -g.ex, dy = -g.ey, dz = -g.ez;
  const T ov = g.R - g.len;
  if (me.cuda) {
    // resolveCollisionsKernel (ParticleModel_kernels.cu:163-199): the same contact test
    // (overlap = r_i + r_j - sqrt(d.d) > 0 holds exactly when sqrt(d.d) < r_i + r_j), then
    // (k ov / (|d| m_i) - 2 m_j (v_i - v_j).d e / ((m_i + m_j) d.d)) d
    const T rvd = dot3(me.v.x - vj.x, me.v.y - vj.y, me.v.z - vj.z, dx, dy, dz);
    const T crel = (((T(2) * vj.w) * rvd) * me.es) / ((me.v.w + vj.w) * g.dd);
    const T cspr = (me.ks * ov) / (g.len * me.v.w);
    const T f = cspr - crel;
    fx = f * dx;
    fy = f * dy;
    fz = f * dz;
    return;
  }
  const T s3 = ((T(2) * vj.w) / (me.v.w + vj.w)) *
               (dot3(me.v.x - vj.x, me.v.y - vj.y, me.v.z - vj.z, dx, dy, dz) / g.dd);
  const T inv = T(1) / g.len;
  const T kov = me.ks * ov;
  fx = (-((s3 * dx) * me.es)) + kov * (dx * inv);
  fy = (-((s3 * dy) * me.es)) + kov * (dy * inv);
  fz = (-((s3 * dz) * me.es)) + kov * (dz * inv);
}

// One candidate j of particle i: its kick on i, or false when they do not touch.
template <class T>
__device__ __forceinline__ bool pair_kick(const PairSelf<T>& me, const Vec4<T>& xj,
                                          const Vec4<T>* __restrict__ vsn, uint32_t q, T& fx,
                                          T& fy, T& fz) {
  PairGeo<T> g;
  if (!(pair_near(me, xj) && pair_touch(me, xj, xj.w, g))) return false;
  pair_force(me, g, vsn[q], fx, fy, fz);
  return true;
}

// Particle i's constants and the nine column slices [lo[d], hi[d]) of its 27 neighbour cells.
//
// Cell culling. A cell is bounded when every member's snapshot position lies inside the cell's box:
// interior cells (not on the grid boundary, where positions outside the domain are clamped in)
// whose triangle list is empty (no container push can move a member: a pushed particle touched a
// triangle within reach of its bucket cell, and box walls only push particles of boundary cells).
// cull[c] marks cells whose 27 neighbours are all bounded; for such a home cell, neighbour cells
// whose box is farther from x_i than reach = r_i + r_max (enlarged by a margin for rounding) hold
// no contact of i and are dropped from the column slices. Only non-contacts are skipped, so the
// contacts and their order are those of the full walk.
template <class T>
__device__ __forceinline__ void pair_setup(const SceneDev<T>* __restrict__ s,
                                           const Vec4<T>* __restrict__ pos,
                                           const Vec4<T>* __restrict__ vsn,
                                           const uint32_t* __restrict__ offset,
                                           const uint32_t* __restrict__ home_of, int32_t p,
                                           PairSelf<T>& me, uint32_t (&lo)[9], uint32_t (&hi)[9]) {
  const GridDev& g = s->grid;
  me.x = pos[p];
  me.v = vsn[p];
  // home cell of the snapshot position (ParticleModel.cpp:106-110, clamped) and its cull flag,
  // from the gather
  const uint32_t hw = home_of[p];
  const int32_t home = (int32_t)(hw & kHomeCell);
  const T rr = me.x.w + s->rmax;
  me.thr = (rr * rr) * PairMargin<T>::value;
  me.ks = s->spring;
  me.es = s->es;
  me.cuda = s->cuda;
  int32_t hx, hy, hz;
  cell_coords(g, home, hx, hy, hz);
  if (hw & kHomeCull) {  // all 27 neighbours exist and are bounded
    // distances (binary64) from x_i to the lower and upper neighbour along each axis
    const double ux = ((double)me.x.x - g.min[0]) - (double)hx * g.h;
    const double uy = ((double)me.x.y - g.min[1]) - (double)hy * g.h;
    const double uz = ((double)me.x.z - g.min[2]) - (double)hz * g.h;
    const double dlo[3] = {fmax(ux, 0.0), fmax(uy, 0.0), fmax(uz, 0.0)};
    const double dhi[3] = {fmax(g.h - ux, 0.0), fmax(g.h - uy, 0.0), fmax(g.h - uz, 0.0)};
    const double reach = ((double)me.x.w + (double)s->rmax) * (1.0 + 1e-4) + 1e-4 * g.h;
    const double r2 = reach * reach;
    const double zm = dlo[2] * dlo[2], zp = dhi[2] * dhi[2];
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const int dx = d / 3 - 1, dy = d % 3 - 1;
      const double ex = dx < 0 ? dlo[0] : (dx > 0 ? dhi[0] : 0.0);
      const double ey = dy < 0 ? dlo[1] : (dy > 0 ? dhi[1] : 0.0);
      const double exy = ex * ex + ey * ey;
      lo[d] = 0;
      hi[d] = 0;
      if (exy <= r2) {  // the column's middle cell is within reach; its ends may be
        const uint32_t* __restrict__ col =
            offset + home + (dx * g.n[1] + dy) * g.n[2];  // cell (hx+dx, hy+dy, hz)
        lo[d] = col[exy + zm <= r2 ? -1 : 0];
        hi[d] = col[exy + zp <= r2 ? 2 : 1];
      }
    }
    return;
  }
  const int32_t zlo = max(hz - 1, 0), zhi = min(hz + 1, g.n[2] - 1);
#pragma unroll
  for (int d = 0; d < 9; ++d) {
    const int32_t bx = hx - 1 + d / 3, by = hy - 1 + d % 3;
    lo[d] = 0;
    hi[d] = 0;
    if (bx >= 0 && bx < g.n[0] && by >= 0 && by < g.n[1]) {
      const int32_t cb = g.cell_base + (bx * g.n[1] + by) * g.n[2];
      lo[d] = offset[cb + zlo];
      hi[d] = offset[cb + zhi + 1];
    }
  }
}

// A candidate that passed the prefilter: its contact test and, if they touch, its kick on
// particle i, added to the sum; the candidate's velocity is loaded only for a contact.
template <class T>
__device__ __forceinline__ void pair_contact(const PairSelf<T>& me, const Vec4<T>& xj,
                                             const Vec4<T>* __restrict__ vsn, uint32_t q, T& ax,
                                             T& ay, T& az) {
  PairGeo<T> g;
  if (pair_touch(me, xj, xj.w, g)) {
    T fx, fy, fz;
    pair_force(me, g, vsn[q], fx, fy, fz);
    ax = ax + fx;
    ay = ay + fy;
    az = az + fz;
  }
}

// Near candidates deferred to a per-lane queue (the one-lane kernel). A chunk whose prefilter
// passes in any lane runs the exact test (a correctly rounded square root and three divisions) and
// the kick for the whole wave; in contact-rich states most chunks of most waves do. The one-lane
// kernel instead files each near candidate's index in its lane's queue in LDS, in walk order, and
// runs the queue's exact tests in one loop after the walk: the wave runs as many exact tests as its
// lane with the most near candidates. The kicks depend on the snapshot only (Jacobi), and the
// additions keep the walk's order, so the results are bit-identical. The walk visits candidates in
// ascending sorted index (columns in ascending cell order, each slice ascending), so a lane whose
// queue fills (crowded cells) keeps the first index it could not file, and after the queue it tests
// the candidates from that index on in line: nothing is walked twice, and the order holds.
// (Emptying the queue inside the walk would put a copy of the exact tests at every filing site of
// the unrolled walk, and walking again would put the in-line walk's registers into the kernel.)
constexpr int kNearQ = 16;  // entries per lane

struct NearQueue {
  uint32_t* slot;  // this lane's entry 0 in LDS; entry k at slot[k * stride]
  uint32_t stride; // lanes of the workgroup
  uint32_t n;      // entries held
  uint32_t rest;   // the first near candidate not filed (~0u: none)
};

__device__ __forceinline__ void near_push(NearQueue& nq, uint32_t j) {
  if (nq.n < (uint32_t)kNearQ) {
    nq.slot[nq.n * nq.stride] = j;
    ++nq.n;
  } else if (nq.rest == ~0u) {
    nq.rest = j;
  }
}

// The exact test and kick of candidate j (position xj, velocity vj) on particle i, added to the sum.
template <class T>
__device__ __forceinline__ void near_add(const PairSelf<T>& me, const Vec4<T>& xj,
                                         const Vec4<T>& vj, T& ax, T& ay, T& az) {
  PairGeo<T> g;
  if (pair_touch(me, xj, xj.w, g)) {
    T fx, fy, fz;
    pair_force(me, g, vj, fx, fy, fz);
    ax = ax + fx;
    ay = ay + fy;
    az = az + fz;
  }
}

// The exact tests and kicks of the queued candidates in queue order, two entries per round with
// two register sets: the next entry's position and velocity are loaded before an entry is tested.
template <class T>
__device__ __forceinline__ void near_flush(const NearQueue& nq, const PairSelf<T>& me,
                                           const Vec4<T>* __restrict__ pos,
                                           const Vec4<T>* __restrict__ vsn, T& ax, T& ay, T& az) {
  if (nq.n == 0u) return;
  const uint32_t last = nq.n - 1u;
  uint32_t j = nq.slot[0];
  Vec4<T> xa = pos[j], va = vsn[j];
  for (uint32_t k = 0; k < nq.n; k += 2) {
    j = nq.slot[min(k + 1u, last) * nq.stride];
    const Vec4<T> xb = pos[j], vb = vsn[j];
    near_add(me, xa, va, ax, ay, az);
    j = nq.slot[min(k + 2u, last) * nq.stride];
    xa = pos[j];
    va = vsn[j];
    if (k + 1u < nq.n) near_add(me, xb, vb, ax, ay, az);
  }
}

// Chunks of four candidates: one address, four independent loads, four prefilters and a single
// branch into the rare path, which applies the contacts of the chunk in ascending order. The
// masks are combined with bitwise operators, so every prefilter runs and the compiler issues the
// four loads before the first test (one round trip per chunk; with short-circuit masks it sinks
// each load under the previous candidate's condition: four round trips). The rare path loads the
// four candidates' velocities in one go and waits for them at its end, so no load is in flight past
// it (its registers would otherwise have to be waited for before the next slice's prefetch could
// be issued). The loads may run up to three entries past the end of a slice (or read four entries
// of an empty slice at offset 0); the sorted position and velocity arrays carry kPairPad spare
// entries so that they stay inside the allocations, and those entries are masked out.
template <class T>
__device__ __forceinline__ void pair_load4(const Vec4<T>* __restrict__ src, Vec4<T> (&c)[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) c[k] = src[k];
}

// vmcnt(0) with the other counters left alone (expcnt 7, lgkmcnt 15): every vector memory load of
// the wave has returned.
constexpr int kWaitVmem = 0xF70;

// Tests the chunk c[0 .. 4) = candidates q .. q + 3 of a slice ending at e.
template <class T>
__device__ __forceinline__ void pair_test4(const PairSelf<T>& me, const Vec4<T> (&c)[4],
                                           const Vec4<T>* __restrict__ vsn, uint32_t q, uint32_t e,
                                           uint32_t self, T& ax, T& ay, T& az) {
  bool m[4];
  bool any = false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t j = q + (uint32_t)k;
    m[k] = (j < e) & (j != self) & pair_near(me, c[k]);
    any = any | m[k];
  }
  if (any) {
    Vec4<T> vj[4];
    pair_load4(vsn + q, vj);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      PairGeo<T> g;
      if (m[k] && pair_touch(me, c[k], c[k].w, g)) {
        T fx, fy, fz;
        pair_force(me, g, vj[k], fx, fy, fz);
        ax = ax + fx;
        ay = ay + fy;
        az = az + fz;
      }
    }
    __builtin_amdgcn_s_waitcnt(kWaitVmem);
  }
}

// The chunk of four candidates starting at q inside a slice, read from memory: only the
// candidates before the slice end e are loaded (the others are zero and masked out), so lanes past
// their slice's end fetch nothing, and the loads are still issued together before the first test.
template <class T>
__device__ __forceinline__ void pair_chunk4(const PairSelf<T>& me, const Vec4<T>* __restrict__ pos,
                                            const Vec4<T>* __restrict__ vsn, uint32_t q,
                                            uint32_t e, uint32_t self, T& ax, T& ay, T& az) {
  Vec4<T> c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    c[k] = Vec4<T>{T(0), T(0), T(0), T(0)};
    if (q + (uint32_t)k < e) c[k] = pos[q + k];
  }
  pair_test4(me, c, vsn, q, e, self, ax, ay, az);
}

// Slice [lo, hi) from index `from` on, the exact tests in line (a lane whose queue filled).
template <class T>
__device__ __forceinline__ void near_tail(const PairSelf<T>& me, const Vec4<T>* __restrict__ pos,
                                          const Vec4<T>* __restrict__ vsn, uint32_t lo, uint32_t hi,
                                          uint32_t from, uint32_t self, T& ax, T& ay, T& az) {
  for (uint32_t j = max(lo, from); j < hi; ++j) {
    const Vec4<T> xj = pos[j];
    if (j != self && pair_near(me, xj)) pair_contact(me, xj, vsn, j, ax, ay, az);
  }
}

// The chunk tests of the one-lane kernel: the near candidates go to the lane's queue.
template <class T>
__device__ __forceinline__ void pair_test4q(const PairSelf<T>& me, const Vec4<T> (&c)[4],
                                            uint32_t q, uint32_t e, uint32_t self, NearQueue& nq) {
  bool m[4];
  bool any = false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t j = q + (uint32_t)k;
    m[k] = (j < e) & pair_near(me, c[k]);
    any = any | m[k];
  }
  if (any) {  // the particle itself (distance 0) is left out here, as in pair_test2q
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (m[k] && q + (uint32_t)k != self) near_push(nq, q + (uint32_t)k);
  }
}

template <class T>
__device__ __forceinline__ void pair_chunk4q(const PairSelf<T>& me, const Vec4<T>* __restrict__ pos,
                                             uint32_t q, uint32_t e, uint32_t self, NearQueue& nq) {
  Vec4<T> c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    c[k] = Vec4<T>{T(0), T(0), T(0), T(0)};
    if (q + (uint32_t)k < e) c[k] = pos[q + k];
  }
  pair_test4q(me, c, q, e, self, nq);
}

// Chunks of two candidates (c0, c1 = candidates q, q + 1 of a slice ending at e), for slices that
// hold one or two spheres: half the prefilters of a four-wide chunk. The masks short-circuit, so
// a candidate past its slice's end costs its lane no load and a wave whose lanes are all past it
// skips the test; the compiler then loads c1 under c0's condition (two round trips for a chunk
// read inside a slice). Measured on MI355X, this beats the eager form of the four-wide chunks on
// the workloads that fill the GPU (the 4.2M-sphere hopper, the 256-scene batch), whose pair
// kernels are bound by instruction issue and load bytes rather than by latency.
template <class T>
__device__ __forceinline__ void pair_test2q(const PairSelf<T>& me, const Vec4<T>& c0,
                                            const Vec4<T>& c1, uint32_t q, uint32_t e,
                                            uint32_t self, NearQueue& nq) {
  // the particle itself passes the prefilter (distance 0) and is left out where candidates are filed,
  // so the rounds that find nothing near, most of them, skip that comparison
  const bool m0 = q < e && pair_near(me, c0);
  const bool m1 = q + 1 < e && pair_near(me, c1);
  if (m0 || m1) {
    if (m0 && q != self) near_push(nq, q);
    if (m1 && q + 1 != self) near_push(nq, q + 1);
  }
}

// The walk of particle p over its nine column slices, in slice order, W (4 or 2) candidates per
// round, its near candidates filed in the lane's queue. In binary32 the first chunk of the next slice is loaded while the current slice is
// tested, so the slices' loads overlap instead of forming a chain of nine round trips (binary64
// lacks the registers). Measured on MI355X: prefetching two or three slices ahead changes the
// 262k-, 4.2M- and 4.6M-sphere workloads by about 1 % either way (more registers, fewer waves per
// SIMD), so one is kept; two-wide chunks are 7-11 % faster on the 4.2M-sphere hopper and 5 %
// slower on the 262k-sphere cylinder, so the chunk width is a tuning choice (psx_tuning.pair_chunk).
template <class T, int W>
__device__ __forceinline__ void pair_sum(const PairSelf<T>& me, const Vec4<T>* __restrict__ pos,
                                         const uint32_t (&lo)[9], const uint32_t (&hi)[9],
                                         uint32_t self, NearQueue& nq) {
  static_assert(W == 2 || W == 4, "two or four candidates per round");
  if constexpr (sizeof(T) == 4) {
    Vec4<T> buf[W];
#pragma unroll
    for (int k = 0; k < W; ++k) buf[k] = pos[lo[0] + k];
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      Vec4<T> c[W];
#pragma unroll
      for (int k = 0; k < W; ++k) c[k] = buf[k];
      if (d + 1 < 9) {
        const Vec4<T>* __restrict__ nx = pos + lo[d + 1];
#pragma unroll
        for (int k = 0; k < W; ++k) buf[k] = nx[k];
      }
      const uint32_t e = hi[d];
      // unconditional (the masks cover an empty slice), so the prefetched chunk is used as loaded
      if constexpr (W == 4) {
        pair_test4q(me, c, lo[d], e, self, nq);
        for (uint32_t q = lo[d] + 4; q < e; q += 4) pair_chunk4q(me, pos, q, e, self, nq);
      } else {
        pair_test2q(me, c[0], c[1], lo[d], e, self, nq);
        for (uint32_t q = lo[d] + 2; q < e; q += 2) {
          const Vec4<T>* __restrict__ src = pos + q;
          pair_test2q(me, src[0], src[1], q, e, self, nq);
        }
      }
    }
  } else {
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const uint32_t e = hi[d];
      if constexpr (W == 4) {
        for (uint32_t q = lo[d]; q < e; q += 4) pair_chunk4q(me, pos, q, e, self, nq);
      } else {
        for (uint32_t q = lo[d]; q < e; q += 2) {
          const Vec4<T>* __restrict__ src = pos + q;
          pair_test2q(me, src[0], src[1], q, e, self, nq);
        }
      }
    }
  }
}

// The binary32 two-wide walk without the nine slices' bounds held in registers. The sphere's walk is
// summarised in a 27-bit code (bit d: column d is walked; 9 + d: its slice starts at cell z - 1,
// else at z; 18 + d: it ends at cell z + 1, else at z), from the same culling rule and the same
// clamping as pair_setup, so the slices are identical. Column d's bounds come from one 16-byte load
// of the offsets of its cells z - 1 .. z + 2 (the offset arrays carry a spare entry on each side),
// issued two slices ahead; each slice's first chunk is loaded one slice ahead as in pair_sum. Half
// the offset loads, and 18 registers fewer for the bounds: measured on MI355X, the 4.2M-sphere
// hopper's pair kernel 147.5 -> 140.2 us, bit-identical.
struct __attribute__((aligned(4))) Off4 {
  uint32_t a, b, c, d;
};

template <class T>
__device__ __forceinline__ uint32_t walk_code(const SceneDev<T>& s, const Vec4<T>& x, uint32_t hw,
                                              int32_t hx, int32_t hy, int32_t hz) {
  const GridDev& g = s.grid;
  uint32_t code = 0;
  if (hw & kHomeCull) {  // as in pair_setup: all 27 neighbours exist and are bounded
    const double ux = ((double)x.x - g.min[0]) - (double)hx * g.h;
    const double uy = ((double)x.y - g.min[1]) - (double)hy * g.h;
    const double uz = ((double)x.z - g.min[2]) - (double)hz * g.h;
    const double dlo[3] = {fmax(ux, 0.0), fmax(uy, 0.0), fmax(uz, 0.0)};
    const double dhi[3] = {fmax(g.h - ux, 0.0), fmax(g.h - uy, 0.0), fmax(g.h - uz, 0.0)};
    const double reach = ((double)x.w + (double)s.rmax) * (1.0 + 1e-4) + 1e-4 * g.h;
    const double r2 = reach * reach;
    const double zm = dlo[2] * dlo[2], zp = dhi[2] * dhi[2];
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const int dx = d / 3 - 1, dy = d % 3 - 1;
      const double ex = dx < 0 ? dlo[0] : (dx > 0 ? dhi[0] : 0.0);
      const double ey = dy < 0 ? dlo[1] : (dy > 0 ? dhi[1] : 0.0);
      const double exy = ex * ex + ey * ey;
      if (exy <= r2) {
        code |= 1u << d;
        if (exy + zm <= r2) code |= 1u << (9 + d);
        if (exy + zp <= r2) code |= 1u << (18 + d);
      }
    }
    return code;
  }
#pragma unroll
  for (int d = 0; d < 9; ++d) {  // the grid's clamps
    const int32_t bx = hx - 1 + d / 3, by = hy - 1 + d % 3;
    if (bx >= 0 && bx < g.n[0] && by >= 0 && by < g.n[1]) {
      code |= 1u << d;
      if (hz > 0) code |= 1u << (9 + d);
      if (hz < g.n[2] - 1) code |= 1u << (18 + d);
    }
  }
  return code;
}

// Offsets of cells z - 1 .. z + 2 of column d around home cell `home`; the lanes of a column not
// walked all read the first entries of the array (one cache line; discarded by walk_bounds). The
// load is unconditional: a load under a branch is merged with its default afterwards, and the
// compiler waited for all loads in flight (the prefetched chunk among them) at that merge.
template <int d>
__device__ __forceinline__ Off4 walk_offsets(const uint32_t* __restrict__ offset, int32_t home,
                                             int32_t ny, int32_t nz, uint32_t code) {
  const int32_t c = (code & (1u << d)) ? home + ((d / 3 - 1) * ny + (d % 3 - 1)) * nz : 1;
  return *reinterpret_cast<const Off4*>(offset + c - 1);
}

template <int d>
__device__ __forceinline__ void walk_bounds(const Off4& o, uint32_t code, uint32_t& lo,
                                            uint32_t& hi) {
  lo = (code & (1u << (9 + d))) ? o.a : o.b;
  hi = (code & (1u << (18 + d))) ? o.d : o.c;
  if (!(code & (1u << d))) lo = hi = 0u;
}

// Column d's slice [lo, hi) for the walk code (d a constant after unrolling).
__device__ __forceinline__ void code_bounds(const uint32_t* __restrict__ offset, int32_t home,
                                            int32_t ny, int32_t nz, uint32_t code, int d,
                                            uint32_t& lo, uint32_t& hi) {
  lo = hi = 0u;
  if (code & (1u << d)) {
    const Off4 o =
        *reinterpret_cast<const Off4*>(offset + home + ((d / 3 - 1) * ny + (d % 3 - 1)) * nz - 1);
    lo = (code & (1u << (9 + d))) ? o.a : o.b;
    hi = (code & (1u << (18 + d))) ? o.d : o.c;
  }
}

// Slice d: its first chunk (buf, loaded one slice earlier), then the rest; meanwhile slice d + 1's
// bounds are decoded and its first chunk loaded, and slice d + 2's offsets fetched.
template <int d, class T>
__device__ __forceinline__ void walk_slice(const PairSelf<T>& me, const Vec4<T>* __restrict__ pos,
                                           const uint32_t* __restrict__ offset, int32_t home,
                                           int32_t ny, int32_t nz, uint32_t code, uint32_t self,
                                           uint32_t& lo, uint32_t& hi, Vec4<T> (&buf)[2], Off4& o2,
                                           NearQueue& nq) {
  const Vec4<T> c0 = buf[0], c1 = buf[1];
  const uint32_t l = lo, e = hi;
  if constexpr (d + 1 < 9) {
    walk_bounds<d + 1>(o2, code, lo, hi);
    buf[0] = pos[lo];
    buf[1] = pos[lo + 1];
  }
  if constexpr (d + 2 < 9) o2 = walk_offsets<d + 2>(offset, home, ny, nz, code);
  pair_test2q(me, c0, c1, l, e, self, nq);
  for (uint32_t q = l + 2; q < e; q += 2) {
    const Vec4<T>* __restrict__ src = pos + q;
    pair_test2q(me, src[0], src[1], q, e, self, nq);
  }
}

template <class T>
__device__ __forceinline__ void pair_sum_coded(const SceneDev<T>& sc, const PairSelf<T>& me,
                                               const Vec4<T>* __restrict__ pos,
                                               const uint32_t* __restrict__ offset, int32_t home,
                                               uint32_t code, uint32_t self, NearQueue& nq) {
  const int32_t ny = sc.grid.n[1], nz = sc.grid.n[2];
  uint32_t lo, hi;
  Off4 o2 = walk_offsets<0>(offset, home, ny, nz, code);
  walk_bounds<0>(o2, code, lo, hi);
  o2 = walk_offsets<1>(offset, home, ny, nz, code);
  Vec4<T> buf[2] = {pos[lo], pos[lo + 1]};
  walk_slice<0>(me, pos, offset, home, ny, nz, code, self, lo, hi, buf, o2, nq);
  walk_slice<1>(me, pos, offset, home, ny, nz, code, self, lo, hi, buf, o2, nq);
  walk_slice<2>(me, pos, offset, home, ny, nz, code, self, lo, hi, buf, o2, nq);
  walk_slice<3>(me, pos, offset, home, ny, nz, code, self, lo, hi, buf, o2, nq);
  walk_slice<4>(me, pos, offset, home, ny, nz, code, self, lo, hi, buf, o2, nq);
  walk_slice<5>(me, pos, offset, home, ny, nz, code, self, lo, hi, buf, o2, nq);
  walk_slice<6>(me, pos, offset, home, ny, nz, code, self, lo, hi, buf, o2, nq);
  walk_slice<7>(me, pos, offset, home, ny, nz, code, self, lo, hi, buf, o2, nq);
  walk_slice<8>(me, pos, offset, home, ny, nz, code, self, lo, hi, buf, o2, nq);
}

// The binary32 two-wide form is held to eight waves per SIMD (at most 64 VGPRs; it takes 53 with
// the near-candidate queue): it runs on the workloads that fill the GPU many times over, where
// more resident waves hide more of each wave's chain of slice loads (1 % on the 4.2M-sphere hopper
// and the 256-scene batch, measured on MI355X, when it needed 67 unbounded).
template <class T, bool kTable, int W>
__global__ void __launch_bounds__(1024, (W == 2 && sizeof(T) == 4) ? 8 : 1)
    k_pairs(const Locator<T> loc, const Vec4<T>* __restrict__ pos,
            const Vec4<T>* __restrict__ vsn, const uint32_t* __restrict__ offset,
            const uint32_t* __restrict__ home, Vec4<T>* __restrict__ vout, const NextBucket next) {
  extern __shared__ uint32_t near_raw[];  // kNearQ entries per lane
  __shared__ WgTiles wt;
  with_slice<T, kTable>(loc, [&](const SceneDev<T>& sc, int32_t p0, int32_t p1) {
    const int32_t p = p0 + (int32_t)threadIdx.x;
    const bool tiles = next.count && next.tiles.tc;
    if (tiles) wt.begin((int32_t)(home[p0] & kHomeCell));
    int32_t c = -1;
    if (p < p1) {
      PairSelf<T> me;
      NearQueue nq{near_raw + threadIdx.x, blockDim.x, 0u, ~0u};
      T ax, ay, az;
      uint32_t hw_next = 0;  // the home word for the incremental hash pass
      if constexpr (W == 2 && sizeof(T) == 4) {
        me.x = pos[p];
        me.v = vsn[p];
        const uint32_t hw = home[p];
        hw_next = hw;
        const int32_t hc = (int32_t)(hw & kHomeCell);
        const T rr = me.x.w + sc.rmax;
        me.thr = (rr * rr) * PairMargin<T>::value;
        me.ks = sc.spring;
        me.es = sc.es;
        me.cuda = sc.cuda;
        int32_t hx, hy, hz;
        cell_coords(sc.grid, hc, hx, hy, hz);
        const uint32_t code = walk_code(sc, me.x, hw, hx, hy, hz);
        ax = me.v.x;
        ay = me.v.y;
        az = me.v.z;
        pair_sum_coded(sc, me, pos, offset, hc, code, (uint32_t)p, nq);
        near_flush(nq, me, pos, vsn, ax, ay, az);
        if (nq.rest != ~0u) {  // the queue filled: the candidates from its first one left out on
#pragma unroll
          for (int d = 0; d < 9; ++d) {
            uint32_t lo, hi;
            code_bounds(offset, hc, sc.grid.n[1], sc.grid.n[2], code, d, lo, hi);
            near_tail(me, pos, vsn, lo, hi, nq.rest, (uint32_t)p, ax, ay, az);
          }
        }
      } else {
        uint32_t lo[9], hi[9];
        pair_setup(&sc, pos, vsn, offset, home, p, me, lo, hi);
        if (next.arr) hw_next = home[p];
        ax = me.v.x;
        ay = me.v.y;
        az = me.v.z;
        pair_sum<T, W>(me, pos, lo, hi, (uint32_t)p, nq);
        near_flush(nq, me, pos, vsn, ax, ay, az);
        if (nq.rest != ~0u) {  // the queue filled: the candidates from its first one left out on
#pragma unroll
          for (int d = 0; d < 9; ++d)
            near_tail(me, pos, vsn, lo[d], hi[d], nq.rest, (uint32_t)p, ax, ay, az);
        }
      }
      const Vec4<T> vn{ax, ay, az, me.v.w};
      vout[p] = vn;
      if (next.count) c = next_bucket(sc, next, me.x, vn, p, hw_next);
    }
    if (tiles) {
      wt.add(next.tiles, c);
      wt.end(next.tiles);
    }
  });
}

// The same contact sum with G lanes per particle (G divides 64, so a group never straddles a
// wave). Each column slice is walked in rounds of G consecutive candidates, lane g of the group
// testing candidate (round start + g); the kicks found in a round are then added by every lane of
// the group in ascending candidate order, fetched from their lanes with shuffles. Every lane of a
// group thus performs the single-lane kernel's additions in the same order and ends with the same
// bit pattern; lane 0 of the group writes it. More lanes per particle shorten the serial chain of
// dependent loads per lane (about 81 candidates / G rounds), which is what bounds the kernel when
// the particles alone are too few to fill the GPU.
template <class T, int G, bool kTable>
__global__ void __launch_bounds__(1024)
    k_pairs_lanes(const Locator<T> loc, const Vec4<T>* __restrict__ pos,
                  const Vec4<T>* __restrict__ vsn, const uint32_t* __restrict__ offset,
                  const uint32_t* __restrict__ home, Vec4<T>* __restrict__ vout,
                  const NextBucket next) {
  static_assert(G >= 2 && G <= 16 && (64 % G) == 0, "lanes per particle: 2, 4, 8 or 16");
  __shared__ WgTiles wt;
  with_slice<T, kTable>(loc, [&](const SceneDev<T>& sc, int32_t p0, int32_t p1) {
    const int32_t g = (int32_t)threadIdx.x % G;
    const int32_t p = p0 + (int32_t)threadIdx.x / G;
    const bool tiles = next.count && next.tiles.tc;
    if (tiles) wt.begin((int32_t)(home[p0] & kHomeCell));
    int32_t c = -1;
    if (p < p1) {  // whole groups take part or not
      PairSelf<T> me;
      uint32_t lo[9], hi[9];
      pair_setup(&sc, pos, vsn, offset, home, p, me, lo, hi);
      const int gbase = lane_id() & ~(G - 1);
      T ax = me.v.x, ay = me.v.y, az = me.v.z;
#pragma unroll
      for (int d = 0; d < 9; ++d) {
        const uint32_t e = hi[d];
        for (uint32_t q0 = lo[d]; q0 < e; q0 += G) {  // the same trip count on a whole group
          const uint32_t q = q0 + (uint32_t)g;
          T fx = T(0), fy = T(0), fz = T(0);
          bool hit = false;
          if (q < e && (int32_t)q != p) hit = pair_kick(me, pos[q], vsn, q, fx, fy, fz);
          uint32_t mine = (uint32_t)(__ballot(hit) >> gbase) & ((1u << G) - 1u);
          while (mine) {  // identical on every lane of the group
            const int src = gbase + __ffs(mine) - 1;
            mine &= mine - 1u;
            const T kx = __shfl(fx, src, 64), ky = __shfl(fy, src, 64), kz = __shfl(fz, src, 64);
            ax = ax + kx;
            ay = ay + ky;
            az = az + kz;
          }
        }
      }
      if (g == 0) {
        const Vec4<T> vn{ax, ay, az, me.v.w};
        vout[p] = vn;
        if (next.count) c = next_bucket(sc, next, me.x, vn, p, next.arr ? home[p] : 0u);
      }
    }
    if (tiles) {
      wt.add(next.tiles, c);
      wt.end(next.tiles);
    }
  });
}

// The one-lane pair kernel with the workgroup's neighbour particles staged in LDS. A workgroup's
// particles are consecutive in cell order, so their home cells lie in the range [ca, cb] of its
// first and last particle's home cells (give or take a container push), and the members of every
// neighbour column slice of a cell in that range lie in one contiguous run of the sorted arrays
// per column offset (dx, dy): the run of cells [ca + off - 1, cb + off + 1]. The nine runs are
// copied to LDS with coalesced loads (plus kPairPad spare entries each, for the four-wide reads);
// each particle then reads its candidates from LDS when its slice lies inside the staged run (any
// particle whose home left the range reads them from memory, as does a workgroup whose runs
// exceed its LDS share). The candidates, their order and the arithmetic are those of k_pairs.
template <class T, bool kTable>
__global__ void __launch_bounds__(512)
    k_pairs_staged(const Locator<T> loc, const Vec4<T>* __restrict__ pos,
                   const Vec4<T>* __restrict__ vsn, const uint32_t* __restrict__ offset,
                   const uint32_t* __restrict__ home, Vec4<T>* __restrict__ vout,
                   const NextBucket next) {
  extern __shared__ __align__(16) unsigned char stage_raw[];
  Vec4<T>* const stage = reinterpret_cast<Vec4<T>*>(stage_raw);
  __shared__ WgTiles wt;
  __shared__ uint32_t s_lo[9], s_hi[9], s_base[9];
  __shared__ uint32_t s_total;
  with_slice<T, kTable>(loc, [&](const SceneDev<T>& sc, int32_t p0, int32_t p1) {
    const int32_t p = p0 + (int32_t)threadIdx.x;
    const bool tiles = next.count && next.tiles.tc;
    if (tiles) wt.begin((int32_t)(home[p0] & kHomeCell));
    if (threadIdx.x < 9) {
      const int d = (int)threadIdx.x;
      const GridDev& g = sc.grid;
      const int32_t h0 = (int32_t)(home[p0] & kHomeCell), h1 = (int32_t)(home[p1 - 1] & kHomeCell);
      const int32_t col = ((d / 3 - 1) * g.n[1] + (d % 3 - 1)) * g.n[2];
      const int32_t first = g.cell_base, last = g.cell_base + g.n[0] * g.n[1] * g.n[2] - 1;
      const int32_t cs = max(min(h0, h1) + col - 1, first), ce = min(max(h0, h1) + col + 1, last);
      uint32_t lo = 0, hi = 0;
      if (cs <= ce) {
        lo = offset[cs];
        hi = offset[ce + 1];
      }
      s_lo[d] = lo;
      s_hi[d] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t b = 0;
      for (int d = 0; d < 9; ++d) {
        s_base[d] = b;
        b += (s_hi[d] - s_lo[d]) + (uint32_t)kPairPad;
      }
      s_total = b;
    }
    __syncthreads();
    const uint32_t total = s_total;
    const bool staged = total <= (uint32_t)(kStageFactor * (int)blockDim.x);
    if (staged) {
      // entry i of the staged runs belongs to run d with s_base[d] <= i < s_base[d + 1]; every
      // thread loads its (at most kStageFactor) entries at once, then stores them
      Vec4<T> buf[kStageFactor];
#pragma unroll
      for (int k = 0; k < kStageFactor; ++k) {
        const uint32_t i = threadIdx.x + (uint32_t)k * blockDim.x;
        if (i < total) {
          int d = 0;
#pragma unroll
          for (int e = 1; e < 9; ++e) d += (i >= s_base[e]) ? 1 : 0;
          // the sorted positions carry kPairPad spare entries, so the pads stay in the array
          buf[k] = pos[s_lo[d] + (i - s_base[d])];
        }
      }
#pragma unroll
      for (int k = 0; k < kStageFactor; ++k) {
        const uint32_t i = threadIdx.x + (uint32_t)k * blockDim.x;
        if (i < total) stage[i] = buf[k];
      }
    }
    __syncthreads();
    int32_t c = -1;
    if (p < p1) {
      PairSelf<T> me;
      uint32_t lo[9], hi[9];
      pair_setup(&sc, pos, vsn, offset, home, p, me, lo, hi);
      T ax = me.v.x, ay = me.v.y, az = me.v.z;
#pragma unroll
      for (int d = 0; d < 9; ++d) {
        const uint32_t e = hi[d];
        if (staged && lo[d] >= s_lo[d] && e <= s_hi[d]) {
          const Vec4<T>* src = stage + s_base[d] - s_lo[d];
          for (uint32_t q = lo[d]; q < e; q += 4) {
            Vec4<T> c[4];
            pair_load4(src + q, c);
            pair_test4(me, c, vsn, q, e, (uint32_t)p, ax, ay, az);
          }
        } else {
          for (uint32_t q = lo[d]; q < e; q += 4) pair_chunk4(me, pos, vsn, q, e, (uint32_t)p, ax, ay, az);
        }
      }
      const Vec4<T> vn{ax, ay, az, me.v.w};
      vout[p] = vn;
      if (next.count) c = next_bucket(sc, next, me.x, vn, p, next.arr ? home[p] : 0u);
    }
    if (tiles) {
      wt.add(next.tiles, c);
      wt.end(next.tiles);
    }
  });
}

// ---------------------------------------------------------------------------------------------
// Statistics of the current state (ParticleStatistics::calculateStatistics,
// ParticleStatistics.cpp:3-43), on a bucketing of the current positions: kinetic energy
// sum 0.5 m (v.v), the AABB of the spheres (x -/+ r), the number of particles i with an overlapping
// partner j of higher index (r_i + r_j - |x_j - x_i| > 0) and the largest such overlap, all in
// binary64. Contacts lie within the 27 cells around a particle's cell (the cell size is at least
// 2 r_max), so the pair search is O(N). One partial record per workgroup; the host combines them in
// workgroup order.
// ---------------------------------------------------------------------------------------------
template <class T, bool kTable>
__global__ void __launch_bounds__(256)
    k_stats(const Locator<T> loc, const Vec4<T>* __restrict__ pos,
            const Vec4<T>* __restrict__ vel, const int32_t* __restrict__ id,
            const uint32_t* __restrict__ offset, const uint32_t* __restrict__ home,
            StatPart* __restrict__ part) {
  __shared__ double s_ke[256], s_ov[256], s_mn[3][256], s_mx[3][256];
  __shared__ uint32_t s_nc[256];
  with_slice<T, kTable>(loc, [&](const SceneDev<T>& sc, int32_t p0, int32_t p1) {
    const int32_t t = (int32_t)threadIdx.x;
    const int32_t p = p0 + t;
    double ke = 0.0, ov = 0.0;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    uint32_t nc = 0;
    if (p < p1) {
      PairSelf<T> me;
      uint32_t lo[9], hi[9];
      pair_setup(&sc, pos, vel, offset, home, p, me, lo, hi);
      const double xi[3] = {(double)me.x.x, (double)me.x.y, (double)me.x.z};
      const double ri = (double)me.x.w;
      const double vv = ((double)me.v.x * (double)me.v.x + (double)me.v.y * (double)me.v.y) +
                        (double)me.v.z * (double)me.v.z;
      ke = 0.5 * (double)me.v.w * vv;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        mn[a] = xi[a] - ri;
        mx[a] = xi[a] + ri;
      }
      const int32_t mine = id[p];
      bool hit = false;
      for (int d = 0; d < 9; ++d)
        for (uint32_t q = lo[d]; q < hi[d]; ++q) {
          if (id[q] <= mine) continue;
          const Vec4<T> xj = pos[q];
          const double e0 = (double)xj.x - xi[0], e1 = (double)xj.y - xi[1],
                       e2 = (double)xj.z - xi[2];
          const double dist = sqrt((e0 * e0 + e1 * e1) + e2 * e2);  // glm::distance(x_i, x_j)
          const double cd = ri + (double)xj.w - dist;
          if (cd > 0.0) {
            hit = true;
            ov = fmax(ov, cd);
          }
        }
      nc = hit ? 1u : 0u;
    }
    s_ke[t] = ke;
    s_ov[t] = ov;
    s_nc[t] = nc;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      s_mn[a][t] = mn[a];
      s_mx[a][t] = mx[a];
    }
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {  // fixed tree
      if (t < w) {
        s_ke[t] += s_ke[t + w];
        s_ov[t] = fmax(s_ov[t], s_ov[t + w]);
        s_nc[t] += s_nc[t + w];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          s_mn[a][t] = fmin(s_mn[a][t], s_mn[a][t + w]);
          s_mx[a][t] = fmax(s_mx[a][t], s_mx[a][t + w]);
        }
      }
      __syncthreads();
    }
    if (t == 0) {
      StatPart r;
      r.ke = s_ke[0];
      r.max_overlap = s_ov[0];
      r.n_collided = s_nc[0];
      r.count = (uint32_t)max(p1 - p0, 0);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        r.mn[a] = s_mn[a][0];
        r.mx[a] = s_mx[a][0];
      }
      part[xcd_tile(blockIdx.x, loc.ntab)] = r;
    }
  });
}

// Per-scene totals of the workgroup partials, one thread per scene, in workgroup order: the same
// operations in the same order as the host's combination of the partials (running sums in
// binary64, std::min / std::max as (b < a) ? b : a and (a < b) ? b : a), so the two agree bit for
// bit. A scene's workgroups are the contiguous run of table entries naming it (found by binary
// search over the ascending scene numbers).
__global__ void __launch_bounds__(64)
    k_stats_combine(int32_t nscenes, int32_t ntab, const BlockInfo* __restrict__ btab,
                    const StatPart* __restrict__ part, psx_stats* __restrict__ out) {
  const int32_t s = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= nscenes) return;
  int32_t lo = 0, hi = ntab;  // first entry with scene >= s
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (btab[mid].scene < s) lo = mid + 1; else hi = mid;
  }
  psx_stats r;
  r.kinetic_energy = 0.0;
  r.n_collided = 0;
  r.max_overlap = 0.0;
  for (int a = 0; a < 3; ++a) {
    r.min[a] = INFINITY;
    r.max[a] = -INFINITY;
  }
  for (int32_t k = lo; k < ntab && btab[k].scene == s; ++k) {
    const StatPart q = part[k];
    r.kinetic_energy += q.ke;
    r.n_collided += (int64_t)q.n_collided;
    r.max_overlap = (r.max_overlap < q.max_overlap) ? q.max_overlap : r.max_overlap;
    for (int a = 0; a < 3; ++a) {
      r.min[a] = (q.mn[a] < r.min[a]) ? q.mn[a] : r.min[a];
      r.max[a] = (r.max[a] < q.mx[a]) ? q.mx[a] : r.max[a];
    }
  }
  out[s] = r;
}

// ---------------------------------------------------------------------------------------------
// state in/out (original order <-> sorted order)
// ---------------------------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256)
    k_unsort(int32_t n, const Vec4<T>* __restrict__ pos, const Vec4<T>* __restrict__ vel,
             const int32_t* __restrict__ id, T* __restrict__ pos3, T* __restrict__ vel3) {
  const int32_t p = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= n) return;
  const int64_t i = id[p];
  if (pos3) {
    const Vec4<T> x = pos[p];
    pos3[3 * i + 0] = x.x;
    pos3[3 * i + 1] = x.y;
    pos3[3 * i + 2] = x.z;
  }
  if (vel3) {
    const Vec4<T> v = vel[p];
    vel3[3 * i + 0] = v.x;
    vel3[3 * i + 1] = v.y;
    vel3[3 * i + 2] = v.z;
  }
}

template <class T>
__global__ void __launch_bounds__(256)
    k_sort_in(int32_t n, Vec4<T>* __restrict__ pos, Vec4<T>* __restrict__ vel,
              const int32_t* __restrict__ id, const T* __restrict__ pos3,
              const T* __restrict__ vel3) {
  const int32_t p = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= n) return;
  const int64_t i = id[p];
  if (pos3) {
    Vec4<T> x = pos[p];
    x.x = pos3[3 * i + 0];
    x.y = pos3[3 * i + 1];
    x.z = pos3[3 * i + 2];
    pos[p] = x;
  }
  if (vel3) {
    Vec4<T> v = vel[p];
    v.x = vel3[3 * i + 0];
    v.y = vel3[3 * i + 1];
    v.z = vel3[3 * i + 2];
    vel[p] = v;
  }
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
static inline int32_t blocks_for(int64_t n, int b) { return (int32_t)((n + b - 1) / b); }

template <class T>
static Locator<T> make_locator(int ntab, int per_block, const BlockInfo* btab,
                               const SceneDev<T>* scenes, const SceneDev<T>* one) {
  Locator<T> loc;
  loc.btab = btab;
  loc.scenes = scenes;
  loc.ntab = ntab;
  loc.per_block = per_block;
  if (one) loc.one = *one;
  else std::memset(&loc.one, 0, sizeof(loc.one));
  return loc;
}

template <class T>
hipError_t launch_integrate(hipStream_t st, int ntab, int block, const BlockInfo* btab,
                            const SceneDev<T>* scenes, const SceneDev<T>* one, bool blocks,
                            const Vec4<T>* pos, const Vec4<T>* vel, const int32_t* id,
                            int32_t* cell_of, uint32_t* count, int32_t* cell_ids,
                            uint32_t id_stride, int2* ovf, uint32_t* ovf_ctr,
                            const TileCounts& tiles, int mode, uint64_t* arr, uint8_t* mark) {
  const Locator<T> loc = make_locator(ntab, block, btab, scenes, one);
  const dim3 grid(xcd_grid(ntab)), blk(block);
  if (arr && (!blocks || tiles.tc || !mark)) return hipErrorInvalidValue;  // id slots, no direct scan
  unsigned long long* const arr_w = reinterpret_cast<unsigned long long*>(arr);
#define PSX_INTEGRATE(TABLE, BLOCKS)                                                             \
  PSX_LAUNCH((k_integrate<T, TABLE, BLOCKS>), grid, blk, 0, st, loc, pos, vel, id, cell_of,     \
             count, cell_ids, id_stride, ovf, ovf_ctr, tiles, mode, arr_w, mark)
  if (one) {
    if (blocks) PSX_INTEGRATE(false, true);
    else PSX_INTEGRATE(false, false);
  } else {
    if (blocks) PSX_INTEGRATE(true, true);
    else PSX_INTEGRATE(true, false);
  }
#undef PSX_INTEGRATE
  return hipGetLastError();
}

template <class T>
hipError_t launch_front(hipStream_t st, int ntab, int block, const BlockInfo* btab,
                        const SceneDev<T>* scenes, const SceneDev<T>* one, bool blocks,
                        Vec4<T>* pos, Vec4<T>* vel, const int32_t* id, const TriEntry<T>* tent,
                        const int32_t* tcell_off, const uint8_t* cflags, int32_t* cell_of,
                        uint32_t* count, int32_t* cell_ids, uint32_t id_stride, int2* ovf,
                        uint32_t* ovf_ctr, const TileCounts& tiles) {
  const Locator<T> loc = make_locator(ntab, block, btab, scenes, one);
  const dim3 grid(xcd_grid(ntab)), blk(block);
#define PSX_FRONT(TABLE, BLOCKS)                                                                 \
  PSX_LAUNCH((k_front<T, TABLE, BLOCKS>), grid, blk, 0, st, loc, pos, vel, id, tent, tcell_off,   \
             cflags, cell_of, count, cell_ids, id_stride, ovf, ovf_ctr, tiles)
  if (one) {
    if (blocks) PSX_FRONT(false, true);
    else PSX_FRONT(false, false);
  } else {
    if (blocks) PSX_FRONT(true, true);
    else PSX_FRONT(true, false);
  }
#undef PSX_FRONT
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_zero_u32(uint32_t* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void __launch_bounds__(64) k_add_u32(uint32_t* __restrict__ p, uint32_t v) {
  if (threadIdx.x == 0) p[0] += v;
}

hipError_t launch_add_u32(hipStream_t st, uint32_t* p, uint32_t v) {
  hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(64), 0, st, p, v);
  return hipGetLastError();
}

hipError_t launch_zero_u32(hipStream_t st, uint32_t* p, int64_t n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_scatter_ids(hipStream_t st, int block, int32_t n, const int32_t* cell_of,
                              uint32_t* count, const uint32_t* offset, const int32_t* id,
                              int32_t* list) {
  const int32_t tiles = (int32_t)((n + block - 1) / block);
  PSX_LAUNCH(k_scatter_ids, dim3(xcd_grid(tiles)), dim3(block), 0, st, n, tiles, cell_of, count,
             offset, id, list);
  return hipGetLastError();
}

hipError_t launch_scan(hipStream_t st, uint32_t* count, uint32_t* offset, uint64_t* tile_state,
                       int32_t ntiles, uint32_t* ovf_ctr, uint32_t* flags, bool clear, uint32_t n) {
  PSX_LAUNCH(k_scan, dim3(ntiles), dim3(kScanThreads), 0, st, count, offset, tile_state, ntiles,
             ovf_ctr, flags, clear ? 1 : 0, n);
  return hipGetLastError();
}

hipError_t launch_tile_totals(hipStream_t st, const uint32_t* count, const TileTotals& tt) {
  if (tt.shift == kDirectShift)
    PSX_LAUNCH(k_tile_totals<1>, dim3(tt.ntiles), dim3(kScanThreads), 0, st, count, tt);
  else if (tt.shift == kDirectShift + 2)
    PSX_LAUNCH(k_tile_totals<4>, dim3(tt.ntiles), dim3(kScanThreads), 0, st, count, tt);
  else if (tt.shift == kScanShift)
    PSX_LAUNCH(k_tile_totals<kScanPerThread / 4>, dim3(tt.ntiles), dim3(kScanThreads), 0, st, count, tt);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_scan_tiles(hipStream_t st, const uint32_t* count, uint32_t* offset,
                             const TileTotals& tt, int32_t ntiles, uint32_t* ovf_ctr, uint32_t* flags,
                             uint32_t n) {
  if (tt.shift == kDirectShift)
    PSX_LAUNCH(k_scan_tiles<1>, dim3(ntiles), dim3(kScanThreads), 0, st, count, offset, tt, ntiles,
               ovf_ctr, flags, n);
  else if (tt.shift == kDirectShift + 2)
    PSX_LAUNCH(k_scan_tiles<4>, dim3(ntiles), dim3(kScanThreads), 0, st, count, offset, tt, ntiles,
               ovf_ctr, flags, n);
  else if (tt.shift == kScanShift)
    PSX_LAUNCH(k_scan_tiles<kScanPerThread / 4>, dim3(ntiles), dim3(kScanThreads), 0, st, count,
               offset, tt, ntiles, ovf_ctr, flags, n);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_scan_direct(hipStream_t st, uint32_t* count, uint32_t* offset,
                              const TileCounts& tiles, uint32_t* ovf_ctr, bool clear) {
  PSX_LAUNCH(k_scan_direct, dim3(tiles.ntiles), dim3(kDirectThreads), 0, st, count, offset, tiles,
             ovf_ctr, clear ? 1 : 0);
  return hipGetLastError();
}

template <class T>
hipError_t launch_rank_gather(hipStream_t st, int ntab, int block, const BlockInfo* btab,
                              const SceneDev<T>* scenes, const SceneDev<T>* one, bool blocks,
                              int32_t n, const int32_t* cell_of, const uint32_t* offset,
                              const int32_t* cell_ids, uint32_t id_stride, const int2* ovf,
                              const uint32_t* ovf_ctr,
                              const int32_t* list, const Vec4<T>* pos, const Vec4<T>* vel,
                              const int32_t* id, const TriEntry<T>* tent,
                              const int32_t* tcell_off, uint32_t* flags, int mode,
                              const uint8_t* cull, Vec4<T>* pos_o, Vec4<T>* vel_o, int32_t* id_o,
                              int32_t* cell_o, uint32_t* home_o, Vec4<T>* xpre_o,
                              uint32_t* gen, const IncRank& inc) {
  const Locator<T> loc = make_locator(ntab, block, btab, scenes, one);
  const dim3 grid(xcd_grid(ntab)), blk(block);
  if (inc.arr && (!blocks || !inc.off_old || !inc.cell_cur_o || !inc.mark || !inc.rank_old || !inc.rank_o))
    return hipErrorInvalidValue;
#define PSX_RANK(TABLE, BLOCKS)                                                                  \
  PSX_LAUNCH((k_rank_gather<T, TABLE, BLOCKS>), grid, blk, 0, st, loc, cell_of, offset,         \
             cell_ids, id_stride, ovf, ovf_ctr, list, pos, vel, id, tent, tcell_off, flags, mode,\
             cull, n,                                                                          \
             pos_o, vel_o, id_o, cell_o, home_o, xpre_o, gen, inc)
  if (one) {
    if (blocks) PSX_RANK(false, true);
    else PSX_RANK(false, false);
  } else {
    if (blocks) PSX_RANK(true, true);
    else PSX_RANK(true, false);
  }
#undef PSX_RANK
  return hipGetLastError();
}

// The one-lane pair kernel with each wave's neighbour columns staged in LDS. A wave holds 64
// consecutive spheres of the cell order; those of one home column (x, y) find all their candidates
// in column (x + dx, y + dy) inside one contiguous run of the sorted arrays, from the cell below the
// lowest home cell to the cell above the highest. For each of the nine column offsets the wave
// loads that run once, with coalesced loads (a lane per entry, at most kWaveWin entries), stores it
// in its LDS window and walks every lane's own slice from there: each candidate is fetched once per
// wave instead of once per lane that tests it, which is what bounds the per-lane kernel (its vector
// memory path is busy throughout). A wave spanning two home columns stages one run for each; a
// candidate outside its lane's window (a third home column, a home cell moved by the container, a run
// longer than the window) is read from memory. The candidates, their order and the arithmetic are
// those of k_pairs, so the results are bit-identical.
constexpr int kWaveWin = 128;

// Stages entries [ws, ws + wn) of the sorted positions into win[0, wn) (wn <= kWaveWin).
template <class T>
__device__ __forceinline__ void wave_stage(const Vec4<T>* __restrict__ pos, uint32_t ws, uint32_t wn,
                                           Vec4<T>* win, int lane) {
  if ((uint32_t)lane < wn) win[lane] = pos[ws + (uint32_t)lane];
  if ((uint32_t)lane + 64u < wn) win[lane + 64] = pos[ws + 64u + (uint32_t)lane];
}

// The run of column (hx + dx, hy + dy) (column offset d) holding the slices of home cells
// (hx, hy, za .. zb): [ws, ws + wn) of the sorted arrays, wn clamped to the window; wn = 0 outside
// the grid. Every argument is uniform over the wave.
__device__ __forceinline__ void wave_run(const GridDev& g, const uint32_t* __restrict__ offset,
                                         int32_t hx, int32_t hy, int32_t za, int32_t zb, int d,
                                         uint32_t& ws, uint32_t& wn) {
  const int32_t bx = hx - 1 + d / 3, by = hy - 1 + d % 3;
  ws = 0;
  wn = 0;
  if (bx < 0 || bx >= g.n[0] || by < 0 || by >= g.n[1]) return;
  const int32_t zl = max(min(za, zb) - 1, 0), zh = min(max(za, zb) + 1, g.n[2] - 1);
  const int32_t cb = g.cell_base + (bx * g.n[1] + by) * g.n[2];
  ws = offset[cb + zl];
  wn = min(offset[cb + zh + 1] - ws, (uint32_t)kWaveWin);
}

template <class T, bool kTable>
__global__ void __launch_bounds__(1024)
    k_pairs_wave(const Locator<T> loc, const Vec4<T>* __restrict__ pos,
                 const Vec4<T>* __restrict__ vsn, const uint32_t* __restrict__ offset,
                 const uint32_t* __restrict__ home, Vec4<T>* __restrict__ vout,
                 const NextBucket next) {
  extern __shared__ __align__(16) unsigned char wave_raw[];
  __shared__ WgTiles wt;
  with_slice<T, kTable>(loc, [&](const SceneDev<T>& sc, int32_t p0, int32_t p1) {
    const int32_t p = p0 + (int32_t)threadIdx.x;
    const int lane = lane_id();
    Vec4<T>* const win =
        reinterpret_cast<Vec4<T>*>(wave_raw) + (size_t)(threadIdx.x >> 6) * (2 * kWaveWin);
    const bool tiles = next.count && next.tiles.tc;
    if (tiles) wt.begin((int32_t)(home[p0] & kHomeCell));
    const bool active = p < p1;  // the active lanes are a prefix of the wave
    const GridDev& g = sc.grid;
    int32_t c = -1;
    const uint64_t act = __ballot(active);
    if (act != 0) {
      PairSelf<T> me;
      uint32_t lo[9], hi[9];
      int32_t hx = 0, hy = 0, hz = 0, col = -1;
#pragma unroll
      for (int d = 0; d < 9; ++d) lo[d] = hi[d] = 0;
      if (active) {
        pair_setup(&sc, pos, vsn, offset, home, p, me, lo, hi);
        cell_coords(g, (int32_t)(home[p] & kHomeCell), hx, hy, hz);
        col = hx * g.n[1] + hy;
      }
      // the wave's first home column (lane 0's) and the first other one; the first and last lane
      // of each give the z range of its run
      const int32_t c0 = __builtin_amdgcn_readfirstlane(col);
      const uint64_t in0 = __ballot(active && col == c0);
      const uint64_t rest = __ballot(active && col != c0);
      const int f1 = rest ? __ffsll((unsigned long long)rest) - 1 : 0;
      const int32_t c1 = __builtin_amdgcn_readlane(col, f1);
      const uint64_t in1 = rest ? __ballot(active && col == c1) : 0ull;
      const int l0 = 63 - __clzll((long long)in0);
      const int l1 = in1 ? 63 - __clzll((long long)in1) : 0;
      const int32_t hx0 = __builtin_amdgcn_readfirstlane(hx), hy0 = __builtin_amdgcn_readfirstlane(hy);
      const int32_t z0a = __builtin_amdgcn_readfirstlane(hz), z0b = __builtin_amdgcn_readlane(hz, l0);
      const int32_t hx1 = __builtin_amdgcn_readlane(hx, f1), hy1 = __builtin_amdgcn_readlane(hy, f1);
      const int32_t z1a = __builtin_amdgcn_readlane(hz, f1), z1b = __builtin_amdgcn_readlane(hz, l1);
      const int grp = (col == c0) ? 0 : ((in1 != 0ull && col == c1) ? 1 : 2);
      T ax = me.v.x, ay = me.v.y, az = me.v.z;
#pragma unroll
      for (int d = 0; d < 9; ++d) {
        uint32_t ws0, wn0, ws1 = 0, wn1 = 0;
        wave_run(g, offset, hx0, hy0, z0a, z0b, d, ws0, wn0);
        if (in1 != 0ull) wave_run(g, offset, hx1, hy1, z1a, z1b, d, ws1, wn1);
        // every lane stores its entries, then reads those of the others: LDS operations of one
        // wave take effect in issue order, and the wait keeps the compiler from moving reads up
        wave_stage(pos, ws0, wn0, win, lane);
        if (in1 != 0ull) wave_stage(pos, ws1, wn1, win + kWaveWin, lane);
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        __builtin_amdgcn_wave_barrier();
        const uint32_t ws = grp == 1 ? ws1 : ws0;
        const uint32_t wn = grp == 0 ? wn0 : (grp == 1 ? wn1 : 0u);
        const Vec4<T>* wb = win + (grp == 1 ? kWaveWin : 0);
        const uint32_t e = hi[d];
        for (uint32_t q = lo[d]; q < e; ++q) {
          const uint32_t k = q - ws;
          Vec4<T> cj;
          if (k < wn) cj = wb[k];
          else cj = pos[q];
          if (q != (uint32_t)p && pair_near(me, cj)) pair_contact(me, cj, vsn, q, ax, ay, az);
        }
        __builtin_amdgcn_wave_barrier();
      }
      if (active) {
        const Vec4<T> vn{ax, ay, az, me.v.w};
        vout[p] = vn;
        if (next.count) c = next_bucket(sc, next, me.x, vn, p, next.arr ? home[p] : 0u);
      }
    }
    if (tiles) {
      wt.add(next.tiles, c);
      wt.end(next.tiles);
    }
  });
}

template <class T, bool kTable>
static hipError_t launch_pairs_t(hipStream_t st, const Locator<T>& loc, int block, int lanes,
                                 int chunk, const Vec4<T>* pos, const Vec4<T>* vsn,
                                 const uint32_t* offset, const uint32_t* home, Vec4<T>* vout,
                                 const NextBucket& next) {
  const dim3 grid(xcd_grid(loc.ntab)), blk(block);
  switch (lanes) {
    case 1: {
      const size_t shm = (size_t)kNearQ * (size_t)block * sizeof(uint32_t);
      if (chunk == 2)
        PSX_LAUNCH((k_pairs<T, kTable, 2>), grid, blk, shm, st, loc, pos, vsn, offset, home, vout, next);
      else
        PSX_LAUNCH((k_pairs<T, kTable, 4>), grid, blk, shm, st, loc, pos, vsn, offset, home, vout, next);
      break;
    }
    case 2:
      PSX_LAUNCH((k_pairs_lanes<T, 2, kTable>), grid, blk, 0, st, loc, pos, vsn, offset, home, vout, next);
      break;
    case 4:
      PSX_LAUNCH((k_pairs_lanes<T, 4, kTable>), grid, blk, 0, st, loc, pos, vsn, offset, home, vout, next);
      break;
    case 8:
      PSX_LAUNCH((k_pairs_lanes<T, 8, kTable>), grid, blk, 0, st, loc, pos, vsn, offset, home, vout, next);
      break;
    case 16:
      PSX_LAUNCH((k_pairs_lanes<T, 16, kTable>), grid, blk, 0, st, loc, pos, vsn, offset, home, vout, next);
      break;
    case kPairWave:
      PSX_LAUNCH((k_pairs_wave<T, kTable>), grid, blk,
                 (size_t)(block / 64) * 2 * kWaveWin * sizeof(Vec4<T>), st, loc, pos, vsn, offset,
                 home, vout, next);
      break;
    case kPairStaged:
      PSX_LAUNCH((k_pairs_staged<T, kTable>), grid, blk,
                 (size_t)kStageFactor * (size_t)block * sizeof(Vec4<T>), st, loc, pos, vsn, offset,
                 home, vout, next);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class T>
hipError_t launch_pairs(hipStream_t st, int ntab, int block, int lanes, int chunk,
                        const BlockInfo* btab, const SceneDev<T>* scenes, const SceneDev<T>* one,
                        const Vec4<T>* pos, const Vec4<T>* vsn, const uint32_t* offset,
                        const uint32_t* home, Vec4<T>* vout, const NextBucket& next) {
  const Locator<T> loc =
      make_locator(ntab, (lanes == kPairStaged || lanes == kPairWave) ? block : block / lanes, btab,
                   scenes, one);
  if (one)
    return launch_pairs_t<T, false>(st, loc, block, lanes, chunk, pos, vsn, offset, home, vout, next);
  return launch_pairs_t<T, true>(st, loc, block, lanes, chunk, pos, vsn, offset, home, vout, next);
}

template <class T>
hipError_t launch_stats(hipStream_t st, int ntab, const BlockInfo* btab, const SceneDev<T>* scenes,
                        const SceneDev<T>* one, const Vec4<T>* pos, const Vec4<T>* vel,
                        const int32_t* id, const uint32_t* offset, const uint32_t* home,
                        StatPart* part) {
  const Locator<T> loc = make_locator(ntab, 256, btab, scenes, one);
  if (one)
    hipLaunchKernelGGL((k_stats<T, false>), dim3(xcd_grid(ntab)), dim3(256), 0, st, loc, pos, vel,
                       id, offset, home, part);
  else
    hipLaunchKernelGGL((k_stats<T, true>), dim3(xcd_grid(ntab)), dim3(256), 0, st, loc, pos, vel,
                       id, offset, home, part);
  return hipGetLastError();
}

hipError_t launch_stats_combine(hipStream_t st, int32_t nscenes, int32_t ntab,
                                const BlockInfo* btab, const StatPart* part, psx_stats* out) {
  hipLaunchKernelGGL(k_stats_combine, dim3(blocks_for(nscenes, 64)), dim3(64), 0, st, nscenes,
                     ntab, btab, part, out);
  return hipGetLastError();
}

template <class T>
hipError_t launch_unsort(hipStream_t st, int32_t n, const Vec4<T>* pos, const Vec4<T>* vel,
                         const int32_t* id, T* pos3, T* vel3) {
  hipLaunchKernelGGL(k_unsort<T>, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, pos, vel, id, pos3,
                     vel3);
  return hipGetLastError();
}

template <class T>
hipError_t launch_sort_in(hipStream_t st, int32_t n, Vec4<T>* pos, Vec4<T>* vel, const int32_t* id,
                          const T* pos3, const T* vel3) {
  hipLaunchKernelGGL(k_sort_in<T>, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, pos, vel, id,
                     pos3, vel3);
  return hipGetLastError();
}

#define PSX_INSTANTIATE(T)                                                                      \
  template hipError_t launch_integrate<T>(hipStream_t, int, int, const BlockInfo*,             \
                                          const SceneDev<T>*, const SceneDev<T>*, bool,        \
                                          const Vec4<T>*, const Vec4<T>*, const int32_t*,      \
                                          int32_t*, uint32_t*, int32_t*, uint32_t, int2*, uint32_t*,     \
                                          const TileCounts&, int, uint64_t*, uint8_t*);                \
  template hipError_t launch_front<T>(hipStream_t, int, int, const BlockInfo*,                  \
                                      const SceneDev<T>*, const SceneDev<T>*, bool, Vec4<T>*,  \
                                      Vec4<T>*, const int32_t*, const TriEntry<T>*,            \
                                      const int32_t*, const uint8_t*, int32_t*, uint32_t*,     \
                                      int32_t*, uint32_t, int2*, uint32_t*, const TileCounts&); \
  template hipError_t launch_rank_gather<T>(hipStream_t, int, int, const BlockInfo*,            \
                                            const SceneDev<T>*, const SceneDev<T>*, bool,      \
                                            int32_t, const int32_t*, const uint32_t*,          \
                                            const int32_t*, uint32_t, const int2*, const uint32_t*,      \
                                            const int32_t*, const Vec4<T>*, const Vec4<T>*,    \
                                            const int32_t*, const TriEntry<T>*,                \
                                            const int32_t*, uint32_t*, int, const uint8_t*,    \
                                            Vec4<T>*, Vec4<T>*, int32_t*, int32_t*,            \
                                            uint32_t*, Vec4<T>*, uint32_t*, const IncRank&);              \
  template hipError_t launch_pairs<T>(hipStream_t, int, int, int, int, const BlockInfo*,            \
                                      const SceneDev<T>*, const SceneDev<T>*, const Vec4<T>*,  \
                                      const Vec4<T>*, const uint32_t*, const uint32_t*,        \
                                      Vec4<T>*, const NextBucket&);                            \
  template hipError_t launch_stats<T>(hipStream_t, int, const BlockInfo*, const SceneDev<T>*,    \
                                      const SceneDev<T>*, const Vec4<T>*, const Vec4<T>*,      \
                                      const int32_t*, const uint32_t*, const uint32_t*,        \
                                      StatPart*);                                              \
  template hipError_t launch_unsort<T>(hipStream_t, int32_t, const Vec4<T>*, const Vec4<T>*,    \
                                       const int32_t*, T*, T*);                                \
  template hipError_t launch_sort_in<T>(hipStream_t, int32_t, Vec4<T>*, Vec4<T>*,               \
                                        const int32_t*, const T*, const T*);

PSX_INSTANTIATE(float)
PSX_INSTANTIATE(double)

}  // namespace psx
