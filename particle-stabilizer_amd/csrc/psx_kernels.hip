// psx_kernels.hip -- the per-step kernels of the HIP backend (gfx950, wave64).
//
// One step (PhysicsEngine.cpp:38-48, CPU order: integrate, bucket from the pre-container positions,
// container, pairs) is these launches on one stream:
//   k_integrate    x += v*dt; v += g*dt; cell hash of the new x; box walls or mesh contact;
//                  per-cell count (atomic, returns the particle's arrival slot)
//   k_scan_reduce  \
//   k_scan_tiles    > exclusive scan of the per-cell counts into offsets; clears the counts
//   k_scan_down    /
//   k_scatter_ids  ids to offset[cell] + arrival slot (an unordered cell list)
//   k_rank_gather  rank of each particle among its cell mates by id (stable order) and gather of
//                  the state into cell-sorted order
//   k_pairs        27-cell spring/restitution velocity kick, Jacobi over the sorted snapshot
// Particle state lives in cell-sorted order between steps; an id array carries original indices.
#include "psx_kernels.hpp"

namespace psx {

// ---------------------------------------------------------------------------------------------
// wave64 helpers
// ---------------------------------------------------------------------------------------------
__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

__device__ inline uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// ---------------------------------------------------------------------------------------------
// container contact (per particle)
// ---------------------------------------------------------------------------------------------

// BoxContainer::resolveCollisions (BoxContainer.cpp:30-65): lo/hi from the incoming position, then
// xmin, ymin, zmin, xmax, ymax, zmax.
template <class T>
__device__ inline void box_walls(const SceneDev<T>* __restrict__ s, Vec4<T>& x, Vec4<T>& v) {
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
// all T triangles, only the particle's triangle-grid cell list is visited; it holds, in ascending
// order, every triangle the particle can touch, and the triangles it omits are exactly ones whose
// contact test fails (they change nothing), so the sequential position update and the velocity sum
// see the same triangles in the same order as the full loop.
template <class T>
__device__ inline void mesh_contact(const SceneDev<T>* __restrict__ s, const MeshDev* __restrict__ meshes,
                                    const TriDev<T>* __restrict__ tris,
                                    const int32_t* __restrict__ tcell_off,
                                    const int32_t* __restrict__ tcell_tri, Vec4<T>& x, Vec4<T>& v) {
  const MeshDev& m = meshes[s->mesh];
  const int32_t tx = axis_index((double)x.x, m.min[0], m.h, m.n[0]);
  const int32_t ty = axis_index((double)x.y, m.min[1], m.h, m.n[1]);
  const int32_t tz = axis_index((double)x.z, m.min[2], m.h, m.n[2]);
  const int32_t tc = m.tcell_base + (tx * m.n[1] + ty) * m.n[2] + tz;
  const int32_t k0 = tcell_off[tc], k1 = tcell_off[tc + 1];
  if (k0 == k1) return;
  const T r = x.w;
  const Vec3<T> p = {x.x, x.y, x.z};
  const T lox = x.x - r, loy = x.y - r, loz = x.z - r;
  const T hix = x.x + r, hiy = x.y + r, hiz = x.z + r;
  const T e = s->ew;
  Vec3<T> y = p;
  T ax = T(0), ay = T(0), az = T(0);
  int32_t cnt = 0;
  for (int32_t k = k0; k < k1; ++k) {
    const TriDev<T> t = tris[tcell_tri[k]];
    if (lox > t.hi[0] || hix < t.lo[0] || loy > t.hi[1] || hiy < t.lo[1] || loz > t.hi[2] ||
        hiz < t.lo[2])
      continue;
    const Vec3<T> q = closest_on_triangle(t, p);
    if (dist_to(p, q) < r) {
      const T s2 = T(2) * dot3(v.x, v.y, v.z, t.n[0], t.n[1], t.n[2]);
      ax = ax + e * (v.x - s2 * t.n[0]);
      ay = ay + e * (v.y - s2 * t.n[1]);
      az = az + e * (v.z - s2 * t.n[2]);
      const T dtp = dot3(-t.n[0], -t.n[1], -t.n[2], y.x - t.a[0], y.y - t.a[1], y.z - t.a[2]);
      const T mod = r - dtp;
      y.x = y.x - mod * t.n[0];
      y.y = y.y - mod * t.n[1];
      y.z = y.z - mod * t.n[2];
      ++cnt;
    }
  }
  if (cnt > 0) {
    const T cf = (T)cnt;
    x.x = y.x;
    x.y = y.y;
    x.z = y.z;
    v.x = ax / cf;
    v.y = ay / cf;
    v.z = az / cf;
  }
}

// ---------------------------------------------------------------------------------------------
// K1: integrate + cell hash + container + count
// ---------------------------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(1024)
    k_integrate(const BlockInfo* __restrict__ btab, const SceneDev<T>* __restrict__ scenes,
                Vec4<T>* __restrict__ pos, Vec4<T>* __restrict__ vel, int32_t* __restrict__ cell_of,
                uint32_t* __restrict__ slot, uint32_t* __restrict__ count,
                const MeshDev* __restrict__ meshes, const TriDev<T>* __restrict__ tris,
                const int32_t* __restrict__ tcell_off, const int32_t* __restrict__ tcell_tri,
                uint32_t* __restrict__ flags, int mode) {
  const BlockInfo bi = btab[blockIdx.x];
  const int32_t p = bi.p_start + (int32_t)threadIdx.x;
  if (p >= bi.p_end) return;
  const SceneDev<T>* __restrict__ s = scenes + bi.scene;
  Vec4<T> x = pos[p];
  Vec4<T> v = vel[p];
  if (mode & kModeIntegrate) {  // PhysicsEngine.cpp:22-25
    const T dt = s->dt;
    x.x = x.x + v.x * dt;
    x.y = x.y + v.y * dt;
    x.z = x.z + v.z * dt;
    v.x = v.x + s->gdt[0];
    v.y = v.y + s->gdt[1];
    v.z = v.z + s->gdt[2];
  }
  // ParticleModel::updateBuckets (ParticleModel.cpp:49-62): the cell of the pre-container position.
  const GridDev& g = s->grid;
  const int32_t ix = axis_index((double)x.x, g.min[0], g.h, g.n[0]);
  const int32_t iy = axis_index((double)x.y, g.min[1], g.h, g.n[1]);
  const int32_t iz = axis_index((double)x.z, g.min[2], g.h, g.n[2]);
  const int32_t c = g.cell_base + (ix * g.n[1] + iy) * g.n[2] + iz;
  if (mode & kModeContainer) {
    if (s->mesh < 0)
      box_walls(s, x, v);
    else
      mesh_contact(s, meshes, tris, tcell_off, tcell_tri, x, v);
  }
  if (mode & kModeCheck) {
    const bool ok = isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(v.x) &&
                    isfinite(v.y) && isfinite(v.z);
    if (!ok) atomicOr(flags, kFlagNonFinite);
  }
  if (mode & (kModeIntegrate | kModeContainer)) {
    pos[p] = x;
    vel[p] = v;
  }
  cell_of[p] = c;
  slot[p] = atomicAdd(count + c, 1u);
}

// ---------------------------------------------------------------------------------------------
// K3: exclusive scan of the per-cell counts (tile = 256 threads x 16 cells). Reduce, scan of the
// tile sums in one workgroup, then a down-sweep that rescans each tile with wave64 inclusive scans
// and LDS wave totals, adds the tile prefix, writes the offsets and clears the counts.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kScanThreads)
    k_scan_reduce(const uint32_t* __restrict__ count, uint32_t* __restrict__ tile_sums) {
  __shared__ uint32_t wsum[kScanThreads / 64];
  const uint4* __restrict__ src =
      reinterpret_cast<const uint4*>(count + (size_t)blockIdx.x * kScanTile);
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanPerThread / 4; ++q) {
    const uint4 u = src[q * kScanThreads + threadIdx.x];
    s += (u.x + u.y) + (u.z + u.w);
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) t += wsum[w];
    tile_sums[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(1024) k_scan_tiles(uint32_t* __restrict__ tile_sums, int32_t ntiles) {
  __shared__ uint32_t wsum[16];
  const int32_t per = (ntiles + 1023) / 1024;
  const int32_t b = (int32_t)threadIdx.x * per;
  const int32_t e = min(b + per, ntiles);
  uint32_t s = 0;
  for (int32_t i = b; i < e; ++i) s += tile_sums[i];
  const uint32_t wi = wave_inclusive_scan(s);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[w] = wi;
  __syncthreads();
  uint32_t wpre = 0;
  for (int q = 0; q < w; ++q) wpre += wsum[q];
  uint32_t run = wpre + wi - s;
  for (int32_t i = b; i < e; ++i) {
    const uint32_t c = tile_sums[i];
    tile_sums[i] = run;
    run += c;
  }
  if (threadIdx.x == 1023) tile_sums[ntiles] = run;
}

__global__ void __launch_bounds__(kScanThreads)
    k_scan_down(uint32_t* __restrict__ count, uint32_t* __restrict__ offset,
                const uint32_t* __restrict__ tile_sums, int32_t ntiles) {
  __shared__ uint32_t wsum[kScanThreads / 64];
  const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPerThread;
  uint4* __restrict__ src = reinterpret_cast<uint4*>(count + base);
  uint32_t v[kScanPerThread];
#pragma unroll
  for (int q = 0; q < kScanPerThread / 4; ++q) {
    const uint4 u = src[q];
    v[4 * q + 0] = u.x;
    v[4 * q + 1] = u.y;
    v[4 * q + 2] = u.z;
    v[4 * q + 3] = u.w;
  }
  uint32_t tot = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) tot += v[k];
  const uint32_t wi = wave_inclusive_scan(tot);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[w] = wi;
  __syncthreads();
  uint32_t run = tile_sums[blockIdx.x] + wi - tot;
  for (int q = 0; q < w; ++q) run += wsum[q];
  const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
  for (int q = 0; q < kScanPerThread / 4; ++q) {
    uint4 o;
    o.x = run;
    run += v[4 * q + 0];
    o.y = run;
    run += v[4 * q + 1];
    o.z = run;
    run += v[4 * q + 2];
    o.w = run;
    run += v[4 * q + 3];
    reinterpret_cast<uint4*>(offset + base)[q] = o;
    src[q] = zero;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) offset[(size_t)ntiles * kScanTile] = tile_sums[ntiles];
}

// ---------------------------------------------------------------------------------------------
// K4: unordered cell lists, then stable rank + gather
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(1024)
    k_scatter_ids(int32_t n, const int32_t* __restrict__ cell_of, const uint32_t* __restrict__ slot,
                  const uint32_t* __restrict__ offset, const int32_t* __restrict__ id,
                  int32_t* __restrict__ tmp) {
  const int32_t p = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= n) return;
  tmp[offset[cell_of[p]] + slot[p]] = id[p];
}

// Stable order inside a cell (the reference appends ids in ascending order,
// ParticleModel.cpp:55-61): rank = number of cell mates with a smaller id.
template <class T>
__global__ void __launch_bounds__(1024)
    k_rank_gather(int32_t n, const int32_t* __restrict__ cell_of, const uint32_t* __restrict__ offset,
                  const int32_t* __restrict__ tmp, const Vec4<T>* __restrict__ pos,
                  const Vec4<T>* __restrict__ vel, const int32_t* __restrict__ id,
                  Vec4<T>* __restrict__ pos_o, Vec4<T>* __restrict__ vel_o,
                  int32_t* __restrict__ id_o) {
  const int32_t p = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= n) return;
  const int32_t c = cell_of[p];
  const uint32_t b = offset[c], e = offset[c + 1];
  const int32_t me = id[p];
  uint32_t rank = 0;
  for (uint32_t k = b; k < e; ++k) rank += (tmp[k] < me) ? 1u : 0u;
  const uint32_t dst = b + rank;
  pos_o[dst] = pos[p];
  vel_o[dst] = vel[p];
  id_o[dst] = me;
}

// ---------------------------------------------------------------------------------------------
// K5: pair contact (ParticleModel::resolveCollisions / reboundWithOtherParticle,
// ParticleModel.cpp:64-144). One lane per particle in sorted order; the 27 neighbour cells are
// walked x-outer, y, z-inner. Cells are numbered z-fastest, so the three z-neighbours of an (x, y)
// column are one contiguous slice [offset[c(zlo)], offset[c(zhi)+1]) of the sorted arrays, and
// walking it visits cells in order and members by ascending id -- the reference's summation order.
// ---------------------------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(1024)
    k_pairs(const BlockInfo* __restrict__ btab, const SceneDev<T>* __restrict__ scenes,
            const Vec4<T>* __restrict__ pos, const Vec4<T>* __restrict__ vsn,
            const uint32_t* __restrict__ offset, Vec4<T>* __restrict__ vout) {
  const BlockInfo bi = btab[blockIdx.x];
  const int32_t p = bi.p_start + (int32_t)threadIdx.x;
  if (p >= bi.p_end) return;
  const SceneDev<T>* __restrict__ s = scenes + bi.scene;
  const GridDev& g = s->grid;
  const Vec4<T> xi = pos[p];
  const Vec4<T> vi = vsn[p];
  const int32_t hx = axis_index((double)xi.x, g.min[0], g.h, g.n[0]);
  const int32_t hy = axis_index((double)xi.y, g.min[1], g.h, g.n[1]);
  const int32_t hz = axis_index((double)xi.z, g.min[2], g.h, g.n[2]);
  const int32_t zlo = max(hz - 1, 0), zhi = min(hz + 1, g.n[2] - 1);
  const T ri = xi.w, mi = vi.w, ks = s->spring, es = s->es;
  const T margin = PairMargin<T>::value;
  T ax = vi.x, ay = vi.y, az = vi.z;
  for (int32_t bx = hx - 1; bx <= hx + 1; ++bx) {
    if (bx < 0 || bx >= g.n[0]) continue;
    for (int32_t by = hy - 1; by <= hy + 1; ++by) {
      if (by < 0 || by >= g.n[1]) continue;
      const int32_t cb = g.cell_base + (bx * g.n[1] + by) * g.n[2];
      const uint32_t q0 = offset[cb + zlo], q1 = offset[cb + zhi + 1];
      for (uint32_t q = q0; q < q1; ++q) {
        if ((int32_t)q == p) continue;
        const Vec4<T> xj = pos[q];
        // isHit: distance(x_i, x_j) = length(x_j - x_i) < r_i + r_j
        const T ex = xj.x - xi.x, ey = xj.y - xi.y, ez = xj.z - xi.z;
        const T dd = dot3(ex, ey, ez, ex, ey, ez);
        const T R = ri + xj.w;
        if (dd > (R * R) * margin) continue;
        const T len = sqrt(dd);
        if (!(len < R)) continue;
        const Vec4<T> vj = vsn[q];
        // d = x_i - x_j; dot(d, d) and length(d) equal dd and len bit for bit.
        const T dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const T ov = R - len;
        const T s3 = ((T(2) * vj.w) / (mi + vj.w)) *
                     (dot3(vi.x - vj.x, vi.y - vj.y, vi.z - vj.z, dx, dy, dz) / dd);
        const T inv = T(1) / len;
        const T kov = ks * ov;
        ax = ax + ((-((s3 * dx) * es)) + kov * (dx * inv));
        ay = ay + ((-((s3 * dy) * es)) + kov * (dy * inv));
        az = az + ((-((s3 * dz) * es)) + kov * (dz * inv));
      }
    }
  }
  vout[p] = Vec4<T>{ax, ay, az, mi};
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
static inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <class T>
hipError_t launch_integrate(hipStream_t st, int nblocks, int block, const BlockInfo* btab,
                            const SceneDev<T>* scenes, Vec4<T>* pos, Vec4<T>* vel,
                            int32_t* cell_of, uint32_t* slot, uint32_t* count,
                            const MeshDev* meshes, const TriDev<T>* tris,
                            const int32_t* tcell_off, const int32_t* tcell_tri, uint32_t* flags,
                            int mode) {
  hipLaunchKernelGGL(k_integrate<T>, dim3(nblocks), dim3(block), 0, st, btab, scenes, pos, vel,
                     cell_of, slot, count, meshes, tris, tcell_off, tcell_tri, flags, mode);
  return hipGetLastError();
}

hipError_t launch_scan_reduce(hipStream_t st, const uint32_t* count, uint32_t* tile_sums,
                              int32_t ntiles) {
  hipLaunchKernelGGL(k_scan_reduce, dim3(ntiles), dim3(kScanThreads), 0, st, count, tile_sums);
  return hipGetLastError();
}

hipError_t launch_scan_tiles(hipStream_t st, uint32_t* tile_sums, int32_t ntiles) {
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, st, tile_sums, ntiles);
  return hipGetLastError();
}

hipError_t launch_scan_down(hipStream_t st, uint32_t* count, uint32_t* offset,
                            const uint32_t* tile_sums, int32_t ntiles) {
  hipLaunchKernelGGL(k_scan_down, dim3(ntiles), dim3(kScanThreads), 0, st, count, offset,
                     tile_sums, ntiles);
  return hipGetLastError();
}

hipError_t launch_scatter_ids(hipStream_t st, int block, int32_t n, const int32_t* cell_of,
                              const uint32_t* slot, const uint32_t* offset, const int32_t* id,
                              int32_t* tmp) {
  hipLaunchKernelGGL(k_scatter_ids, dim3(blocks_for(n, block)), dim3(block), 0, st, n, cell_of,
                     slot, offset, id, tmp);
  return hipGetLastError();
}

template <class T>
hipError_t launch_rank_gather(hipStream_t st, int block, int32_t n, const int32_t* cell_of,
                              const uint32_t* offset, const int32_t* tmp, const Vec4<T>* pos,
                              const Vec4<T>* vel, const int32_t* id, Vec4<T>* pos_o,
                              Vec4<T>* vel_o, int32_t* id_o) {
  hipLaunchKernelGGL(k_rank_gather<T>, dim3(blocks_for(n, block)), dim3(block), 0, st, n, cell_of,
                     offset, tmp, pos, vel, id, pos_o, vel_o, id_o);
  return hipGetLastError();
}

template <class T>
hipError_t launch_pairs(hipStream_t st, int nblocks, int block, const BlockInfo* btab,
                        const SceneDev<T>* scenes, const Vec4<T>* pos, const Vec4<T>* vsn,
                        const uint32_t* offset, Vec4<T>* vout) {
  hipLaunchKernelGGL(k_pairs<T>, dim3(nblocks), dim3(block), 0, st, btab, scenes, pos, vsn, offset,
                     vout);
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

#define PSX_INSTANTIATE(T)                                                                       \
  template hipError_t launch_integrate<T>(hipStream_t, int, int, const BlockInfo*,              \
                                          const SceneDev<T>*, Vec4<T>*, Vec4<T>*, int32_t*,     \
                                          uint32_t*, uint32_t*, const MeshDev*, const TriDev<T>*, \
                                          const int32_t*, const int32_t*, uint32_t*, int);      \
  template hipError_t launch_rank_gather<T>(hipStream_t, int, int32_t, const int32_t*,           \
                                            const uint32_t*, const int32_t*, const Vec4<T>*,    \
                                            const Vec4<T>*, const int32_t*, Vec4<T>*, Vec4<T>*, \
                                            int32_t*);                                          \
  template hipError_t launch_pairs<T>(hipStream_t, int, int, const BlockInfo*,                  \
                                      const SceneDev<T>*, const Vec4<T>*, const Vec4<T>*,       \
                                      const uint32_t*, Vec4<T>*);                               \
  template hipError_t launch_unsort<T>(hipStream_t, int32_t, const Vec4<T>*, const Vec4<T>*,     \
                                       const int32_t*, T*, T*);                                 \
  template hipError_t launch_sort_in<T>(hipStream_t, int32_t, Vec4<T>*, Vec4<T>*,                \
                                        const int32_t*, const T*, const T*);

PSX_INSTANTIATE(float)
PSX_INSTANTIATE(double)

}  // namespace psx
