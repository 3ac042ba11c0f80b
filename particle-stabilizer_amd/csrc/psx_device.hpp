// psx_device.hpp -- device-side types and per-particle math of the HIP backend.
//
// Every arithmetic expression here is evaluated in the order the reference's CPU engine evaluates it
// (glm's scalar definitions: dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z, length = sqrt(dot),
// normalize(v) = v * (1/sqrt(dot(v,v))), distance(p,q) = length(q-p)). Sources are compiled with
// -ffp-contract=off and correctly rounded division and square root, so a kernel evaluated in
// binary64 reproduces the reference's double results bit for bit, and in binary32 it reproduces the
// same expression order rounded to float. The cell hash is always evaluated in binary64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace psx {

// 16-byte (float) / 32-byte (double) particle records: pos4 = (x, y, z, radius),
// vel4 = (vx, vy, vz, mass). Two vector loads per particle.
template <class T>
struct alignas(4 * sizeof(T)) Vec4 {
  T x, y, z, w;
};

template <class T>
struct Vec3 {
  T x, y, z;
};

// Uniform grid of one scene (ParticleBucket.cpp:17-38): min corner, cell size h (binary64),
// cell counts and the scene's first cell in the global cell numbering.
struct GridDev {
  double min[3];
  double h;
  int32_t n[3];
  int32_t cell_base;
  // division by n[2] and n[1] as multiply-high and shifts (Granlund and Montgomery, "Division by
  // invariant integers using multiplication", 1994, fig. 4.1): magic[a] and shift[a] = s1 | s2 << 8
  // for a = 0 (n[2]) and 1 (n[1])
  uint32_t magic[2];
  uint32_t shift[2];
};

// floor(u / d) for the divisor behind (m, sh): t = mulhi(m, u); (t + ((u - t) >> s1)) >> s2.
__device__ __forceinline__ uint32_t div_invariant(uint32_t u, uint32_t m, uint32_t sh) {
  const uint32_t t = __umulhi(m, u);
  return (t + ((u - t) >> (sh & 0xffu))) >> (sh >> 8);
}

// (ix, iy, iz) of cell c of grid g.
__device__ __forceinline__ void cell_coords(const GridDev& g, int32_t c, int32_t& ix, int32_t& iy,
                                            int32_t& iz) {
  const uint32_t u = (uint32_t)(c - g.cell_base);
  const uint32_t t = div_invariant(u, g.magic[0], g.shift[0]);  // u / nz
  iz = (int32_t)(u - t * (uint32_t)g.n[2]);
  const uint32_t x = div_invariant(t, g.magic[1], g.shift[1]);  // t / ny
  iy = (int32_t)(t - x * (uint32_t)g.n[1]);
  ix = (int32_t)x;
}

// Per-scene constants on the device, in the state's arithmetic type T.
template <class T>
struct SceneDev {
  GridDev grid;
  T bmin[3], bmax[3];
  // BoxContainer face coefficients (BoxContainer.cpp:14-19) by axis: emin[a] for the min face of
  // axis a, emax[a] for the max face.
  T emin[3], emax[3];
  T gdt[3];  // gravity * dt, evaluated once in T (PhysicsEngine.cpp:24 evaluates the same product)
  T dt;
  T spring;
  T es;      // sphere restitution
  T ew;      // mesh restitution (walls.zMin, PhysicsEngine.cpp:107-109)
  T rmax;    // largest radius of the scene
  int32_t p_begin, p_end;  // the scene's slice of the (sorted) particle arrays
  int32_t mesh;            // -1: box container; otherwise the scene has a triangle container
  int32_t tcell_base;      // first entry of the scene's cells in the triangle-cell offset table
  int32_t cuda;            // 1: the reference CUDA backend's step (PSX_FLAG_SEMANTICS_CUDA)
};

// Triangle of a mesh container: corners and normal (PolygonContainer.cpp:52-68).
template <class T>
struct TriGeo {
  T a[3], b[3], c[3], n[3];
};

// One candidate of a grid cell's triangle list: the triangle's AABB (PolygonContainer.cpp:70-89)
// stored inline, so the reference's AABB reject costs one load, and the triangle id.
template <class T>
struct alignas(sizeof(T) == 4 ? 32 : 64) TriBox {
  T lo[3], hi[3];
  int32_t tri;
};

// One candidate of a grid cell's triangle list as the forward container kernel reads it: the
// triangle's AABB first (one short load for the reject test), then its geometry in the same or the
// next cache line (80 bytes in binary32, 144 in binary64).
template <class T>
struct alignas(16) TriEntry {
  T lo[3], hi[3];
  TriGeo<T> g;
};

// Work assignment of one workgroup: its scene and slice of particles.
struct BlockInfo {
  int32_t scene;
  int32_t p_start;
  int32_t p_end;
  int32_t pad;
};

// ParticleBucket::toBucketIndex{X,Y,Z} (ParticleBucket.cpp:66-106): (int64)((x - min)/h) clamped
// to [0, n-1]. The cast follows x86-64 cvttsd2si, which the reference's build relies on: NaN and
// |q| >= 2^63 give INT64_MIN and therefore clamp to 0.
__device__ __host__ inline int32_t axis_index(double x, double mn, double h, int32_t n) {
  const double q = (x - mn) / h;
  if (!(q >= 1.0 && q < 9223372036854775808.0)) return 0;
  if (q >= (double)n) return n - 1;
  return (int32_t)q;
}

template <class T>
__device__ __host__ inline T dot3(T ax, T ay, T az, T bx, T by, T bz) {
  return (ax * bx + ay * by) + az * bz;
}

// PolygonContainer::getClosestPointOnPolygon and helpers (PolygonContainer.cpp:92-176).
template <class T>
__device__ inline Vec3<T> closest_on_segment(const Vec3<T>& p, const T* a, const T* b) {
  const T abx = b[0] - a[0], aby = b[1] - a[1], abz = b[2] - a[2];
  const T apx = p.x - a[0], apy = p.y - a[1], apz = p.z - a[2];
  T t = dot3(apx, apy, apz, abx, aby, abz) / dot3(abx, aby, abz, abx, aby, abz);
  const T m = (t < T(1)) ? t : T(1);
  t = (T(0) < m) ? m : T(0);
  return {a[0] + abx * t, a[1] + aby * t, a[2] + abz * t};
}

template <class T>
__device__ inline T dist_to(const Vec3<T>& p, const Vec3<T>& q) {
  const T dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
  return sqrt(dot3(dx, dy, dz, dx, dy, dz));
}

// kDotPick: the edge point is picked by dot(p, q) instead of the distance |q - p|, as the
// reference's CUDA backend does (getClosestPointOnPolygonKernelFunc,
// PolygonContainerCUDA_kernels.cu:125-154); everything else is the same computation there.
template <class T, bool kDotPick = false>
__device__ inline T edge_metric(const Vec3<T>& p, const Vec3<T>& q) {
  if constexpr (kDotPick) return dot3(p.x, p.y, p.z, q.x, q.y, q.z);
  return dist_to(p, q);
}

template <class T, bool kDotPick = false>
__device__ inline Vec3<T> closest_on_triangle(const TriGeo<T>& t, const Vec3<T>& p) {
  // projectPointOntoPlane(p, a, n)
  const T vx = p.x - t.a[0], vy = p.y - t.a[1], vz = p.z - t.a[2];
  const T ip = dot3(vx, vy, vz, t.n[0], t.n[1], t.n[2]);
  Vec3<T> proj = {p.x - t.n[0] * ip, p.y - t.n[1] * ip, p.z - t.n[2] * ip};
  // isPointOnPolygon(proj, a, b, c): v0 = c - a, v1 = b - a, v2 = proj - a
  {
    const T v0x = t.c[0] - t.a[0], v0y = t.c[1] - t.a[1], v0z = t.c[2] - t.a[2];
    const T v1x = t.b[0] - t.a[0], v1y = t.b[1] - t.a[1], v1z = t.b[2] - t.a[2];
    const T v2x = proj.x - t.a[0], v2y = proj.y - t.a[1], v2z = proj.z - t.a[2];
    const T d00 = dot3(v0x, v0y, v0z, v0x, v0y, v0z);
    const T d01 = dot3(v0x, v0y, v0z, v1x, v1y, v1z);
    const T d02 = dot3(v0x, v0y, v0z, v2x, v2y, v2z);
    const T d11 = dot3(v1x, v1y, v1z, v1x, v1y, v1z);
    const T d12 = dot3(v1x, v1y, v1z, v2x, v2y, v2z);
    const T inv = T(1) / (d00 * d11 - d01 * d01);
    const T u = (d11 * d02 - d01 * d12) * inv;
    const T w = (d00 * d12 - d01 * d02) * inv;
    if ((u >= T(0)) && (w >= T(0)) && (u + w <= T(1))) return proj;
  }
  const Vec3<T> qab = closest_on_segment(p, t.a, t.b);
  T best = edge_metric<T, kDotPick>(p, qab);
  proj = qab;
  const Vec3<T> qbc = closest_on_segment(p, t.b, t.c);
  const T dbc = edge_metric<T, kDotPick>(p, qbc);
  if (dbc < best) {
    best = dbc;
    proj = qbc;
  }
  const Vec3<T> qca = closest_on_segment(p, t.c, t.a);
  const T dca = edge_metric<T, kDotPick>(p, qca);
  if (dca < best) {
    best = dca;
    proj = qca;
  }
  return proj;
}

// Prefilter margin for the pair test: sqrt_rn(dd) < R can only hold when dd < R^2 (1 + margin).
template <class T>
struct PairMargin;
template <>
struct PairMargin<float> {
  static constexpr float value = 1.0f + 1e-5f;
};
template <>
struct PairMargin<double> {
  static constexpr double value = 1.0 + 1e-12;
};

}  // namespace psx
