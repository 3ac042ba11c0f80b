// psx_oracle.cpp -- CPU restatement of particle-stabilizer's CPU step (the reference's own
// implementation, src/ParticleStabilizer/*). TEST INFRASTRUCTURE ONLY: it is the checker used by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. Nothing in the product links,
// loads or calls it.
//
// Semantics restated (paths relative to the reference repository):
//   step order            PhysicsEngine.cpp:38-48 (integrate, bucket, container, pairs)
//   integrate             PhysicsEngine.cpp:19-26
//   grid geometry         ParticleBucket.cpp:17-38, ParticleBucket.hpp:17 (COEFF = (double)1.05f),
//                         ParticleModel.cpp:28-47 (r_max over all radii, start 0.0)
//   cell hash             ParticleBucket.cpp:66-106 (truncate toward zero, then clamp)
//   bucketing             ParticleModel.cpp:49-62, ParticleBucket.cpp:56-64,151-160
//   box container         BoxContainer.cpp:4-65
//   mesh container        PolygonContainer.cpp:8-90 (normals, AABBs), 92-286 (contact),
//                         288-318 (bounds); brute force over every triangle, OpenMP over
//                         particles exactly as PolygonContainer.cpp:244
//   pair contact          ParticleModel.cpp:64-144 (Jacobi snapshot, 27 cells x-outer, ascending
//                         member ids, skip self)
//   mass                  ParticleIO.cpp:131
// Vector arithmetic follows glm's scalar definitions: dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z,
// length(v) = sqrt(dot(v,v)), distance(p,q) = length(q - p), normalize(v) = v * (1 / sqrt(dot(v,v))),
// cross as glm::cross. The scalar type T is double (the reference's) or float (the same expression
// order evaluated in binary32); the cell hash is always evaluated in double.
//
// Pinning (each function is checked by tests/test_oracle_*.py): free-fall closed form, two-body
// closed forms and momentum of the restitution term, per-axis wall closed forms, brute-force O(N^2)
// pair enumeration in grid order, stable argsort / cumsum for bucketing, the textbook
// closest-point-on-triangle (Ericson, Real-Time Collision Detection 5.1.5), plane / crease closed
// forms for the mesh contact. glm's and assimp's exact versions are unknown (the reference's
// submodule is empty), so bit-level agreement with a glm build is "parity unpinned" beyond the glm
// definitions above.
//
// The reference's CUDA backend (src/ParticleStabilizerCUDA/*) is restated as well, for the
// PSX_FLAG_SEMANTICS_CUDA mode of the product (semantics "cuda" here):
//   step order            PhysicsEngineCUDA.cpp:36-43 (integrate, container, bucket, pairs)
//   grid geometry         ParticleModelCUDA.cpp:62-99 (width r_max * 2.0 * 1.05, exact 1.05)
//   pair contact          ParticleModel_kernels.cu:122-199 (spring term / m_i; kicks added onto v)
//   mesh container        PolygonContainerCUDA.cpp:30-99 (normals cross / length),
//                         PolygonContainerCUDA_kernels.cu:69-294 (edge point picked by
//                         dot(center, q), hits |q - x|^2 < r^2, summed corrections from x)
// The reference adds the pair kicks and the mesh terms with atomics in no fixed order and lets
// nvcc contract multiply-adds; here they are added in one of those orders (cells x-outer, y,
// z-inner, members by ascending id, triangles ascending) with every operation rounded on its own.
// The CUDA count kernel does not clamp a cell index (ParticleModel_kernels.cu:49-56); positions
// outside the grid are clamped here, as the neighbour search clamps them.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp, no -march: plain SSE2 double/float
// arithmetic, as the reference's Release build).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class T>
struct V3 {
  T x, y, z;
};

template <class T>
inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T>
inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T>
inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T>
inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T>
inline T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T>
inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T>
inline T distance(V3<T> p, V3<T> q) { return length(q - p); }
template <class T>
inline V3<T> normalize(V3<T> v) { return v * (T(1) / std::sqrt(dot(v, v))); }
template <class T>
inline V3<T> cross(V3<T> x, V3<T> y) {
  return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
template <class T>
inline V3<T> vmin(V3<T> a, V3<T> b) {  // glm::min: component (b < a) ? b : a
  return {b.x < a.x ? b.x : a.x, b.y < a.y ? b.y : a.y, b.z < a.z ? b.z : a.z};
}
template <class T>
inline V3<T> vmax(V3<T> a, V3<T> b) {  // glm::max: component (a < b) ? b : a
  return {a.x < b.x ? b.x : a.x, a.y < b.y ? b.y : a.y, a.z < b.z ? b.z : a.z};
}

// x86-64 cvttsd2si: truncation toward zero; NaN and |q| >= 2^63 give INT64_MIN (the reference's
// (int64_t) cast compiled by gcc for x86-64).
inline int64_t trunc_i64(double q) {
  if (q >= -9223372036854775808.0 && q < 9223372036854775808.0) return (int64_t)q;
  return INT64_MIN;
}

struct Grid {
  double min[3];
  double max[3];
  double h;
  int64_t n[3];
};

// ParticleBucket::initBucket, ParticleBucket.cpp:17-27; COEFF_FOR_INTERVAL_OF_BUCKETS = 1.05f stored
// in a double (ParticleBucket.hpp:17). cuda: ParticleModelCUDA::initBuckets
// (ParticleModelCUDA.cpp:62-99), bucketInterval = maxRadius * 2.0 * 1.05 and
// (int)((max - min) / interval) + 1 cells per axis.
Grid make_grid(double rmax, const double* mn, const double* mx, bool cuda = false) {
  Grid g;
  const double coeff = (double)1.05f;
  g.h = cuda ? rmax * 2.0 * 1.05 : rmax * 2.0 * coeff;
  for (int a = 0; a < 3; ++a) {
    g.min[a] = mn[a];
    g.max[a] = mx[a];
    g.n[a] = trunc_i64((mx[a] - mn[a]) / g.h) + 1;
  }
  return g;
}

// ParticleBucket::toBucketIndex{X,Y,Z}, ParticleBucket.cpp:66-106.
inline int64_t axis_index(double x, double mn, double h, int64_t n) {
  int64_t index = trunc_i64((x - mn) / h);
  if (index < 0) index = 0;
  if (index >= n) index = n - 1;
  return index;
}

template <class T>
inline void cell3(const Grid& g, V3<T> p, int64_t* c) {
  c[0] = axis_index((double)p.x, g.min[0], g.h, g.n[0]);
  c[1] = axis_index((double)p.y, g.min[1], g.h, g.n[1]);
  c[2] = axis_index((double)p.z, g.min[2], g.h, g.n[2]);
}

inline int64_t linear_cell(const Grid& g, const int64_t* c) {
  return (c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
}

template <class T>
struct Mesh {
  int64_t nt = 0;
  std::vector<V3<T>> a, b, c, n, lo, hi;
  double min[3], max[3];
  // Optional index for long test runs: the triangles whose AABB touches each cell of a uniform
  // grid over the mesh bounds (cell of a coordinate: floor((x - org) / cs), clamped). It only skips
  // triangles whose AABB is disjoint from the sphere's, which the loop below rejects anyway, so the
  // results are those of the full loop.
  bool indexed = false;
  double org[3], cs = 1.0;
  int64_t dims[3] = {1, 1, 1};
  std::vector<int64_t> cell_off;
  std::vector<int32_t> cell_tris;

  int64_t axis_cell(double x, int a) const {
    const double q = std::floor((x - org[a]) / cs);
    return q < 0.0 ? 0 : (q >= (double)(dims[a] - 1) ? dims[a] - 1 : (int64_t)q);
  }
  void build_index() {
    double ext = 0.0;
    for (int a = 0; a < 3; ++a) {
      org[a] = min[a];
      ext = std::max(ext, max[a] - min[a]);
    }
    cs = std::max(ext / 64.0, 1e-9);
    for (int a = 0; a < 3; ++a) dims[a] = (int64_t)((max[a] - min[a]) / cs) + 1;
    const int64_t ncell = dims[0] * dims[1] * dims[2];
    std::vector<std::vector<int32_t>> lists((size_t)ncell);
    for (int64_t t = 0; t < nt; ++t) {
      int64_t l[3], h[3];
      const double tl[3] = {(double)lo[t].x, (double)lo[t].y, (double)lo[t].z};
      const double th[3] = {(double)hi[t].x, (double)hi[t].y, (double)hi[t].z};
      for (int a = 0; a < 3; ++a) {
        l[a] = axis_cell(tl[a], a);
        h[a] = axis_cell(th[a], a);
      }
      for (int64_t i = l[0]; i <= h[0]; ++i)
        for (int64_t j = l[1]; j <= h[1]; ++j)
          for (int64_t k = l[2]; k <= h[2]; ++k)
            lists[(size_t)((i * dims[1] + j) * dims[2] + k)].push_back((int32_t)t);
    }
    cell_off.assign((size_t)ncell + 1, 0);
    for (int64_t c = 0; c < ncell; ++c) cell_off[(size_t)c + 1] = cell_off[(size_t)c] + (int64_t)lists[(size_t)c].size();
    cell_tris.clear();
    for (auto& l : lists) cell_tris.insert(cell_tris.end(), l.begin(), l.end());
    indexed = true;
  }
  // ascending ids of the triangles whose AABB may touch [l, h]
  void candidates(const V3<T>& l, const V3<T>& h, std::vector<int32_t>& out) const {
    out.clear();
    int64_t a0[3], a1[3];
    const double dl[3] = {(double)l.x, (double)l.y, (double)l.z};
    const double dh[3] = {(double)h.x, (double)h.y, (double)h.z};
    for (int a = 0; a < 3; ++a) {
      a0[a] = axis_cell(dl[a], a);
      a1[a] = axis_cell(dh[a], a);
    }
    for (int64_t i = a0[0]; i <= a1[0]; ++i)
      for (int64_t j = a0[1]; j <= a1[1]; ++j)
        for (int64_t k = a0[2]; k <= a1[2]; ++k) {
          const int64_t c = (i * dims[1] + j) * dims[2] + k;
          out.insert(out.end(), cell_tris.begin() + cell_off[(size_t)c],
                     cell_tris.begin() + cell_off[(size_t)c + 1]);
        }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
  }
};

// PolygonContainer ctor (PolygonContainer.cpp:52-89) and bounds (288-318). Normals are evaluated in
// double from the double vertices and then stored in T; AABBs are taken over the T vertices.
// cuda: PolygonContainerCUDA ctor (PolygonContainerCUDA.cpp:30-99): the same cross product
// (relVec0 x relVec1 with relVec0 = v1 - v0, relVec1 = v2 - v0), divided component by component
// by its length; the same AABBs and bounds (getMinCoords / getMaxCoords, :228-262).
template <class T>
Mesh<T> make_mesh(int64_t nt, const double* v, bool cuda = false) {
  Mesh<T> m;
  m.nt = nt;
  m.a.resize(nt);
  m.b.resize(nt);
  m.c.resize(nt);
  m.n.resize(nt);
  m.lo.resize(nt);
  m.hi.resize(nt);
  for (int64_t t = 0; t < nt; ++t) {
    const double* p = v + 9 * t;
    V3<double> a{p[0], p[1], p[2]}, b{p[3], p[4], p[5]}, c{p[6], p[7], p[8]};
    const V3<double> cr = cross(b - a, c - a);
    const V3<double> nn = cuda ? cr / length(cr) : normalize(cr);
    m.a[t] = {(T)a.x, (T)a.y, (T)a.z};
    m.b[t] = {(T)b.x, (T)b.y, (T)b.z};
    m.c[t] = {(T)c.x, (T)c.y, (T)c.z};
    m.n[t] = {(T)nn.x, (T)nn.y, (T)nn.z};
    m.lo[t] = vmin(vmin(m.a[t], m.b[t]), m.c[t]);
    m.hi[t] = vmax(vmax(m.a[t], m.b[t]), m.c[t]);
  }
  for (int k = 0; k < 3; ++k) {
    m.min[k] = v[k];
    m.max[k] = v[k];
  }
  for (int64_t i = 0; i < 3 * nt; ++i) {
    for (int k = 0; k < 3; ++k) {
      const double x = v[3 * i + k];
      if (x < m.min[k]) m.min[k] = x;
      if (m.max[k] < x) m.max[k] = x;
    }
  }
  return m;
}

// PolygonContainer::projectPointOntoPlane / isPointOnPolygon / closestPointOnLineSegment /
// getClosestPointOnPolygon, PolygonContainer.cpp:92-176.
template <class T>
inline V3<T> project_onto_plane(V3<T> p, V3<T> a, V3<T> n) {
  const V3<T> v = p - a;
  const T ip = dot(v, n);
  return p - n * ip;
}

template <class T>
inline bool point_on_triangle(V3<T> p, V3<T> a, V3<T> b, V3<T> c) {
  const V3<T> v0 = c - a, v1 = b - a, v2 = p - a;
  const T d00 = dot(v0, v0), d01 = dot(v0, v1), d02 = dot(v0, v2), d11 = dot(v1, v1),
          d12 = dot(v1, v2);
  const T inv = T(1) / (d00 * d11 - d01 * d01);
  const T u = (d11 * d02 - d01 * d12) * inv;
  const T w = (d00 * d12 - d01 * d02) * inv;
  return (u >= T(0)) && (w >= T(0)) && (u + w <= T(1));
}

template <class T>
inline V3<T> closest_on_segment(V3<T> p, V3<T> a, V3<T> b) {
  const V3<T> ab = b - a, ap = p - a;
  T t = dot(ap, ab) / dot(ab, ab);
  const T m = (t < T(1)) ? t : T(1);  // std::min(1.0, t)
  t = (T(0) < m) ? m : T(0);           // std::max(0.0, m)
  return a + ab * t;
}

template <class T>
inline V3<T> closest_on_triangle(V3<T> a, V3<T> b, V3<T> c, V3<T> n, V3<T> p) {
  V3<T> proj = project_onto_plane(p, a, n);
  if (point_on_triangle(proj, a, b, c)) return proj;
  const V3<T> qab = closest_on_segment(p, a, b);
  T best = distance(p, qab);
  proj = qab;
  const V3<T> qbc = closest_on_segment(p, b, c);
  const T dbc = distance(p, qbc);
  if (dbc < best) {
    best = dbc;
    proj = qbc;
  }
  const V3<T> qca = closest_on_segment(p, c, a);
  const T dca = distance(p, qca);
  if (dca < best) {
    best = dca;
    proj = qca;
  }
  return proj;
}

// getClosestPointOnPolygonKernelFunc (PolygonContainerCUDA_kernels.cu:125-154): the same
// projection, inside test and segment points as closest_on_triangle, but the edge point is picked by
// dot(center, q), not by the distance |q - center|.
template <class T>
inline V3<T> closest_on_triangle_cuda(V3<T> a, V3<T> b, V3<T> c, V3<T> n, V3<T> p) {
  V3<T> proj = project_onto_plane(p, a, n);
  if (point_on_triangle(proj, a, b, c)) return proj;
  const V3<T> qab = closest_on_segment(p, a, b);
  T best = dot(p, qab);
  proj = qab;
  const V3<T> qbc = closest_on_segment(p, b, c);
  const T dbc = dot(p, qbc);
  if (dbc < best) {
    best = dbc;
    proj = qbc;
  }
  const V3<T> qca = closest_on_segment(p, c, a);
  const T dca = dot(p, qca);
  if (dca < best) {
    best = dca;
    proj = qca;
  }
  return proj;
}

template <class T>
struct Params {
  V3<T> g;
  T dt, spring, es, ew_mesh;
  T walls[6];  // zMin, zMax, xMin, xMax, yMin, yMax
  V3<T> bmin, bmax;
};

template <class T>
struct Sim {
  int64_t n;
  std::vector<V3<T>> x, v;
  std::vector<T> r, m;
  Params<T> p;
  Grid g;
  const Mesh<T>* mesh = nullptr;
  std::vector<std::vector<int64_t>> buckets;
  // fast (test runs): the pair loop's particles are split among threads (each particle's sum is
  // its own, read from the snapshot, so the results are those of the serial loop)
  bool fast = false;
  int nthreads = 1;
  bool cuda = false;  // the reference CUDA backend's step
  // scratch
  std::vector<V3<T>> sx, sv, bv, ybuf, accbuf;
  std::vector<int64_t> counter;
};

template <class T>
void integrate(Sim<T>& s) {  // PhysicsEngine.cpp:19-26
  for (int64_t i = 0; i < s.n; ++i) {
    s.x[i] = s.x[i] + s.v[i] * s.p.dt;
    s.v[i] = s.v[i] + s.p.g * s.p.dt;
  }
}

template <class T>
void update_buckets(Sim<T>& s) {  // ParticleModel.cpp:49-62
  for (auto& b : s.buckets) b.clear();
  int64_t c[3];
  for (int64_t i = 0; i < s.n; ++i) {
    cell3(s.g, s.x[i], c);
    s.buckets[linear_cell(s.g, c)].push_back(i);
  }
}

template <class T>
void box_walls(Sim<T>& s) {  // BoxContainer.cpp:22-65
  const T* e = s.p.walls;    // [0]=zMin [1]=zMax [2]=xMin [3]=xMax [4]=yMin [5]=yMax
  for (int64_t i = 0; i < s.n; ++i) {
    V3<T>& x = s.x[i];
    V3<T>& v = s.v[i];
    const T r = s.r[i];
    const V3<T> lo = {x.x - r, x.y - r, x.z - r};
    const V3<T> hi = {x.x + r, x.y + r, x.z + r};
    if (lo.x < s.p.bmin.x) {
      x.x = s.p.bmin.x + r;
      v.x = -v.x * e[2];
    }
    if (lo.y < s.p.bmin.y) {
      x.y = s.p.bmin.y + r;
      v.y = -v.y * e[4];
    }
    if (lo.z < s.p.bmin.z) {
      x.z = s.p.bmin.z + r;
      v.z = -v.z * e[0];
    }
    if (hi.x > s.p.bmax.x) {
      x.x = s.p.bmax.x - r;
      v.x = -v.x * e[3];
    }
    if (hi.y > s.p.bmax.y) {
      x.y = s.p.bmax.y - r;
      v.y = -v.y * e[5];
    }
    if (hi.z > s.p.bmax.z) {
      x.z = s.p.bmax.z - r;
      v.z = -v.z * e[1];
    }
  }
}

template <class T>
void mesh_contact(Sim<T>& s, int nthreads) {  // PolygonContainer.cpp:237-286
  const Mesh<T>& m = *s.mesh;
  const T e = s.p.ew_mesh;
  const int64_t n = s.n;
  s.ybuf = s.x;                          // copyParticleBufferPosition
  s.accbuf.assign(n, V3<T>{0, 0, 0});    // clearParticleBufferVelocity
  s.counter.assign(n, 0);                // clearIntersectCounter
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
  {
  std::vector<int32_t> cand;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
  for (int64_t i = 0; i < n; ++i) {
    const V3<T> x = s.x[i];
    const V3<T> v = s.v[i];
    const T r = s.r[i];
    const V3<T> lo = {x.x - r, x.y - r, x.z - r};
    const V3<T> hi = {x.x + r, x.y + r, x.z + r};
    V3<T> y = s.ybuf[i];
    V3<T> acc = s.accbuf[i];
    int64_t cnt = 0;
    if (m.indexed) m.candidates(lo, hi, cand);
    const int64_t ncand = m.indexed ? (int64_t)cand.size() : m.nt;
    for (int64_t k = 0; k < ncand; ++k) {
      const int64_t t = m.indexed ? (int64_t)cand[(size_t)k] : k;
      if (lo.x > m.hi[t].x || hi.x < m.lo[t].x || lo.y > m.hi[t].y || hi.y < m.lo[t].y ||
          lo.z > m.hi[t].z || hi.z < m.lo[t].z)
        continue;
      const V3<T> q = closest_on_triangle(m.a[t], m.b[t], m.c[t], m.n[t], x);
      const V3<T> rel = q - x;
      const T dist = length(rel);
      if (dist < r) {
        const V3<T> nn = m.n[t];
        const V3<T> refl = v - (T(2) * dot(v, nn)) * nn;
        acc = acc + e * refl;
        const T dtp = dot(-nn, y - m.a[t]);
        const T mod = r - dtp;
        y = y - mod * nn;
        ++cnt;
      }
    }
    s.ybuf[i] = y;
    s.accbuf[i] = acc;
    s.counter[i] = cnt;
  }
  }
  for (int64_t i = 0; i < n; ++i) {
    if (s.counter[i] > 0) {
      s.x[i] = s.ybuf[i];
      s.v[i] = s.accbuf[i] / (T)s.counter[i];
    }
  }
}

// PolygonContainerCUDA::resolveCollisions (PolygonContainerCUDA.cpp:130-214): isIntersectedKernel
// (PolygonContainerCUDA_kernels.cu:179-211: AABB overlap with <= / >=, then |q - x|^2 < r^2),
// calcModifiedVelocitySumKernel and calcAveragedVelocityKernel (:214-252: v = mean of
// e (v - (2 v.n) n) over the hits, when there is one), calcPosModificationAmountKernel and
// modifyPositionKernel (:254-286: x += sum of -(r - dot(-n, x - a)) n, for every particle). All terms
// come from the incoming x and v; the atomics' sums are taken in ascending triangle order.
template <class T>
void mesh_contact_cuda(Sim<T>& s, int nthreads) {
  const Mesh<T>& m = *s.mesh;
  const T e = s.p.ew_mesh;
  const int64_t n = s.n;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
  {
  std::vector<int32_t> cand;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
  for (int64_t i = 0; i < n; ++i) {
    const V3<T> x = s.x[i];
    const V3<T> v = s.v[i];
    const T r = s.r[i];
    const V3<T> lo = {x.x - r, x.y - r, x.z - r};
    const V3<T> hi = {x.x + r, x.y + r, x.z + r};
    V3<T> vb{0, 0, 0}, mod{0, 0, 0};
    int64_t cnt = 0;
    if (m.indexed) m.candidates(lo, hi, cand);
    const int64_t ncand = m.indexed ? (int64_t)cand.size() : m.nt;
    for (int64_t k = 0; k < ncand; ++k) {
      const int64_t t = m.indexed ? (int64_t)cand[(size_t)k] : k;
      if (!(lo.x <= m.hi[t].x && hi.x >= m.lo[t].x && lo.y <= m.hi[t].y && hi.y >= m.lo[t].y &&
            lo.z <= m.hi[t].z && hi.z >= m.lo[t].z))
        continue;
      const V3<T> q = closest_on_triangle_cuda(m.a[t], m.b[t], m.c[t], m.n[t], x);
      const V3<T> rel = q - x;
      if (!(dot(rel, rel) < r * r)) continue;
      const V3<T> nn = m.n[t];
      const V3<T> refl = v - (T(2) * dot(v, nn)) * nn;
      vb = vb + e * refl;
      const T sc = r - dot(T(-1) * nn, x - m.a[t]);
      mod = mod + (-sc) * nn;
      ++cnt;
    }
    if (cnt > 0) s.v[i] = vb / (T)cnt;
    s.x[i] = x + mod;
  }
  }
}

template <class T>
void pair_contact(Sim<T>& s) {  // ParticleModel.cpp:64-144
  const int64_t n = s.n;
  s.sx = s.x;  // the Jacobi snapshot S (copyParticleData into _particlesBuffer, then read _particles)
  s.sv = s.v;
  s.bv = s.v;
  const T k = s.p.spring, es = s.p.es;
#ifdef _OPENMP
#pragma omp parallel for num_threads(s.nthreads) schedule(static) if (s.fast)
#endif
  for (int64_t i = 0; i < n; ++i) {
    int64_t hc[3];
    cell3(s.g, s.sx[i], hc);
    const V3<T> xi = s.sx[i], vi = s.sv[i];
    const T ri = s.r[i], mi = s.m[i];
    V3<T> acc = s.bv[i];
    for (int64_t ox = 0; ox < 3; ++ox) {
      const int64_t bx = hc[0] - 1 + ox;
      for (int64_t oy = 0; oy < 3; ++oy) {
        const int64_t by = hc[1] - 1 + oy;
        for (int64_t oz = 0; oz < 3; ++oz) {
          const int64_t bz = hc[2] - 1 + oz;
          if (!(bx >= 0 && bx < s.g.n[0] && by >= 0 && by < s.g.n[1] && bz >= 0 &&
                bz < s.g.n[2]))
            continue;
          const int64_t c3[3] = {bx, by, bz};
          for (const int64_t j : s.buckets[linear_cell(s.g, c3)]) {
            if (j == i) continue;
            const V3<T> xj = s.sx[j];
            const T rj = s.r[j];
            if (s.cuda) {  // resolveCollisionsKernel, ParticleModel_kernels.cu:163-199
              const V3<T> d = xi - xj;
              const T dsq = (d.x * d.x + d.y * d.y) + d.z * d.z;
              const T dist = std::sqrt(dsq);
              const T ovc = (ri + rj) - dist;
              if (!(ovc > T(0))) continue;
              const V3<T> vj = s.sv[j];
              const T mj = s.m[j];
              const V3<T> rv = vi - vj;
              const T rvd = (rv.x * d.x + rv.y * d.y) + rv.z * d.z;
              const T crel = (((T(2) * mj) * rvd) * es) / ((mi + mj) * dsq);
              const T cspr = (k * ovc) / (dist * mi);
              const T f = cspr - crel;
              acc = V3<T>{acc.x + f * d.x, acc.y + f * d.y, acc.z + f * d.z};
              continue;
            }
            if (!(distance(xi, xj) < ri + rj)) continue;  // isHit
            const V3<T> vj = s.sv[j];
            const T mj = s.m[j];
            const V3<T> d = xi - xj;
            const T ov = ri + rj - length(d);
            const V3<T> rel = ((T(2) * mj / (mi + mj)) * (dot(vi - vj, d) / dot(d, d))) * d * es;
            const V3<T> spr = k * ov * normalize(d);
            acc = acc + ((-rel) + spr);
          }
        }
      }
    }
    s.bv[i] = acc;
  }
  s.v = s.bv;
}

template <class T>
void step(Sim<T>& s, int nthreads) {  // PhysicsEngine.cpp:38-48
  if (s.cuda) {  // PhysicsEngineCUDA.cpp:36-43: container before bucketing
    integrate(s);
    if (s.mesh)
      mesh_contact_cuda(s, nthreads);
    else
      box_walls(s);
    update_buckets(s);
    pair_contact(s);
    return;
  }
  integrate(s);
  update_buckets(s);
  if (s.mesh)
    mesh_contact(s, nthreads);
  else
    box_walls(s);
  pair_contact(s);
}

template <class T>
Params<T> make_params(const double* p) {
  // p: gravity[3], walls[6], spring, restitution, dt, box_min[3], box_max[3]
  Params<T> q;
  q.g = {(T)p[0], (T)p[1], (T)p[2]};
  for (int k = 0; k < 6; ++k) q.walls[k] = (T)p[3 + k];
  q.spring = (T)p[9];
  q.es = (T)p[10];
  q.dt = (T)p[11];
  q.bmin = {(T)p[12], (T)p[13], (T)p[14]};
  q.bmax = {(T)p[15], (T)p[16], (T)p[17]};
  q.ew_mesh = (T)p[3];  // walls.zMin, PhysicsEngine.cpp:107-109
  return q;
}

template <class T>
int run(const double* params, int64_t n, T* pos, T* vel, const T* r, const T* m, int64_t ntri,
        const double* tri, int64_t nsteps, int nthreads, int opts) {
  const int fast = opts & 1;
  const bool cuda = (opts & 2) != 0;
  if (n <= 0 || !params || !pos || !vel || !r || !m) return 1;
  Sim<T> s;
  s.n = n;
  s.x.resize(n);
  s.v.resize(n);
  s.r.assign(r, r + n);
  s.m.assign(m, m + n);
  for (int64_t i = 0; i < n; ++i) {
    s.x[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    s.v[i] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
  }
  s.p = make_params<T>(params);
  s.fast = fast != 0;
  s.nthreads = std::max(nthreads, 1);
  s.cuda = cuda;
  Mesh<T> mesh;
  double mn[3], mx[3];
  if (ntri > 0) {
    mesh = make_mesh<T>(ntri, tri, cuda);
    if (fast) mesh.build_index();
    s.mesh = &mesh;
    for (int k = 0; k < 3; ++k) {
      mn[k] = mesh.min[k];
      mx[k] = mesh.max[k];
    }
  } else {
    for (int k = 0; k < 3; ++k) {
      mn[k] = params[12 + k];
      mx[k] = params[15 + k];
    }
  }
  double rmax = 0.0;  // ParticleModel::initBuckets, ParticleModel.cpp:35-38
  for (int64_t i = 0; i < n; ++i) rmax = std::max(rmax, (double)r[i]);
  s.g = make_grid(rmax, mn, mx, cuda);
  s.buckets.resize((size_t)(s.g.n[0] * s.g.n[1] * s.g.n[2]));
  for (int64_t it = 0; it < nsteps; ++it) step(s, nthreads);
  for (int64_t i = 0; i < n; ++i) {
    pos[3 * i] = s.x[i].x;
    pos[3 * i + 1] = s.x[i].y;
    pos[3 * i + 2] = s.x[i].z;
    vel[3 * i] = s.v[i].x;
    vel[3 * i + 1] = s.v[i].y;
    vel[3 * i + 2] = s.v[i].z;
  }
  return 0;
}

template <class T>
int bucketing(double rmax, const double* mn, const double* mx, int64_t n, const T* pos,
              int32_t* cell, uint32_t* count, uint32_t* offset, int32_t* perm) {
  Grid g = make_grid(rmax, mn, mx);
  const int64_t C = g.n[0] * g.n[1] * g.n[2];
  std::vector<std::vector<int64_t>> buckets((size_t)C);
  int64_t c[3];
  for (int64_t i = 0; i < n; ++i) {
    cell3(g, V3<T>{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}, c);
    const int64_t lc = linear_cell(g, c);
    if (cell) cell[i] = (int32_t)lc;
    buckets[lc].push_back(i);
  }
  uint32_t run_off = 0;
  for (int64_t k = 0; k < C; ++k) {
    if (count) count[k] = (uint32_t)buckets[k].size();
    if (offset) offset[k] = run_off;
    for (size_t q = 0; q < buckets[k].size(); ++q)
      if (perm) perm[run_off + q] = (int32_t)buckets[k][q];
    run_off += (uint32_t)buckets[k].size();
  }
  if (offset) offset[C] = run_off;
  return 0;
}

}  // namespace

extern "C" {

// Grid geometry of ParticleBucket::initBucket (cuda: ParticleModelCUDA::initBuckets): h and cell
// counts.
void oracle_grid(double rmax, const double* mn, const double* mx, int cuda, double* h, int64_t* n3) {
  Grid g = make_grid(rmax, mn, mx, cuda != 0);
  *h = g.h;
  for (int a = 0; a < 3; ++a) n3[a] = g.n[a];
}

int64_t oracle_axis_index(double x, double mn, double h, int64_t n) {
  return axis_index(x, mn, h, n);
}

// Full runs. params = gravity[3], walls[6], spring, restitution, dt, box_min[3], box_max[3]
// (18 doubles). tri: 9*ntri double vertices (may be NULL when ntri == 0). pos/vel are updated in
// place (n x 3).
// opts bit 0 (fast, long test runs): the mesh loop visits only the triangles a triangle-grid
// index offers and the pair loop runs on nthreads threads; both give the results of the
// reference's loops. Without it: every triangle and a serial pair loop, as the reference runs
// them. opts bit 1: the reference CUDA backend's step instead of the CPU backend's.
int oracle_run_f64(const double* params, int64_t n, double* pos, double* vel, const double* r,
                   const double* m, int64_t ntri, const double* tri, int64_t nsteps,
                   int nthreads, int opts) {
  return run<double>(params, n, pos, vel, r, m, ntri, tri, nsteps, nthreads, opts);
}

int oracle_run_f32(const double* params, int64_t n, float* pos, float* vel, const float* r,
                   const float* m, int64_t ntri, const double* tri, int64_t nsteps, int nthreads,
                   int opts) {
  return run<float>(params, n, pos, vel, r, m, ntri, tri, nsteps, nthreads, opts);
}

int oracle_bucketing_f64(double rmax, const double* mn, const double* mx, int64_t n,
                         const double* pos, int32_t* cell, uint32_t* count, uint32_t* offset,
                         int32_t* perm) {
  return bucketing<double>(rmax, mn, mx, n, pos, cell, count, offset, perm);
}

int oracle_bucketing_f32(double rmax, const double* mn, const double* mx, int64_t n,
                         const float* pos, int32_t* cell, uint32_t* count, uint32_t* offset,
                         int32_t* perm) {
  return bucketing<float>(rmax, mn, mx, n, pos, cell, count, offset, perm);
}

// getClosestPointOnPolygon for one triangle (normal from the reference's formula).
void oracle_closest_point_f64(const double* a, const double* b, const double* c, const double* p,
                              double* q) {
  V3<double> A{a[0], a[1], a[2]}, B{b[0], b[1], b[2]}, C{c[0], c[1], c[2]},
      P{p[0], p[1], p[2]};
  V3<double> N = normalize(cross(B - A, C - A));
  V3<double> Q = closest_on_triangle(A, B, C, N, P);
  q[0] = Q.x;
  q[1] = Q.y;
  q[2] = Q.z;
}

// getClosestPointOnPolygonKernelFunc for one triangle (normal cross / length, the CUDA backend's).
void oracle_closest_point_cuda_f64(const double* a, const double* b, const double* c,
                                   const double* p, double* q) {
  V3<double> A{a[0], a[1], a[2]}, B{b[0], b[1], b[2]}, C{c[0], c[1], c[2]},
      P{p[0], p[1], p[2]};
  const V3<double> cr = cross(B - A, C - A);
  V3<double> N = cr / length(cr);
  V3<double> Q = closest_on_triangle_cuda(A, B, C, N, P);
  q[0] = Q.x;
  q[1] = Q.y;
  q[2] = Q.z;
}

// Triangle normals (PolygonContainer.cpp:52-68) in double.
void oracle_mesh_normals_f64(int64_t ntri, const double* tri, double* normals) {
  for (int64_t t = 0; t < ntri; ++t) {
    const double* p = tri + 9 * t;
    V3<double> a{p[0], p[1], p[2]}, b{p[3], p[4], p[5]}, c{p[6], p[7], p[8]};
    V3<double> nn = normalize(cross(b - a, c - a));
    normals[3 * t] = nn.x;
    normals[3 * t + 1] = nn.y;
    normals[3 * t + 2] = nn.z;
  }
}

// ParticleStatistics::calculateStatistics (ParticleStatistics.cpp:3-43) over n particles in index
// order: out = {kinetic energy, min x, y, z, max x, y, z, collided count, max overlap}. The
// kinetic energy is the sequential sum of 0.5 * m * dot(v, v); the AABB uses glm::min / glm::max
// ((b < a) ? b : a and (a < b) ? b : a) starting from particle 0; particle i counts as collided
// when some j > i has r_i + r_j - distance(x_i, x_j) > 0, distance = length(x_j - x_i).
void oracle_statistics(int64_t n, const double* pos, const double* vel, const double* radius,
                       const double* mass, double* out) {
  double ke = 0.0;
  double mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = pos[a] - radius[0];
    mx[a] = pos[a] + radius[0];
  }
  for (int64_t i = 0; i < n; ++i) {
    const V3<double> v{vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    ke += 0.5 * mass[i] * dot(v, v);
    for (int a = 0; a < 3; ++a) {
      const double lo = pos[3 * i + a] - radius[i], hi = pos[3 * i + a] + radius[i];
      mn[a] = (lo < mn[a]) ? lo : mn[a];
      mx[a] = (mx[a] < hi) ? hi : mx[a];
    }
  }
  int64_t collided = 0;
  double maxov = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const V3<double> xi{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    bool hit = false;
    for (int64_t j = i + 1; j < n; ++j) {
      const V3<double> xj{pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
      const double dist = length(xj - xi);
      const double cd = radius[i] + radius[j] - dist;
      if (cd > 0.0) {
        hit = true;
        maxov = std::max(maxov, cd);
      }
    }
    if (hit) ++collided;
  }
  out[0] = ke;
  for (int a = 0; a < 3; ++a) {
    out[1 + a] = mn[a];
    out[4 + a] = mx[a];
  }
  out[7] = (double)collided;
  out[8] = maxov;
}

// ParticleModel::resolveOverlaps (ParticleModel.cpp:154-252), literally: rank by the coordinate
// along `axis` with std::sort; for every rank i, the partner scan starts at the first rank j < i
// whose top (coordinate + radius) reaches the bottom of particle i (rank 0 when none does); each
// overlapping partner lifts particle i to x_j + sqrt((R^2 - d0^2) - d1^2) along `axis`; a lift
// whose new coordinate falls strictly between the coordinates at ranks k-1 and k (first such
// k > i) re-sorts ranks [i, k] and processes rank i again. Then the O(N^2) overlap check:
// out = {collided count, max overlap}.
void oracle_resolve_overlaps(int64_t n, double* pos, const double* radius, int32_t axis,
                             double* out) {
  std::vector<int64_t> rank((size_t)n);
  for (int64_t i = 0; i < n; ++i) rank[(size_t)i] = i;
  auto less = [&](int64_t a, int64_t b) { return pos[3 * a + axis] < pos[3 * b + axis]; };
  std::sort(rank.begin(), rank.end(), less);
  int other[2], k = 0;
  for (int a = 0; a < 3; ++a)
    if (a != axis) other[k++] = a;
  for (int64_t i = 0; i < n; ++i) {
    double* pi = pos + 3 * rank[(size_t)i];
    const double ri = radius[rank[(size_t)i]];
    int64_t first = 0;
    for (int64_t j = 0; j < i; ++j) {
      const int64_t q = rank[(size_t)j];
      if (pi[axis] - ri <= pos[3 * q + axis] + radius[q]) {
        first = j;
        break;
      }
    }
    for (int64_t j = first; j < i; ++j) {
      const double* pj = pos + 3 * rank[(size_t)j];
      const double sum = ri + radius[rank[(size_t)j]];
      const double dx = pi[0] - pj[0], dy = pi[1] - pj[1], dz = pi[2] - pj[2];
      if (!(sum * sum - ((dx * dx + dy * dy) + dz * dz) > 0.0)) continue;
      const double e0 = pi[other[0]] - pj[other[0]], e1 = pi[other[1]] - pj[other[1]];
      pi[axis] = pj[axis] + std::sqrt((sum * sum - e0 * e0) - e1 * e1);
      int64_t last = i;
      for (int64_t m = i + 1; m < n; ++m)
        if (pos[3 * rank[(size_t)(m - 1)] + axis] < pi[axis] && pi[axis] < pos[3 * rank[(size_t)m] + axis]) {
          last = m;
          break;
        }
      if (last != i) {
        std::sort(rank.begin() + i, rank.begin() + last + 1, less);
        --i;
        break;
      }
    }
  }
  int64_t collided = 0;
  double worst = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    bool hit = false;
    for (int64_t j = i + 1; j < n; ++j) {
      const double dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1],
                   dz = pos[3 * i + 2] - pos[3 * j + 2];
      const double cd = radius[i] + radius[j] - std::sqrt((dx * dx + dy * dy) + dz * dz);
      if (cd > 0.0) {
        hit = true;
        worst = std::max(worst, cd);
      }
    }
    collided += hit;
  }
  out[0] = (double)collided;
  out[1] = worst;
}

// Mass of ParticleData::readFromCSV (ParticleIO.cpp:131).
double oracle_mass(double density, double radius) {
  return density * 4.0 / 3.0 * M_PI * std::pow(radius, 3);
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
