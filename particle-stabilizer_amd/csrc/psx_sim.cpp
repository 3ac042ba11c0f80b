// psx_sim.cpp -- host side of libparticle_stabilizer_hip.so: scene set-up (grid geometry, triangle
// grid), device buffers, the step sequence, hipGraph replay, per-kernel event timing and the C ABI
// declared in include/psx.h.
//
// A single simulation is a batch of one scene: every kernel takes a per-workgroup (scene, slice)
// table, so plain and batched runs share the same launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "psx.h"
#include "psx_adjoint.hpp"
#include "psx_kernels.hpp"
#include "psx_host.hpp"

namespace psx {

// ---------------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------------
static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PSX_HIP(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw Error(PSX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

// ---------------------------------------------------------------------------------------------
// grid geometry (ParticleBucket::initBucket, ParticleBucket.cpp:17-27; with cuda,
// ParticleModelCUDA::initBuckets, ParticleModelCUDA.cpp:62-99, whose width uses an exact 1.05)
// ---------------------------------------------------------------------------------------------
HostGrid make_grid(double rmax, const double* mn, const double* mx, bool cuda) {
  HostGrid g;
  const double coeff = (double)1.05f;  // COEFF_FOR_INTERVAL_OF_BUCKETS, ParticleBucket.hpp:17
  g.h = cuda ? rmax * 2.0 * 1.05 : rmax * 2.0 * coeff;
  if (!(g.h > 0.0) || !std::isfinite(g.h))
    throw Error(PSX_ERR_INVALID, "largest radius must be positive and finite");
  for (int a = 0; a < 3; ++a) {
    g.min[a] = mn[a];
    g.max[a] = mx[a];
    if (!(mx[a] > mn[a]) || !std::isfinite(mn[a]) || !std::isfinite(mx[a]))
      throw Error(PSX_ERR_INVALID, "container bounds must satisfy min < max on every axis");
    const double q = (mx[a] - mn[a]) / g.h;
    if (!(q < 1.0e9)) throw Error(PSX_ERR_INVALID, "grid too large: container / cell size");
    g.n[a] = (int64_t)q + 1;
  }
  return g;
}

// ---------------------------------------------------------------------------------------------
// triangle grid (replaces the O(N*T) candidate loop of PolygonContainer.cpp:244-278)
// ---------------------------------------------------------------------------------------------
namespace {

struct D3 {
  double x, y, z;
};
inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 scale(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dotd(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Distance from p to triangle abc by region classification (Ericson, Real-Time Collision Detection,
// 5.1.5). Used only to decide which triangles a grid cell lists, with a safety margin.
double point_triangle_distance(D3 p, D3 a, D3 b, D3 c) {
  const D3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  const double d1 = dotd(ab, ap), d2 = dotd(ac, ap);
  D3 q;
  if (d1 <= 0 && d2 <= 0) {
    q = a;
  } else {
    const D3 bp = sub(p, b);
    const double d3 = dotd(ab, bp), d4 = dotd(ac, bp);
    if (d3 >= 0 && d4 <= d3) {
      q = b;
    } else {
      const double vc = d1 * d4 - d3 * d2;
      if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        q = add(a, scale(ab, d1 / (d1 - d3)));
      } else {
        const D3 cp = sub(p, c);
        const double d5 = dotd(ab, cp), d6 = dotd(ac, cp);
        if (d6 >= 0 && d5 <= d6) {
          q = c;
        } else {
          const double vb = d5 * d2 - d1 * d6;
          if (vb <= 0 && d2 >= 0 && d6 <= 0) {
            q = add(a, scale(ac, d2 / (d2 - d6)));
          } else {
            const double va = d3 * d6 - d5 * d4;
            if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
              q = add(b, scale(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
            } else {
              const double den = va + vb + vc;
              if (!(std::fabs(den) > 0)) {  // degenerate triangle: nearest of the three edges
                auto seg = [&](D3 s0, D3 s1) {
                  const D3 e = sub(s1, s0);
                  const double ee = dotd(e, e);
                  double t = ee > 0 ? dotd(sub(p, s0), e) / ee : 0.0;
                  t = std::min(1.0, std::max(0.0, t));
                  const D3 d = sub(p, add(s0, scale(e, t)));
                  return std::sqrt(dotd(d, d));
                };
                return std::min(seg(a, b), std::min(seg(b, c), seg(c, a)));
              }
              const double v = vb / den, w = vc / den;
              q = add(a, add(scale(ab, v), scale(ac, w)));
            }
          }
        }
      }
    }
  }
  const D3 d = sub(p, q);
  return std::sqrt(dotd(d, d));
}

}  // namespace

// Builds the triangle lists of one mesh over the scene's particle grid g (whose domain is the mesh's
// bounding box, PhysicsEngine.cpp:112). Cell k lists triangle t (ascending) when a sphere of radius
// <= rmax hashed to cell k can touch t. The hash clamps positions outside the domain into the
// boundary cells, but a sphere farther than rmax from the domain cannot touch any triangle, so the
// boundary cells' boxes are extended outward by rmax + margin only. A cell keeps t when
//   distance(box centre, t) <= half diagonal of the box + rmax + margin,
// a superset of the triangles within rmax of the box. margin absorbs the rounding of the hash
// division and of the closest-point evaluation in T.
void build_tri_grid(MeshHost& m, const HostGrid& g, double rmax, bool fp32) {
  const int64_t nt = m.nt;
  double lmax = 0.0;
  for (int a = 0; a < 3; ++a) lmax = std::max(lmax, std::max(std::fabs(m.min[a]), std::fabs(m.max[a])));
  const double margin = fp32 ? 1e-4 * (lmax + 1.0) + 1e-3 * rmax : 1e-9 * (lmax + 1.0);
  const double reach = rmax + margin;
  const int64_t ncell = g.n[0] * g.n[1] * g.n[2];
  std::vector<std::vector<int32_t>> lists((size_t)ncell);
  auto box_lo = [&](int64_t i, int a) {
    return i == 0 ? g.min[a] - reach : g.min[a] + (double)i * g.h;
  };
  auto box_hi = [&](int64_t i, int a) {
    return i == g.n[a] - 1 ? g.min[a] + (double)g.n[a] * g.h + reach : g.min[a] + (double)(i + 1) * g.h;
  };
  for (int64_t t = 0; t < nt; ++t) {
    const double* p = &m.v[9 * t];
    const D3 A{p[0], p[1], p[2]}, B{p[3], p[4], p[5]}, C{p[6], p[7], p[8]};
    int64_t i0[3], i1[3];
    for (int a = 0; a < 3; ++a) {
      const double lo = std::min(p[a], std::min(p[3 + a], p[6 + a]));
      const double hi = std::max(p[a], std::max(p[3 + a], p[6 + a]));
      i0[a] = axis_index(lo - reach, g.min[a], g.h, (int32_t)g.n[a]);
      i1[a] = axis_index(hi + reach, g.min[a], g.h, (int32_t)g.n[a]);
    }
    for (int64_t ix = i0[0]; ix <= i1[0]; ++ix)
      for (int64_t iy = i0[1]; iy <= i1[1]; ++iy)
        for (int64_t iz = i0[2]; iz <= i1[2]; ++iz) {
          const int64_t ii[3] = {ix, iy, iz};
          double ctr[3], hd2 = 0.0;
          for (int a = 0; a < 3; ++a) {
            const double lo = box_lo(ii[a], a), hi = box_hi(ii[a], a);
            ctr[a] = 0.5 * (lo + hi);
            hd2 += 0.25 * (hi - lo) * (hi - lo);
          }
          if (point_triangle_distance(D3{ctr[0], ctr[1], ctr[2]}, A, B, C) <= std::sqrt(hd2) + reach)
            lists[(size_t)((ix * g.n[1] + iy) * g.n[2] + iz)].push_back((int32_t)t);
        }
  }
  // the lists' offsets are int32 on the device
  size_t entries = 0;
  for (const auto& l : lists) entries += l.size();
  if (entries > (size_t)INT32_MAX)
    throw Error(PSX_ERR_INVALID, "triangle grid exceeds 2^31 cell entries (mesh too fine for the grid)");
  m.off.assign((size_t)ncell + 1, 0);
  m.list.clear();
  m.list.reserve(entries);
  for (int64_t c = 0; c < ncell; ++c) {
    m.off[(size_t)c] = (int32_t)m.list.size();
    m.list.insert(m.list.end(), lists[(size_t)c].begin(), lists[(size_t)c].end());
  }
  m.off[(size_t)ncell] = (int32_t)m.list.size();
}

// The per-cell flag byte of the scene's grid. kCellCull: the cell's 27 neighbours may be culled by
// distance in the pair kernel. A cell is bounded when every member's snapshot position lies inside
// its box: interior cells (index 1 .. n-2 on every axis; positions outside the domain are clamped
// into the boundary cells, and box walls only push particles of boundary cells) with an empty
// triangle list (a particle pushed by the mesh touched a triangle, which its bucket cell therefore
// lists). kCellCull marks cells whose 3x3x3 neighbourhood is bounded: the bounded mask eroded along
// x, y, z. kCellTris: the cell lists at least one triangle (the rank/gather kernel reads the list
// bounds of those cells only).
std::vector<uint8_t> cull_cells(const HostGrid& g, const MeshHost* m) {
  const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const size_t nc = (size_t)(nx * ny * nz);
  std::vector<uint8_t> a(nc, 0), b(nc, 0);
  auto at = [&](int64_t i, int64_t j, int64_t k) { return (size_t)((i * ny + j) * nz + k); };
  for (int64_t i = 1; i + 1 < nx; ++i)
    for (int64_t j = 1; j + 1 < ny; ++j)
      for (int64_t k = 1; k + 1 < nz; ++k) {
        const size_t c = at(i, j, k);
        a[c] = (!m || m->off[c + 1] == m->off[c]) ? 1 : 0;
      }
  // erosion: a cell survives when it and both neighbours along the axis are set (none at the edges)
  auto erode = [&](const std::vector<uint8_t>& src, std::vector<uint8_t>& dst, int axis) {
    std::fill(dst.begin(), dst.end(), 0);
    for (int64_t i = 0; i < nx; ++i)
      for (int64_t j = 0; j < ny; ++j)
        for (int64_t k = 0; k < nz; ++k) {
          const int64_t idx[3] = {i, j, k};
          const int64_t len = g.n[axis];
          if (idx[axis] == 0 || idx[axis] == len - 1) continue;
          int64_t lo[3] = {i, j, k}, hi[3] = {i, j, k};
          --lo[axis];
          ++hi[axis];
          const size_t c = at(i, j, k);
          dst[c] = src[c] & src[at(lo[0], lo[1], lo[2])] & src[at(hi[0], hi[1], hi[2])];
        }
  };
  erode(a, b, 0);
  erode(b, a, 1);
  erode(a, b, 2);
  if (m)
    for (size_t c = 0; c < nc; ++c)
      if (m->off[c + 1] != m->off[c]) b[c] |= kCellTris;
  return b;
}

// ---------------------------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t b) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    bytes = std::max<size_t>(b, 256);
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(PSX_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(bytes) +
                                     " bytes) failed: " + hipGetErrorString(e));
    }
  }
  template <class U>
  U* as() const {
    return static_cast<U*>(p);
  }
};

// Particles per workgroup of the statistics kernel (its table is built at set-up).
constexpr int kStatBlock = 256;

static const char* kKernelNames[kNumKernels] = {"k_integrate", "k_scan", "k_rank_gather",
                                                "k_pairs", "k_scatter_ids"};

// ---------------------------------------------------------------------------------------------
// the simulation
// ---------------------------------------------------------------------------------------------
template <class T>
struct SimImpl final : SimBase {
  // particle state, double-buffered; pos/id alternate every step, vel alternates only when a
  // bucketing pass runs without the pair kernel
  DevBuf pos[2], vel[2], id[2];
  DevBuf cell_of, cell_ids, ovf, ovf_ctr, home, list, stat_part;
  DevBuf tile_count, scan_gen;  // direct scan: per-tile counts (two generations), generation
  int32_t ntiles_direct = 0;    // tiles of the direct scan, 0: look-back scan
  DevBuf count, offset, tile_state, flags;
  // incremental bucketing (sort_variant 3): per-cell arrival words and change marks, the second
  // offset array (the offsets alternate with the position buffers), the bucket cells of
  // container-moved spheres and the kept per-tile totals of the counts (two generations)
  DevBuf arr, mark, offset2, cell_cur, tile_tot, rank8[2];
  DevBuf scenes_d, tgeo_d, tbox_d, tent_d, tcell_off_d, cull_d;
  std::vector<SceneDev<T>> scenes_h;  // host copy of scenes_d
  DevBuf io;  // 6n T staging for state transfer
  std::map<int, std::unique_ptr<DevBuf>> btab;
  std::map<int, int> btab_blocks;
  int cur_pos = 0, cur_vel = 0;
  // Runs of steps replay 64-step graphs, then 16-step ones, then single steps (both lengths are even,
  // so a chunk returns to the same buffer parity). Measured on MI355X: 64-step graphs take the
  // 262,144-sphere step from 24.4 to 24.0 us (fewer graph launches); the 16-step graphs keep short
  // runs on graph replay.
  static constexpr int kGraphSteps[2] = {64, 16};
  hipGraphExec_t graphs[2][2][2] = {};  // [length][position buffer][velocity buffer]
  // Graphs of the 1-15 steps a run leaves after its 64- and 16-step chunks, captured by prepare()
  // only (a run that was not prepared takes them as single steps): [steps][position][velocity]. An
  // odd-length graph ends on the other position buffer.
  hipGraphExec_t tail_graphs[16][2][2] = {};

  ~SimImpl() override {
    destroy_graphs();
    if (own_stream) (void)hipStreamDestroy(own_stream);
  }

  void destroy_graphs() {
    destroy_window_graphs();
    for (auto& len : graphs)
      for (auto& row : len)
        for (auto& g : row)
          if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
          }
    for (auto& len : tail_graphs)
      for (auto& row : len)
        for (auto& g : row)
          if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
          }
  }

  int32_t dtype() const override { return sizeof(T) == 4 ? PSX_F32 : PSX_F64; }

  const BlockInfo* table(int block, int* nblocks) {
    auto it = btab.find(block);
    if (it == btab.end()) {
      std::vector<BlockInfo> t;
      for (size_t s = 0; s < scenes.size(); ++s) {
        const int64_t b = scenes[s].p_begin, e = b + scenes[s].n;
        for (int64_t p = b; p < e; p += block)
          t.push_back(BlockInfo{(int32_t)s, (int32_t)p, (int32_t)std::min<int64_t>(p + block, e), 0});
      }
      auto buf = std::make_unique<DevBuf>();
      buf->alloc(t.size() * sizeof(BlockInfo));
      PSX_HIP(hipMemcpy(buf->p, t.data(), t.size() * sizeof(BlockInfo), hipMemcpyHostToDevice));
      btab_blocks[block] = (int)t.size();
      it = btab.emplace(block, std::move(buf)).first;
    }
    *nblocks = btab_blocks[block];
    return it->second->as<BlockInfo>();
  }

  void setup(const std::vector<double>& hpos, const std::vector<double>& hvel,
             const std::vector<double>& hr, const std::vector<double>& hm) override {
    const int64_t nn = n;
    // particle arrays in the caller's order (which already groups scenes contiguously)
    std::vector<Vec4<T>> p4((size_t)nn), v4((size_t)nn);
    std::vector<int32_t> ids((size_t)nn);
    for (int64_t i = 0; i < nn; ++i) {
      p4[i] = Vec4<T>{(T)hpos[3 * i], (T)hpos[3 * i + 1], (T)hpos[3 * i + 2], (T)hr[i]};
      v4[i] = Vec4<T>{(T)hvel[3 * i], (T)hvel[3 * i + 1], (T)hvel[3 * i + 2], (T)hm[i]};
      ids[i] = (int32_t)i;
    }
    // scenes: r_max in T, grid, cell bases (ParticleModel::initBuckets, ParticleModel.cpp:28-47)
    int64_t cells = 0;
    std::vector<SceneDev<T>> sd(scenes.size());
    std::vector<TriGeo<T>> tgeo;
    std::vector<TriBox<T>> tbox;
    std::vector<TriEntry<T>> tent;
    std::vector<int32_t> toff;
    std::vector<uint8_t> cullh;  // one byte per cell over all scenes
    for (size_t s = 0; s < scenes.size(); ++s) {
      SceneHost& sh = scenes[s];
      double rmax = 0.0;
      for (int64_t i = sh.p_begin; i < sh.p_begin + sh.n; ++i)
        rmax = std::max(rmax, (double)(T)hr[(size_t)i]);
      sh.rmax = rmax;
      double mn[3], mx[3];
      if (sh.mesh >= 0) {
        const MeshHost& m = meshes[(size_t)sh.mesh];
        for (int a = 0; a < 3; ++a) {
          mn[a] = m.min[a];
          mx[a] = m.max[a];
        }
      } else {
        for (int a = 0; a < 3; ++a) {
          mn[a] = sh.prm.box_min[a];
          mx[a] = sh.prm.box_max[a];
        }
      }
      sh.grid = make_grid(rmax, mn, mx, cuda);
      sh.cell_base = cells;
      const int64_t c = sh.grid.n[0] * sh.grid.n[1] * sh.grid.n[2];
      if (c > ((int64_t)1 << 30)) throw Error(PSX_ERR_INVALID, "grid of a scene exceeds 2^30 cells");
      cells += c;
      if (cells > ((int64_t)1 << 30)) throw Error(PSX_ERR_INVALID, "total grid exceeds 2^30 cells");

      SceneDev<T>& d = sd[s];
      std::memset(&d, 0, sizeof(d));
      for (int a = 0; a < 3; ++a) {
        d.grid.min[a] = sh.grid.min[a];
        d.grid.n[a] = (int32_t)sh.grid.n[a];
      }
      d.grid.h = sh.grid.h;
      d.grid.cell_base = (int32_t)sh.cell_base;
      for (int a = 0; a < 2; ++a) {  // divisors n_z, n_y (Granlund and Montgomery, fig. 4.1)
        const uint64_t dv = (uint64_t)sh.grid.n[2 - a];
        uint32_t l = 0;
        while ((uint64_t(1) << l) < dv) ++l;
        const uint64_t m = ((uint64_t(1) << 32) * ((uint64_t(1) << l) - dv)) / dv + 1;
        d.grid.magic[a] = (uint32_t)m;
        d.grid.shift[a] = std::min<uint32_t>(l, 1u) | (std::max<uint32_t>(l, 1u) - 1u) << 8;
      }
      fill_coefficients(d, sh.prm);
      d.rmax = (T)rmax;
      d.p_begin = (int32_t)sh.p_begin;
      d.p_end = (int32_t)(sh.p_begin + sh.n);
      d.mesh = -1;
      d.tcell_base = 0;
      d.cuda = cuda ? 1 : 0;
      if (sh.mesh >= 0) build_tri_grid(meshes[(size_t)sh.mesh], sh.grid, rmax, sizeof(T) == 4);
      {
        const std::vector<uint8_t> cs =
            cull_cells(sh.grid, sh.mesh >= 0 ? &meshes[(size_t)sh.mesh] : nullptr);
        cullh.insert(cullh.end(), cs.begin(), cs.end());
      }
      if (sh.mesh >= 0) {
        MeshHost& m = meshes[(size_t)sh.mesh];
        d.mesh = sh.mesh;
        d.tcell_base = (int32_t)toff.size();
        const int32_t tri_base = (int32_t)tgeo.size();
        // triangle records (PolygonContainer.cpp:52-89): normal in double, stored in T; AABB over
        // the T corners (glm::min / glm::max). With cuda the normal is cross / length, component by
        // component (PolygonContainerCUDA.cpp:64-82); the cross product is the same.
        std::vector<TriBox<T>> boxes((size_t)m.nt);
        for (int64_t t = 0; t < m.nt; ++t) {
          const double* q = &m.v[(size_t)(9 * t)];
          const D3 A{q[0], q[1], q[2]}, B{q[3], q[4], q[5]}, C{q[6], q[7], q[8]};
          const D3 e1 = sub(B, A), e2 = sub(C, A);
          // glm::cross(x, y) and glm::normalize(v) = v * (1 / sqrt(dot(v, v)))
          const D3 cr{e1.y * e2.z - e2.y * e1.z, e1.z * e2.x - e2.z * e1.x,
                      e1.x * e2.y - e2.x * e1.y};
          const double len = std::sqrt((cr.x * cr.x + cr.y * cr.y) + cr.z * cr.z);
          const double inv = 1.0 / len;
          double nrm[3] = {cr.x * inv, cr.y * inv, cr.z * inv};
          if (cuda) {
            nrm[0] = cr.x / len;
            nrm[1] = cr.y / len;
            nrm[2] = cr.z / len;
          }
          TriGeo<T> tg;
          TriBox<T> tb;
          std::memset(&tb, 0, sizeof(tb));
          for (int a = 0; a < 3; ++a) {
            tg.a[a] = (T)q[a];
            tg.b[a] = (T)q[3 + a];
            tg.c[a] = (T)q[6 + a];
            tg.n[a] = (T)nrm[a];
            T lo = (tg.b[a] < tg.a[a]) ? tg.b[a] : tg.a[a];
            lo = (tg.c[a] < lo) ? tg.c[a] : lo;
            T hi = (tg.a[a] < tg.b[a]) ? tg.b[a] : tg.a[a];
            hi = (hi < tg.c[a]) ? tg.c[a] : hi;
            tb.lo[a] = lo;
            tb.hi[a] = hi;
          }
          tb.tri = tri_base + (int32_t)t;
          tgeo.push_back(tg);
          boxes[(size_t)t] = tb;
        }
        if (tbox.size() + m.list.size() > (size_t)INT32_MAX)
          throw Error(PSX_ERR_INVALID, "the scenes' triangle grids exceed 2^31 cell entries in all");
        const int32_t entry_base = (int32_t)tbox.size();
        for (size_t c2 = 0; c2 + 1 < m.off.size(); ++c2) toff.push_back(entry_base + m.off[c2]);
        for (int32_t t : m.list) {
          const TriBox<T>& bx = boxes[(size_t)t];
          tbox.push_back(bx);
          TriEntry<T> te;
          std::memset(&te, 0, sizeof(te));
          te.g = tgeo[(size_t)bx.tri];
          for (int a = 0; a < 3; ++a) {
            te.lo[a] = bx.lo[a];
            te.hi[a] = bx.hi[a];
          }
          tent.push_back(te);
        }
        toff.push_back((int32_t)tbox.size());  // end of this scene's last cell
      }
    }
    ncells = cells;
    cpad = (cells + kScanTile - 1) / kScanTile * kScanTile;
    ntiles = (int32_t)(cpad / kScanTile);
    id_stride = (cells + kIdPlane - 1) / kIdPlane * kIdPlane;
    ntiles_direct = cpad / kDirectTile <= kDirectTiles ? (int32_t)(cpad / kDirectTile) : 0;

    // device allocations
    for (int k = 0; k < 2; ++k) {
      pos[k].alloc(((size_t)nn + kPairPad) * sizeof(Vec4<T>));
      vel[k].alloc(((size_t)nn + kPairPad) * sizeof(Vec4<T>));
      id[k].alloc((size_t)nn * sizeof(int32_t));
    }
    cell_of.alloc((size_t)nn * 4);
    home.alloc((size_t)nn * 4);
    list.alloc(((size_t)nn + kListPad) * 4);
    cell_ids.alloc((size_t)id_stride * (kCellIds + 1) * 4);  // id slots + overflow chain heads
    ovf.alloc((size_t)nn * 8);
    ovf_ctr.alloc(16);  // overflow entries, their published number, the bucketing generation
    count.alloc((size_t)cpad * 4);
    offset.alloc(((size_t)cpad + 1 + kOffsetPad) * 4);
    tile_state.alloc((size_t)ntiles * 8);
    tile_count.alloc(std::max<size_t>(2 * (size_t)ntiles_direct * kTileStride, 1) * 4);
    scan_gen.alloc(4);
    flags.alloc(16);
    io.alloc((size_t)nn * 6 * sizeof(T));
    scenes_d.alloc(sd.size() * sizeof(SceneDev<T>));
    tgeo_d.alloc(std::max<size_t>(tgeo.size(), 1) * sizeof(TriGeo<T>));
    tbox_d.alloc(std::max<size_t>(tbox.size(), 1) * sizeof(TriBox<T>));
    tent_d.alloc((tent.size() + kTriPad) * sizeof(TriEntry<T>));
    cull_d.alloc((size_t)cpad);
    tcell_off_d.alloc(std::max<size_t>(toff.size(), 1) * 4);

    PSX_HIP(hipMemcpy(pos[0].p, p4.data(), (size_t)nn * sizeof(Vec4<T>), hipMemcpyHostToDevice));
    PSX_HIP(hipMemcpy(vel[0].p, v4.data(), (size_t)nn * sizeof(Vec4<T>), hipMemcpyHostToDevice));
    PSX_HIP(hipMemcpy(id[0].p, ids.data(), (size_t)nn * 4, hipMemcpyHostToDevice));
    PSX_HIP(hipMemset(count.p, 0, count.bytes));
    PSX_HIP(hipMemset(offset.p, 0, offset.bytes));
    PSX_HIP(hipMemset(tile_state.p, 0, tile_state.bytes));
    PSX_HIP(hipMemset(tile_count.p, 0, tile_count.bytes));
    PSX_HIP(hipMemset(scan_gen.p, 0, scan_gen.bytes));
    PSX_HIP(hipMemset(ovf_ctr.p, 0, ovf_ctr.bytes));
    PSX_HIP(hipMemset(cell_ids.p, 0, cell_ids.bytes));
    PSX_HIP(hipMemset(flags.p, 0, flags.bytes));
    PSX_HIP(hipMemcpy(scenes_d.p, sd.data(), sd.size() * sizeof(SceneDev<T>), hipMemcpyHostToDevice));
    scenes_h = sd;
    if (!tgeo.empty())
      PSX_HIP(hipMemcpy(tgeo_d.p, tgeo.data(), tgeo.size() * sizeof(TriGeo<T>), hipMemcpyHostToDevice));
    if (!tbox.empty())
      PSX_HIP(hipMemcpy(tbox_d.p, tbox.data(), tbox.size() * sizeof(TriBox<T>), hipMemcpyHostToDevice));
    PSX_HIP(hipMemset(tent_d.p, 0, tent_d.bytes));  // the padding entries are read, never used
    PSX_HIP(hipMemset(cull_d.p, 0, cull_d.bytes));
    PSX_HIP(hipMemcpy(cull_d.p, cullh.data(), cullh.size(), hipMemcpyHostToDevice));
    if (!tent.empty())
      PSX_HIP(hipMemcpy(tent_d.p, tent.data(), tent.size() * sizeof(TriEntry<T>), hipMemcpyHostToDevice));
    if (!toff.empty())
      PSX_HIP(hipMemcpy(tcell_off_d.p, toff.data(), toff.size() * 4, hipMemcpyHostToDevice));
    tri_count = (int64_t)tgeo.size();
    tcell_entries = (int64_t)tbox.size();
    int nb;
    table(tuning.block_integrate, &nb);
    table(tuning.block_sort, &nb);
    table(pair_particles_per_block(), &nb);
    table(kStatBlock, &nb);
    inc_buffers();
    PSX_HIP(hipDeviceSynchronize());
  }

  // Incremental bucketing (sort_variant 3): kept counts and tile totals, atomics only for spheres
  // that change cell (NextBucket, IncRank), and the scan from the kept totals (k_scan_tiles) in place
  // of the direct or look-back scan, whose bookkeeping every hash pass rebuilds; the full form's
  // bucketings of the same handle (statistics, windows) keep the grid's scan. Every change between
  // the two goes through drop_hash, which clears the direct scan's per-tile counts. The CUDA
  // backend's step uses the id slots of sort_variant 1.
  bool incremental() const { return tuning.sort_variant == 3 && !cuda; }

  // The buffers of the incremental bucketing, allocated when the tuning asks for it (at set-up or
  // retune, outside any step or capture) and zeroed: arrival words of generation 0 are older than
  // every bucketing.
  void inc_buffers() {
    if (!incremental() || arr.p) return;
    arr.alloc((size_t)cpad * 8);
    mark.alloc((size_t)cpad);
    offset2.alloc(offset.bytes);
    cell_cur.alloc((size_t)n * 4);
    tile_tot.alloc((size_t)2 * inc_tiles() * 4);
    for (DevBuf& r : rank8) r.alloc((size_t)n);
    PSX_HIP(hipMemset(arr.p, 0, arr.bytes));
    PSX_HIP(hipMemset(mark.p, 0, mark.bytes));
    PSX_HIP(hipMemset(offset2.p, 0, offset2.bytes));
    PSX_HIP(hipDeviceSynchronize());
  }

  // Tiles of the incremental bucketing's scan: kDirectTile cells on grids the direct scan covers,
  // 4,096 cells on grids of at most 4,096 such tiles, kScanTile cells on larger ones (the most
  // workgroups for which a workgroup's sum of the kept totals before its tile stays short; measured
  // on MI355X, the 11.5M-cell hopper grid scans in 19.5 us with 4,096-cell tiles, 21.5 with 16,384).
  int32_t inc_shift() const {
    if (ntiles_direct) return kDirectShift;
    return (cpad >> (kDirectShift + 2)) <= 4096 ? kDirectShift + 2 : kScanShift;
  }
  int32_t inc_tiles() const { return (int32_t)(cpad >> inc_shift()); }

  TileTotals tile_totals() const {
    TileTotals t;
    t.tot = tile_tot.as<uint32_t>();
    t.gen = scan_gen.as<uint32_t>();
    t.ntiles = inc_tiles();
    t.shift = inc_shift();
    return t;
  }

  // Offsets of the incremental bucketing's state in position buffer k (k = 0: the offsets every
  // other path uses).
  uint32_t* offs_at(int k) const {
    return k == 0 ? offs() : offset2.as<uint32_t>() + kOffsetLead;
  }

  // One kernel launch, optionally bracketed by timing events.
  // One kernel launch; when profiling, its own dispatch timestamps go to the profiler's events.
  template <class F>
  void launch(int k, Profiler* prof, F&& f) {
    if (prof) set_launch_events(prof->begin(k), prof->end(k));
    const hipError_t e = f();
    set_launch_events(nullptr, nullptr);
    PSX_HIP(e);
  }

  // Bucketing: k_integrate hashes the positions of (ps, vs) after the integration step (when `mode`
  // says so) and files the ids `is` into the cells' id slots; scan into `off`; k_rank_gather ranks,
  // integrates, applies the container (kModeContainer) and the check (kModeCheck) to (ps, vs) and
  // gathers into (pd, vd, idd).
  // Cell lists of the bucketing: per-cell id slots, slot-major (no scatter kernel; consecutive
  // spheres fill neighbouring words) or dense lists (a scatter kernel, the hash pass keeps only the
  // counts). Measured on MI355X, the id slots are faster on all four standard workloads (262k to
  // 4.6M spheres, 4e5 to 3.5e7 cells), so they are the automatic choice.
  bool id_blocks() const { return tuning.sort_variant != 2; }

  // The direct scan's per-tile counts for a hash pass; none for the incremental bucketing's passes,
  // whose scan keeps tile totals of its own (TileTotals).
  TileCounts tiles(bool inc = false) const {
    TileCounts t;
    t.tc = (ntiles_direct && !inc) ? tile_count.as<uint32_t>() : nullptr;
    t.gen = scan_gen.as<uint32_t>();
    t.ntiles = ntiles_direct;
    return t;
  }

  // Bucketing of a state: hash pass (k_integrate) then ordering (scan, lists, rank/gather).
  void bucket(const Vec4<T>* ps, const Vec4<T>* vs, const int32_t* is, Vec4<T>* pd, Vec4<T>* vd,
              int32_t* idd, uint32_t* off, int32_t* cell_sorted, int mode, Profiler* prof,
              Vec4<T>* xpre_sorted = nullptr) {
    hash(ps, vs, is, mode, prof);
    order(ps, vs, is, pd, vd, idd, off, cell_sorted, mode, prof, xpre_sorted);
  }

  // Hash pass: cell, count and id-block entry or slot of every particle (counts must be zero).
  // inc: the first pass of the incremental bucketing (every sphere arrives in its cell).
  void hash(const Vec4<T>* ps, const Vec4<T>* vs, const int32_t* is, int mode, Profiler* prof,
            bool inc = false) {
    int nb_i;
    const BlockInfo* bt_i = table(tuning.block_integrate, &nb_i);
    const SceneDev<T>* one = single() ? scenes_h.data() : nullptr;
    launch(0, prof, [&] {
      return launch_integrate<T>(stream, nb_i, tuning.block_integrate, one ? nullptr : bt_i,
                                 scenes_d.as<SceneDev<T>>(), one, id_blocks(), ps, vs, is,
                                 cell_of.as<int32_t>(), count.as<uint32_t>(),
                                 cell_ids.as<int32_t>(), (uint32_t)id_stride, ovf.as<int2>(),
                                 ovf_ctr.as<uint32_t>(),
                                 tiles(inc), mode, inc ? arr.as<uint64_t>() : nullptr,
                                 inc ? mark.as<uint8_t>() : nullptr);
    });
  }

  // Ordering after a hash pass: offsets (clearing the counts), lists, rank/gather into pd/vd/idd.
  // off_old: the incremental bucketing's order, from the offsets of the state (ps, vs, is) and its
  // slots' ranks (rank_old), writing the ranks of the sorted state to rank_o; its counts are kept.
  void order(const Vec4<T>* ps, const Vec4<T>* vs, const int32_t* is, Vec4<T>* pd, Vec4<T>* vd,
             int32_t* idd, uint32_t* off, int32_t* cell_sorted, int mode, Profiler* prof,
             Vec4<T>* xpre_sorted, const uint32_t* off_old = nullptr,
             const uint8_t* rank_old = nullptr, uint8_t* rank_o = nullptr) {
    int nb_s;
    const BlockInfo* bt_s = table(tuning.block_sort, &nb_s);
    const SceneDev<T>* one = single() ? scenes_h.data() : nullptr;
    const int32_t nn = (int32_t)n;
    const bool blocks = id_blocks();
    IncRank inc;
    if (off_old) {
      inc.arr = arr.as<uint64_t>();
      inc.mark = mark.as<uint8_t>();
      inc.off_old = off_old;
      inc.cell_cur_o = cell_cur.as<int32_t>();
      inc.rank_old = rank_old;
      inc.rank_o = rank_o;
    }
    launch(1, prof, [&] {
      if (off_old)
        return launch_scan_tiles(stream, count.as<uint32_t>(), off, tile_totals(), inc_tiles(),
                                 ovf_ctr.as<uint32_t>(), flags.as<uint32_t>(), (uint32_t)n);
      if (ntiles_direct)
        return launch_scan_direct(stream, count.as<uint32_t>(), off, tiles(), ovf_ctr.as<uint32_t>(),
                                  blocks);
      return launch_scan(stream, count.as<uint32_t>(), off, tile_state.as<uint64_t>(), ntiles,
                         ovf_ctr.as<uint32_t>(), flags.as<uint32_t>(), blocks && !off_old,
                         (uint32_t)n);
    });
    if (!blocks)
      launch(4, prof, [&] {
        return launch_scatter_ids(stream, tuning.block_sort, nn, cell_of.as<int32_t>(),
                                  count.as<uint32_t>(), off, is, list.as<int32_t>());
      });
    launch(2, prof, [&] {
      return launch_rank_gather<T>(stream, nb_s, tuning.block_sort, one ? nullptr : bt_s,
                                   scenes_d.as<SceneDev<T>>(), one, blocks, nn,
                                   cell_of.as<int32_t>(), off, cell_ids.as<int32_t>(), (uint32_t)id_stride,
                                   ovf.as<int2>(), ovf_ctr.as<uint32_t>(), list.as<int32_t>(), ps,
                                   vs, is, tent_d.as<TriEntry<T>>(), tcell_off_d.as<int32_t>(),
                                   flags.as<uint32_t>(), mode, cull_d.as<uint8_t>(), pd, vd, idd,
                                   cell_sorted, home.as<uint32_t>(), xpre_sorted,
                                   scan_gen.as<uint32_t>(), inc);
    });
  }

  void bucket_and_sort(int mode, Profiler* prof) {
    const int a = cur_pos, vc = cur_vel;
    Vec4<T>* p = pos[a].as<Vec4<T>>();
    Vec4<T>* v = vel[vc].as<Vec4<T>>();
    bucket(p, v, id[a].as<int32_t>(), pos[1 - a].as<Vec4<T>>(), vel[1 - vc].as<Vec4<T>>(),
           id[1 - a].as<int32_t>(), offs(), nullptr, mode, prof);
  }

  // A single scene launches without the workgroup table (constants by value, slices by index).
  bool single() const { return scenes_h.size() == 1; }

  // The cell offsets (C_pad + 1 entries) inside their buffer, which holds spare entries around them.
  uint32_t* offs() const { return offset.as<uint32_t>() + kOffsetLead; }

  // Form of the pair kernel (tuning.pair_variant; 0 means 1): lanes per particle, or kPairStaged
  // (one lane per particle, neighbour particles staged in LDS).
  int pair_lanes() const { return tuning.pair_variant > 1 ? tuning.pair_variant : 1; }
  int pair_particles_per_block() const {
    return (pair_lanes() == kPairStaged || pair_lanes() == kPairWave)
               ? tuning.block_pairs
               : tuning.block_pairs / pair_lanes();
  }

  // Arguments of a hash pass done inside another kernel, for the particles with ids `ids` (in the
  // kernel's particle order); ids == nullptr: none. inc: the incremental bucketing's.
  NextBucket next_bucket_args(const int32_t* ids, bool inc = false) const {
    NextBucket nb;
    std::memset(&nb, 0, sizeof(nb));
    if (ids) {
      nb.id = ids;
      nb.cell_of = cell_of.as<int32_t>();
      nb.count = count.as<uint32_t>();
      nb.cell_ids = cell_ids.as<int32_t>();
      nb.id_stride = (uint32_t)id_stride;
      nb.ovf = ovf.as<int2>();
      nb.ovf_ctr = ovf_ctr.as<uint32_t>();
      nb.blocks = id_blocks() ? 1 : 0;
      nb.tiles = tiles(inc);
      if (inc) {
        nb.arr = arr.as<unsigned long long>();
        nb.mark = mark.as<uint8_t>();
        nb.home = home.as<uint32_t>();
        nb.cell_cur = cell_cur.as<int32_t>();
        nb.tt = tile_totals();
      }
    }
    return nb;
  }

  // next_ids: with the sorted ids of the pair kernel's state, it also does the next step's hash
  // pass (NextBucket); nullptr: velocities only.
  void pairs(const Vec4<T>* ps, const Vec4<T>* vs, const uint32_t* off, Vec4<T>* vout,
             Profiler* prof, const int32_t* next_ids = nullptr, bool inc = false) {
    int nb_p;
    const BlockInfo* bt_p = table(pair_particles_per_block(), &nb_p);
    const NextBucket nb = next_bucket_args(next_ids, inc);
    launch(3, prof, [&] {
      return launch_pairs<T>(stream, nb_p, tuning.block_pairs, pair_lanes(),
                             tuning.pair_chunk == 2 ? 2 : 4, single() ? nullptr : bt_p, scenes_d.as<SceneDev<T>>(),
                             single() ? scenes_h.data() : nullptr, ps, vs, off,
                             home.as<uint32_t>(), vout, nb);
    });
  }

  // The hash pass of the current state is done (by k_integrate or by the previous step's pair
  // kernel): the counts hold it and the next step skips k_integrate.
  bool hashed = false;

  // The incremental bucketing starts with every sphere arriving in its cell and with empty ranges
  // in the offsets of the state (no sphere stays): those offsets are zeroed. Its generation is
  // advanced first: a hash pass dropped since the last scan (drop_hash) has left arrival words of
  // the generation the next scan would complete, which this pass must not add to. The kept tile
  // totals are then summed from the pass's counts (a reduction, where per-sphere atomics on a few
  // hundred totals would contend on an unsorted state).
  void ensure_hashed(Profiler* prof) {
    if (hashed || cuda) return;  // the CUDA backend's step hashes after its container
    const bool inc = incremental();
    if (inc) {
      PSX_HIP(launch_add_u32(stream, ovf_ctr.as<uint32_t>() + 2, 1u));
      PSX_HIP(launch_zero_u32(stream, offs_at(cur_pos) - kOffsetLead, (int64_t)(offset.bytes / 4)));
    }
    hash(pos[cur_pos].as<Vec4<T>>(), vel[cur_vel].as<Vec4<T>>(), id[cur_pos].as<int32_t>(),
         kModeIntegrate, prof, inc);
    if (inc) PSX_HIP(launch_tile_totals(stream, count.as<uint32_t>(), tile_totals()));
    hashed = true;
  }

  // Discards a pending hash pass (before the state, the parameters or the bucketing form change,
  // or before a bucketing that starts from zero counts). A fill kernel, not a memset, in case the
  // stream is being captured by the caller.
  void drop_hash() {
    if (!hashed) return;
    PSX_HIP(launch_zero_u32(stream, count.as<uint32_t>(), (int64_t)cpad));
    PSX_HIP(launch_zero_u32(stream, ovf_ctr.as<uint32_t>(), 1));
    PSX_HIP(launch_zero_u32(stream, tile_count.as<uint32_t>(), 2 * (int64_t)ntiles_direct * kTileStride));
    hashed = false;
  }

  // One step of the reference CUDA backend (PhysicsEngineCUDA.cpp:36-43): integration and the
  // container in k_front, which also hashes the post-container state; scan; rank and gather (no
  // integration or container there); the pair kick, with no hash pass of the next step (the next
  // step's container comes first).
  void step_once_cuda(Profiler* prof) {
    const int a = cur_pos, vc = cur_vel;
    int nb_i;
    const BlockInfo* bt_i = table(tuning.block_integrate, &nb_i);
    const SceneDev<T>* one = single() ? scenes_h.data() : nullptr;
    launch(0, prof, [&] {
      return launch_front<T>(stream, nb_i, tuning.block_integrate, one ? nullptr : bt_i,
                             scenes_d.as<SceneDev<T>>(), one, id_blocks(), pos[a].as<Vec4<T>>(),
                             vel[vc].as<Vec4<T>>(), id[a].as<int32_t>(), tent_d.as<TriEntry<T>>(),
                             tcell_off_d.as<int32_t>(), cull_d.as<uint8_t>(),
                             cell_of.as<int32_t>(), count.as<uint32_t>(), cell_ids.as<int32_t>(),
                             (uint32_t)id_stride, ovf.as<int2>(), ovf_ctr.as<uint32_t>(), tiles());
    });
    order(pos[a].as<Vec4<T>>(), vel[vc].as<Vec4<T>>(), id[a].as<int32_t>(), pos[1 - a].as<Vec4<T>>(),
          vel[1 - vc].as<Vec4<T>>(), id[1 - a].as<int32_t>(), offs(), nullptr, kModeCheck, prof,
          nullptr);
    pairs(pos[1 - a].as<Vec4<T>>(), vel[1 - vc].as<Vec4<T>>(), offs(), vel[vc].as<Vec4<T>>(), prof);
    cur_pos = 1 - a;
  }

  // One step from a hashed state; leaves the next state hashed.
  void step_once(Profiler* prof) {
    if (cuda) {
      step_once_cuda(prof);
      return;
    }
    const int a = cur_pos, vc = cur_vel;
    ensure_hashed(prof);
    const bool inc = incremental();
    // the incremental bucketing reads the offsets and slot ranks of the state it orders (position
    // buffer a) and writes those of the sorted state into the other arrays
    uint32_t* const off = inc ? offs_at(1 - a) : offs();
    order(pos[a].as<Vec4<T>>(), vel[vc].as<Vec4<T>>(), id[a].as<int32_t>(), pos[1 - a].as<Vec4<T>>(),
          vel[1 - vc].as<Vec4<T>>(), id[1 - a].as<int32_t>(), off, nullptr,
          kModeIntegrate | kModeContainer | kModeCheck, prof, nullptr, inc ? offs_at(a) : nullptr,
          rank8[a].as<uint8_t>(), rank8[1 - a].as<uint8_t>());
    pairs(pos[1 - a].as<Vec4<T>>(), vel[1 - vc].as<Vec4<T>>(), off,
          vel[vc].as<Vec4<T>>(), prof, id[1 - a].as<int32_t>(), inc);
    cur_pos = 1 - a;
  }

  // ---- differentiable windows --------------------------------------------------------------
  // Forward: the state is saved every win_every steps (ck_*, slot nseg = final state). Backward,
  // per segment: the segment's start state is copied to seg_*, the segment is recomputed with every
  // step recorded (its bucket-sorted post-container snapshot rs_*, offsets rs_off, bucket cells
  // rs_cell, and the new velocities rs_vnew), then the adjoint sweeps the steps in reverse. Step k's
  // pre-step state is seg_* for k = 0 and (rs_pos, rs_vnew, rs_id)[k-1] otherwise, so the
  // recompute copies nothing. Both halves are replayed from hipGraphs captured per segment length.
  int64_t win_steps = 0, win_every = 0, win_nseg = 0;
  DevBuf ck_pos, ck_vel, ck_id, seg_pos, seg_vel, seg_id;
  DevBuf rs_pos, rs_vel, rs_id, rs_cell, rs_off, rs_vnew, rs_xpre;
  DevBuf posH, velH, idH, offH;
  DevBuf lxA, lvA, lxB, lvB, gk, ge, gg, gout;
  std::map<int64_t, hipGraphExec_t> g_replay, g_adjoint;
  double win_replay_ms = 0.0, win_adjoint_ms = 0.0;
  // Ids of forward windows: each psx_window_forward takes the next one; win_id names the window
  // whose checkpoints are held (0: none, or its coefficients have changed since).
  int64_t win_counter = 0, win_id = 0;

  void destroy_window_graphs() {
    for (auto* m : {&g_replay, &g_adjoint}) {
      for (auto& kv : *m) (void)hipGraphExecDestroy(kv.second);
      m->clear();
    }
  }

  void save_state(int64_t slot_k) {
    const size_t nn = (size_t)n;
    PSX_HIP(hipMemcpyAsync(ck_pos.as<Vec4<T>>() + slot_k * nn, pos[cur_pos].p, nn * sizeof(Vec4<T>), hipMemcpyDeviceToDevice, stream));
    PSX_HIP(hipMemcpyAsync(ck_vel.as<Vec4<T>>() + slot_k * nn, vel[cur_vel].p, nn * sizeof(Vec4<T>), hipMemcpyDeviceToDevice, stream));
    PSX_HIP(hipMemcpyAsync(ck_id.as<int32_t>() + slot_k * nn, id[cur_pos].p, nn * 4, hipMemcpyDeviceToDevice, stream));
  }

  void load_state(int64_t slot_k, void* dp, void* dv, void* di) {
    const size_t nn = (size_t)n;
    PSX_HIP(hipMemcpyAsync(dp, ck_pos.as<Vec4<T>>() + slot_k * nn, nn * sizeof(Vec4<T>), hipMemcpyDeviceToDevice, stream));
    PSX_HIP(hipMemcpyAsync(dv, ck_vel.as<Vec4<T>>() + slot_k * nn, nn * sizeof(Vec4<T>), hipMemcpyDeviceToDevice, stream));
    PSX_HIP(hipMemcpyAsync(di, ck_id.as<int32_t>() + slot_k * nn, nn * 4, hipMemcpyDeviceToDevice, stream));
  }

  int64_t window_forward(int64_t nsteps, int64_t every) override {
    if (cuda)
      throw Error(PSX_ERR_STATE, "the differentiable window implements the CPU backend's step; this "
                                 "handle was created with PSX_FLAG_SEMANTICS_CUDA");
    const size_t nn = (size_t)n;
    const int64_t nseg = (nsteps + every - 1) / every;
    if (nseg != win_nseg || every != win_every) {
      destroy_window_graphs();
      ck_pos.alloc((size_t)(nseg + 1) * nn * sizeof(Vec4<T>));
      ck_vel.alloc((size_t)(nseg + 1) * nn * sizeof(Vec4<T>));
      ck_id.alloc((size_t)(nseg + 1) * nn * 4);
      for (DevBuf* b : {&seg_pos, &seg_vel, &posH, &velH})
        b->alloc(nn * sizeof(Vec4<T>));
      seg_id.alloc(nn * 4);
      idH.alloc(nn * 4);
      offH.alloc(((size_t)cpad + 1) * 4);
      rs_pos.alloc(((size_t)every * nn + kPairPad) * sizeof(Vec4<T>));
      rs_vel.alloc(((size_t)every * nn + kPairPad) * sizeof(Vec4<T>));
      rs_vnew.alloc((size_t)every * nn * sizeof(Vec4<T>));
      rs_xpre.alloc((size_t)every * nn * sizeof(Vec4<T>));
      rs_id.alloc((size_t)every * nn * 4);
      rs_cell.alloc((size_t)every * nn * 4);
      // k_scan writes cpad + 1 per record; spare entries around the records as around `offset`
      rs_off.alloc(((size_t)every * ((size_t)cpad + 1) + kOffsetPad) * 4);
      for (DevBuf* b : {&lxA, &lvA, &lxB, &lvB}) b->alloc(3 * nn * sizeof(T));
      for (DevBuf* b : {&gk, &ge, &gg}) b->alloc(nn * sizeof(T));
      gout.alloc(3 * sizeof(double) * scenes.size());
    }
    win_steps = nsteps;
    win_every = every;
    win_nseg = nseg;
    win_id = 0;  // the checkpoints are being replaced
    for (int64_t sgm = 0; sgm < nseg; ++sgm) {
      save_state(sgm);
      step(std::min(every, nsteps - sgm * every));
    }
    save_state(nseg);
    win_id = ++win_counter;
    return win_id;
  }

  // pre-step state of recorded step k
  const Vec4<T>* pre_pos(int64_t k) {
    return k == 0 ? seg_pos.as<Vec4<T>>() : rs_pos.as<Vec4<T>>() + (k - 1) * n;
  }
  const Vec4<T>* pre_vel(int64_t k) {
    return k == 0 ? seg_vel.as<Vec4<T>>() : rs_vnew.as<Vec4<T>>() + (k - 1) * n;
  }
  const int32_t* pre_id(int64_t k) {
    return k == 0 ? seg_id.as<int32_t>() : rs_id.as<int32_t>() + (k - 1) * n;
  }

  // The recompute: as a run of steps, the hash pass of step k + 1 is done by step k's pair kernel
  // (its pre-step state is step k's snapshot, new velocities and ids); the last step leaves none
  // pending.
  void replay_segment(int64_t ks) {
    const size_t nn = (size_t)n, nc = (size_t)cpad + 1;  // offset record stride
    if (ks > 0) hash(pre_pos(0), pre_vel(0), pre_id(0), kModeIntegrate, nullptr);
    for (int64_t k = 0; k < ks; ++k) {
      Vec4<T>* sp = rs_pos.as<Vec4<T>>() + k * nn;
      Vec4<T>* sv = rs_vel.as<Vec4<T>>() + k * nn;
      uint32_t* so = rs_off.as<uint32_t>() + kOffsetLead + k * nc;
      int32_t* si = rs_id.as<int32_t>() + k * nn;
      order(pre_pos(k), pre_vel(k), pre_id(k), sp, sv, si, so, rs_cell.as<int32_t>() + k * nn,
            kModeIntegrate | kModeContainer | kModeCheck, nullptr, rs_xpre.as<Vec4<T>>() + k * nn);
      pairs(sp, sv, so, rs_vnew.as<Vec4<T>>() + k * nn, nullptr, k + 1 < ks ? si : nullptr);
    }
  }

  // Adjoint sweep of a recorded segment, last step first; the cotangent alternates between the A
  // and B buffers and ends in A.
  void adjoint_segment(int64_t ks) {
    const size_t nn = (size_t)n, nc = (size_t)cpad + 1;  // offset record stride
    int nb;
    const BlockInfo* bt = table(kAdjPerBlock, &nb);
    T *lx = lxA.as<T>(), *lv = lvA.as<T>(), *lx_o = lxB.as<T>(), *lv_o = lvB.as<T>();
    // home-cell bucketing of each snapshot (no integration, no container: nothing is written to
    // the snapshot); the hash pass of step k - 1's is done by step k's adjoint kernel
    if (ks > 0)
      hash(rs_pos.as<Vec4<T>>() + (ks - 1) * nn, rs_vel.as<Vec4<T>>() + (ks - 1) * nn,
           rs_id.as<int32_t>() + (ks - 1) * nn, 0, nullptr);
    for (int64_t k = ks - 1; k >= 0; --k) {
      Vec4<T>* sp = rs_pos.as<Vec4<T>>() + k * nn;
      Vec4<T>* sv = rs_vel.as<Vec4<T>>() + k * nn;
      const int32_t* si = rs_id.as<int32_t>() + k * nn;
      order(sp, sv, si, posH.as<Vec4<T>>(), velH.as<Vec4<T>>(), idH.as<int32_t>(),
            offH.as<uint32_t>(), nullptr, 0, nullptr, nullptr);
      const NextBucket hn = next_bucket_args(k > 0 ? si - nn : nullptr);
      PSX_HIP(launch_adj_step<T>(stream, nb, bt, scenes_d.as<SceneDev<T>>(), sp, sv, si,
                                 rs_off.as<uint32_t>() + kOffsetLead + k * nc,
                                 rs_cell.as<int32_t>() + k * nn,
                                 rs_xpre.as<Vec4<T>>() + k * nn, posH.as<Vec4<T>>(),
                                 velH.as<Vec4<T>>(), idH.as<int32_t>(), offH.as<uint32_t>(),
                                 tgeo_d.as<TriGeo<T>>(), tbox_d.as<TriBox<T>>(),
                                 tcell_off_d.as<int32_t>(), lx, lv, lx_o, lv_o, gk.as<T>(),
                                 ge.as<T>(), gg.as<T>(), hn, k > 0 ? sp - nn : nullptr));
      std::swap(lx, lx_o);
      std::swap(lv, lv_o);
    }
    if (lx != lxA.as<T>())
      PSX_HIP(launch_copy2<T>(stream, 3 * (int64_t)nn, lx, lv, lxA.as<T>(), lvA.as<T>()));
  }

  template <class F>
  hipGraphExec_t capture_graph(F&& f) {
    hipGraph_t g = nullptr;
    PSX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      f();
    } catch (...) {
      hipGraph_t dummy = nullptr;
      (void)hipStreamEndCapture(stream, &dummy);
      if (dummy) (void)hipGraphDestroy(dummy);
      throw;
    }
    PSX_HIP(hipStreamEndCapture(stream, &g));
    hipGraphExec_t ge_ = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) throw Error(PSX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return ge_;
  }

  // The graph of fn(ks), captured and instantiated on first use (its own call, so that a failure
  // names the capture or the launch).
  __attribute__((noinline)) hipGraphExec_t cached_graph(std::map<int64_t, hipGraphExec_t>& cache,
                                                        int64_t ks, void (SimImpl::*fn)(int64_t)) {
    auto it = cache.find(ks);
    if (it == cache.end()) it = cache.emplace(ks, capture_graph([&] { (this->*fn)(ks); })).first;
    return it->second;
  }

  void run_cached(std::map<int64_t, hipGraphExec_t>& cache, int64_t ks, void (SimImpl::*fn)(int64_t)) {
    if (!tuning.use_graph) {
      (this->*fn)(ks);
      return;
    }
    const hipGraphExec_t ge_ = cached_graph(cache, ks, fn);
    const hipError_t e = hipGraphLaunch(ge_, stream);
    if (e != hipSuccess)
      throw Error(PSX_ERR_HIP, std::string("hipGraphLaunch (window segment of ") + std::to_string(ks) +
                                   " steps): " + hipGetErrorString(e));
  }

  void window_backward(int64_t window_id, const void* dlx, const void* dlv, void* dlx0, void* dlv0,
                       double* gparams) override {
    if (win_nseg <= 0 || win_id == 0)
      throw Error(PSX_ERR_STATE, "psx_window_backward needs a psx_window_forward first (and no "
                                 "psx_set_params since)");
    if (window_id != win_id)
      throw Error(PSX_ERR_STATE, "psx_window_backward: window " + std::to_string(window_id) +
                                     " is stale; the checkpoints held are those of window " +
                                     std::to_string(win_id));
    drop_hash();  // the recompute buckets from zero counts
    const size_t nn = (size_t)n;
    PSX_HIP(hipMemcpyAsync(lxA.p, dlx, 3 * nn * sizeof(T), hipMemcpyDeviceToDevice, stream));
    PSX_HIP(hipMemcpyAsync(lvA.p, dlv, 3 * nn * sizeof(T), hipMemcpyDeviceToDevice, stream));
    for (DevBuf* b : {&gk, &ge, &gg}) PSX_HIP(launch_fill<T>(stream, (int64_t)nn, b->as<T>(), T(0)));
    int nb;
    table(kAdjPerBlock, &nb);  // created here, outside any graph capture
    std::vector<hipEvent_t> ev((size_t)(3 * win_nseg), nullptr);
    for (auto& e : ev) PSX_HIP(hipEventCreate(&e));
    struct EvGuard {
      std::vector<hipEvent_t>& v;
      ~EvGuard() {
        for (auto e : v)
          if (e) (void)hipEventDestroy(e);
      }
    } guard{ev};
    for (int64_t sgm = win_nseg - 1; sgm >= 0; --sgm) {
      const int64_t ks = std::min(win_every, win_steps - sgm * win_every);
      load_state(sgm, seg_pos.p, seg_vel.p, seg_id.p);
      PSX_HIP(hipEventRecord(ev[(size_t)(3 * sgm)], stream));
      run_cached(g_replay, ks, &SimImpl::replay_segment);
      PSX_HIP(hipEventRecord(ev[(size_t)(3 * sgm + 1)], stream));
      run_cached(g_adjoint, ks, &SimImpl::adjoint_segment);
      PSX_HIP(hipEventRecord(ev[(size_t)(3 * sgm + 2)], stream));
    }
    PSX_HIP(hipMemcpyAsync(dlx0, lxA.p, 3 * nn * sizeof(T), hipMemcpyDeviceToDevice, stream));
    PSX_HIP(hipMemcpyAsync(dlv0, lvA.p, 3 * nn * sizeof(T), hipMemcpyDeviceToDevice, stream));
    for (size_t sc = 0; sc < scenes.size(); ++sc) {
      const int64_t b0 = scenes[sc].p_begin;
      PSX_HIP(launch_reduce3<T>(stream, (int32_t)scenes[sc].n, gk.as<T>() + b0, ge.as<T>() + b0,
                                gg.as<T>() + b0, gout.as<double>() + 3 * sc));
    }
    std::vector<double> h(3 * scenes.size());
    PSX_HIP(hipMemcpyAsync(h.data(), gout.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    PSX_HIP(hipStreamSynchronize(stream));
    if (gparams) std::memcpy(gparams, h.data(), h.size() * sizeof(double));
    win_replay_ms = win_adjoint_ms = 0.0;
    for (int64_t sgm = 0; sgm < win_nseg; ++sgm) {
      float a = 0.f, b = 0.f;
      PSX_HIP(hipEventElapsedTime(&a, ev[(size_t)(3 * sgm)], ev[(size_t)(3 * sgm + 1)]));
      PSX_HIP(hipEventElapsedTime(&b, ev[(size_t)(3 * sgm + 1)], ev[(size_t)(3 * sgm + 2)]));
      win_replay_ms += a;
      win_adjoint_ms += b;
    }
  }

  void window_stats(double* replay_ms, double* adjoint_ms) override {
    if (replay_ms) *replay_ms = win_replay_ms;
    if (adjoint_ms) *adjoint_ms = win_adjoint_ms;
  }

  void set_params(int32_t scene, const psx_params& p) override {
    SceneHost& sh = scenes[(size_t)scene];
    for (int a = 0; a < 3; ++a)
      if (p.box_min[a] != sh.prm.box_min[a] || p.box_max[a] != sh.prm.box_max[a])
        throw Error(PSX_ERR_INVALID, "psx_set_params cannot change the container bounds");
    if (p.dtype != sh.prm.dtype) throw Error(PSX_ERR_INVALID, "psx_set_params cannot change dtype");
    if (p.flags != sh.prm.flags) throw Error(PSX_ERR_INVALID, "psx_set_params cannot change flags");
    sh.prm = p;
    win_id = 0;   // a held window was run with the old coefficients
    drop_hash();  // hashed with the old dt
    PSX_HIP(hipStreamSynchronize(stream));
    SceneDev<T>& d = scenes_h[(size_t)scene];
    fill_coefficients(d, sh.prm);
    PSX_HIP(hipMemcpy(scenes_d.as<SceneDev<T>>() + scene, &d, sizeof(d), hipMemcpyHostToDevice));
    destroy_graphs();  // single-scene launches carry the constants by value in captured graphs
  }

  static void fill_coefficients(SceneDev<T>& d, const psx_params& prm) {
    for (int a = 0; a < 3; ++a) {
      d.bmin[a] = (T)prm.box_min[a];
      d.bmax[a] = (T)prm.box_max[a];
      const T ga = (T)prm.gravity[a];
      const T dta = (T)prm.dt;
      d.gdt[a] = ga * dta;
    }
    // walls: zMin, zMax, xMin, xMax, yMin, yMax
    d.emin[0] = (T)prm.walls[2];
    d.emax[0] = (T)prm.walls[3];
    d.emin[1] = (T)prm.walls[4];
    d.emax[1] = (T)prm.walls[5];
    d.emin[2] = (T)prm.walls[0];
    d.emax[2] = (T)prm.walls[1];
    d.dt = (T)prm.dt;
    d.spring = (T)prm.spring;
    d.es = (T)prm.restitution;
    d.ew = (T)prm.walls[0];
  }

  // Captures nsteps steps from buffer parity (a, vc) into *out; the host state is left as it was.
  void capture(int nsteps, int a, int vc, hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    PSX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      for (int k = 0; k < nsteps; ++k) step_once(nullptr);
    } catch (...) {
      hipGraph_t dummy = nullptr;
      (void)hipStreamEndCapture(stream, &dummy);
      if (dummy) (void)hipGraphDestroy(dummy);
      cur_pos = a;
      cur_vel = vc;
      throw;
    }
    PSX_HIP(hipStreamEndCapture(stream, &g));
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) throw Error(PSX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    // the graph's device-side set-up now, in stream order, rather than at its first launch
    PSX_HIP(hipGraphUpload(*out, stream));
    cur_pos = a;
    cur_vel = vc;
  }

  void step(int64_t nsteps) override {
    int64_t left = nsteps;
    if (left > 0) ensure_hashed(nullptr);  // outside any capture: the graphs start hashed
    if (tuning.use_graph) {
      for (int li = 0; li < 2; ++li) {
        while (left >= kGraphSteps[li]) {
          hipGraphExec_t& g = graphs[li][cur_pos][cur_vel];
          if (!g) capture(kGraphSteps[li], cur_pos, cur_vel, &g);
          PSX_HIP(hipGraphLaunch(g, stream));
          left -= kGraphSteps[li];
        }
      }
      if (left > 0 && tail_graphs[left][cur_pos][cur_vel]) {  // prepared
        PSX_HIP(hipGraphLaunch(tail_graphs[left][cur_pos][cur_vel], stream));
        cur_pos ^= (int)(left & 1);
        left = 0;
      }
    }
    for (; left > 0; --left) step_once(nullptr);
  }

  // A run of nsteps from the current state replays 64- and 16-step graphs of the current buffer
  // parity (both lengths are even) and then the graph of the steps left: capture those it will use.
  void prepare(int64_t nsteps) override {
    if (nsteps <= 0 || !tuning.use_graph) return;
    ensure_hashed(nullptr);
    int64_t left = nsteps;
    for (int li = 0; li < 2; ++li) {
      hipGraphExec_t& g = graphs[li][cur_pos][cur_vel];
      if (left >= kGraphSteps[li] && !g) capture(kGraphSteps[li], cur_pos, cur_vel, &g);
      left %= kGraphSteps[li];
    }
    hipGraphExec_t& t = tail_graphs[left][cur_pos][cur_vel];
    if (left > 0 && !t) capture((int)left, cur_pos, cur_vel, &t);
  }

  void profile(int64_t nsteps, double* ms, double* per_step) override {
    Profiler prof(kNumKernels, nsteps);
    for (int64_t s = 0; s < nsteps; ++s) {
      prof.next_step();
      step_once(&prof);
    }
    PSX_HIP(hipStreamSynchronize(stream));
    prof.collect(ms, per_step);
  }

  void set_stream(hipStream_t s) override {
    PSX_HIP(hipStreamSynchronize(stream));
    destroy_graphs();
    stream = s ? s : own_stream;
  }

  void retune(const psx_tuning& t) override {
    drop_hash();  // the bucketing form may change
    PSX_HIP(hipStreamSynchronize(stream));
    destroy_graphs();
    tuning = t;
    int nb;
    table(tuning.block_integrate, &nb);
    table(tuning.block_sort, &nb);
    table(pair_particles_per_block(), &nb);
    inc_buffers();
  }

  void get_state(double* hp, double* hv) override {
    const int64_t nn = n;
    T* d = io.as<T>();
    PSX_HIP(launch_unsort<T>(stream, (int32_t)nn, pos[cur_pos].as<Vec4<T>>(), vel[cur_vel].as<Vec4<T>>(),
                             id[cur_pos].as<int32_t>(), d, d + 3 * nn));
    std::vector<T> h((size_t)(6 * nn));
    PSX_HIP(hipMemcpyAsync(h.data(), d, h.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
    PSX_HIP(hipStreamSynchronize(stream));
    for (int64_t k = 0; k < 3 * nn; ++k) {
      if (hp) hp[k] = (double)h[(size_t)k];
      if (hv) hv[k] = (double)h[(size_t)(3 * nn + k)];
    }
  }

  void set_state(const double* hp, const double* hv) override {
    const int64_t nn = n;
    std::vector<T> h((size_t)(6 * nn));
    for (int64_t k = 0; k < 3 * nn; ++k) {
      h[(size_t)k] = hp ? (T)hp[k] : T(0);
      h[(size_t)(3 * nn + k)] = hv ? (T)hv[k] : T(0);
    }
    T* d = io.as<T>();
    drop_hash();
    PSX_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream));
    PSX_HIP(launch_sort_in<T>(stream, (int32_t)nn, pos[cur_pos].as<Vec4<T>>(), vel[cur_vel].as<Vec4<T>>(),
                              id[cur_pos].as<int32_t>(), hp ? d : nullptr, hv ? d + 3 * nn : nullptr));
    PSX_HIP(hipMemsetAsync(flags.p, 0, flags.bytes, stream));
    PSX_HIP(hipStreamSynchronize(stream));
  }

  void get_state_device(void* dp, void* dv) override {
    PSX_HIP(launch_unsort<T>(stream, (int32_t)n, pos[cur_pos].as<Vec4<T>>(), vel[cur_vel].as<Vec4<T>>(),
                             id[cur_pos].as<int32_t>(), static_cast<T*>(dp), static_cast<T*>(dv)));
  }

  void set_state_device(const void* dp, const void* dv) override {
    drop_hash();
    PSX_HIP(launch_sort_in<T>(stream, (int32_t)n, pos[cur_pos].as<Vec4<T>>(), vel[cur_vel].as<Vec4<T>>(),
                              id[cur_pos].as<int32_t>(), static_cast<const T*>(dp),
                              static_cast<const T*>(dv)));
  }

  void debug_bucketing(int32_t* cell, uint32_t* cnt, uint32_t* off, int32_t* perm) override {
    const int a = cur_pos, vc = cur_vel;
    const int64_t nn = n;
    drop_hash();
    bucket_and_sort(0, nullptr);
    std::vector<int32_t> hc((size_t)nn), hid((size_t)nn), hperm((size_t)nn);
    std::vector<uint32_t> hoff((size_t)ncells + 1);
    PSX_HIP(hipMemcpyAsync(hc.data(), cell_of.p, (size_t)nn * 4, hipMemcpyDeviceToHost, stream));
    PSX_HIP(hipMemcpyAsync(hid.data(), id[a].p, (size_t)nn * 4, hipMemcpyDeviceToHost, stream));
    PSX_HIP(hipMemcpyAsync(hperm.data(), id[1 - a].p, (size_t)nn * 4, hipMemcpyDeviceToHost, stream));
    PSX_HIP(hipMemcpyAsync(hoff.data(), offs(), ((size_t)ncells) * 4, hipMemcpyDeviceToHost, stream));
    PSX_HIP(hipStreamSynchronize(stream));
    hoff[(size_t)ncells] = (uint32_t)nn;  // offset[C] == n (padding cells are empty)
    cur_pos = 1 - a;
    cur_vel = 1 - vc;
    if (cell)
      for (int64_t p = 0; p < nn; ++p) cell[hid[(size_t)p]] = hc[(size_t)p];
    if (off) std::memcpy(off, hoff.data(), hoff.size() * 4);
    if (cnt)
      for (int64_t c = 0; c < ncells; ++c) cnt[c] = hoff[(size_t)c + 1] - hoff[(size_t)c];
    if (perm) std::memcpy(perm, hperm.data(), (size_t)nn * 4);
  }

  // ParticleStatistics::calculateStatistics (ParticleStatistics.cpp:3-43) per scene: bucketing of
  // the current positions, per-workgroup partials on the device, combined here in workgroup order.
  void statistics(psx_stats* out) override {
    // bucket the current state; the cell-sorted copy lands in the spare buffers, which the next
    // step overwrites, so the state itself is left as it is
    drop_hash();
    bucket_and_sort(0, nullptr);
    int nb;
    const BlockInfo* bt = table(256, &nb);
    stat_part.alloc((size_t)nb * sizeof(StatPart));
    PSX_HIP(launch_stats<T>(stream, nb, single() ? nullptr : bt, scenes_d.as<SceneDev<T>>(),
                            single() ? scenes_h.data() : nullptr, pos[1 - cur_pos].as<Vec4<T>>(),
                            vel[1 - cur_vel].as<Vec4<T>>(), id[1 - cur_pos].as<int32_t>(),
                            offs(), home.as<uint32_t>(), stat_part.as<StatPart>()));
    std::vector<StatPart> h((size_t)nb);
    PSX_HIP(hipMemcpyAsync(h.data(), stat_part.p, h.size() * sizeof(StatPart), hipMemcpyDeviceToHost,
                           stream));
    PSX_HIP(hipStreamSynchronize(stream));
    size_t b = 0;  // workgroups follow the scenes in order, ceil(n_s / 256) each
    for (size_t sc = 0; sc < scenes.size(); ++sc) {
      psx_stats r;
      r.kinetic_energy = 0.0;
      r.n_collided = 0;
      r.max_overlap = 0.0;
      for (int a = 0; a < 3; ++a) {
        r.min[a] = INFINITY;
        r.max[a] = -INFINITY;
      }
      for (int64_t k = 0; k < (scenes[sc].n + 255) / 256; ++k, ++b) {
        const StatPart& q = h[b];
        r.kinetic_energy += q.ke;
        r.n_collided += q.n_collided;
        r.max_overlap = std::max(r.max_overlap, q.max_overlap);
        for (int a = 0; a < 3; ++a) {
          r.min[a] = std::min(r.min[a], q.mn[a]);
          r.max[a] = std::max(r.max[a], q.mx[a]);
        }
      }
      out[sc] = r;
    }
  }

  // statistics() without allocating: the partials and per-scene totals go to `scratch`, the totals
  // are combined on the device in the host combination's order and copied to out.
  void statistics_into(psx_stats* out, void* scratch_dev, size_t bytes) override {
    int nb;
    const BlockInfo* bt = table(kStatBlock, &nb);  // built at set-up: no allocation here
    const size_t parts = ((size_t)nb * sizeof(StatPart) + 255) / 256 * 256;
    const size_t need = parts + scenes.size() * sizeof(psx_stats);
    if (need > bytes)
      throw Error(PSX_ERR_INVALID, "psx_statistics_noalloc: " + std::to_string(n) + " particles in " +
                                       std::to_string(scenes.size()) + " scenes need " +
                                       std::to_string(need) + " bytes of scratch, more than its " +
                                       std::to_string(bytes));
    StatPart* part = static_cast<StatPart*>(scratch_dev);
    psx_stats* tot = reinterpret_cast<psx_stats*>(static_cast<char*>(scratch_dev) + parts);
    drop_hash();
    bucket_and_sort(0, nullptr);
    PSX_HIP(launch_stats<T>(stream, nb, single() ? nullptr : bt, scenes_d.as<SceneDev<T>>(),
                            single() ? scenes_h.data() : nullptr, pos[1 - cur_pos].as<Vec4<T>>(),
                            vel[1 - cur_vel].as<Vec4<T>>(), id[1 - cur_pos].as<int32_t>(),
                            offs(), home.as<uint32_t>(), part));
    PSX_HIP(launch_stats_combine(stream, (int32_t)scenes.size(), nb, bt, part, tot));
    PSX_HIP(hipMemcpyAsync(out, tot, scenes.size() * sizeof(psx_stats), hipMemcpyDeviceToHost,
                           stream));
    PSX_HIP(hipStreamSynchronize(stream));
  }

  uint32_t read_flags() override {
    uint32_t f[4];
    PSX_HIP(hipMemcpyAsync(f, flags.p, 16, hipMemcpyDeviceToHost, stream));
    PSX_HIP(hipStreamSynchronize(stream));
    return f[0];
  }
};

std::unique_ptr<SimBase> make_sim(int dtype) {
  if (dtype == PSX_F32) return std::make_unique<SimImpl<float>>();
  return std::make_unique<SimImpl<double>>();
}

const char* kernel_name(int k) {
  if (k < 0 || k >= kNumKernels) return nullptr;
  return kKernelNames[k];
}

// ---------------------------------------------------------------------------------------------
// Profiler
// ---------------------------------------------------------------------------------------------
Profiler::Profiler(int nk, int64_t nsteps) : nk_(nk) {
  ev_.resize((size_t)(nk * std::max<int64_t>(nsteps, 1) * 2), nullptr);
  rec_.assign((size_t)(nk * std::max<int64_t>(nsteps, 1)), 0);
  for (auto& e : ev_) PSX_HIP(hipEventCreate(&e));
}
Profiler::~Profiler() {
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
}
void Profiler::next_step() { ++step_; }
hipEvent_t Profiler::begin(int k) {
  rec_[(size_t)((step_ - 1) * nk_ + k)] = 1;
  return ev_[(size_t)(((step_ - 1) * nk_ + k) * 2)];
}
hipEvent_t Profiler::end(int k) { return ev_[(size_t)(((step_ - 1) * nk_ + k) * 2 + 1)]; }
void Profiler::collect(double* ms, double* per_step) {
  for (int k = 0; k < nk_; ++k) {
    double tot = 0.0;
    int64_t cnt = 0;
    for (int64_t s = 0; s < step_; ++s) {
      if (!rec_[(size_t)(s * nk_ + k)]) continue;
      float t = 0.0f;
      PSX_HIP(hipEventElapsedTime(&t, ev_[(size_t)((s * nk_ + k) * 2)], ev_[(size_t)((s * nk_ + k) * 2 + 1)]));
      tot += t;
      ++cnt;
    }
    ms[k] = cnt ? tot / (double)cnt : 0.0;
    if (per_step) per_step[k] = step_ ? (double)cnt / (double)step_ : 0.0;
  }
}

}  // namespace psx

// =================================================================================================
// C ABI
// =================================================================================================
using psx::Error;
using psx::set_error;

struct psx_sim {
  std::unique_ptr<psx::SimBase> impl;
};

// Scratch memory of psx_statistics_noalloc: the per-workgroup partials (72 bytes per 256
// particles) and the per-scene totals, in a fixed device array that every GPU running the library
// holds until the process ends (about 14 million particles per call at most). Calls that overlap
// on one GPU share it and race.
static constexpr size_t kScratchBytes = (size_t)4 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of `scratch` on the current device. The array's name is a host-side stand-in
// and never a device address: the address is looked up once per device, with that device current,
// and used for every kernel argument and copy.
static constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[kScratchDevices] = {};

static void* scratch_address() {
  int dev = -1;
  PSX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices)
    throw Error(PSX_ERR_INVALID, "psx_statistics_noalloc: device index " + std::to_string(dev) +
                                     " beyond the " + std::to_string(kScratchDevices) + " supported");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(PSX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  return g_scratch_dev[dev];
}

#define PSX_TRY(...)                                                          \
  try {                                                                       \
    __VA_ARGS__                                                               \
  } catch (const Error& e) {                                                  \
    return set_error(e.code, e.what());                                       \
  } catch (const std::bad_alloc&) {                                           \
    return set_error(PSX_ERR_ALLOC, "host allocation failed");                \
  } catch (const std::exception& e) {                                         \
    return set_error(PSX_ERR_INVALID, e.what());                              \
  }

static int check_params(const psx_params* p) {
  if (!p) return set_error(PSX_ERR_INVALID, "params is NULL");
  if (p->dtype != PSX_F32 && p->dtype != PSX_F64) return set_error(PSX_ERR_INVALID, "dtype must be PSX_F32 or PSX_F64");
  if (!(p->dt > 0.0) || !std::isfinite(p->dt)) return set_error(PSX_ERR_INVALID, "dt must be positive and finite");
  if (p->flags & ~PSX_FLAG_SEMANTICS_CUDA)
    return set_error(PSX_ERR_INVALID, "params.flags: unknown bits (only PSX_FLAG_SEMANTICS_CUDA)");
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(p->gravity[a])) return set_error(PSX_ERR_INVALID, "gravity must be finite");
  return PSX_OK;
}

static int create_impl(int32_t n_scenes, const psx_scene* scenes, const double* pos,
                       const double* vel, const double* radius, const double* mass,
                       psx_sim** out) {
  if (!out) return set_error(PSX_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (n_scenes <= 0 || !scenes) return set_error(PSX_ERR_INVALID, "need at least one scene");
  if (!pos || !vel || !radius || !mass) return set_error(PSX_ERR_INVALID, "particle arrays must not be NULL");
  const int32_t dtype = scenes[0].params.dtype;
  int64_t total = 0, tris = 0;
  for (int32_t s = 0; s < n_scenes; ++s) {
    int rc = check_params(&scenes[s].params);
    if (rc) return rc;
    if (scenes[s].params.dtype != dtype) return set_error(PSX_ERR_INVALID, "all scenes must share one dtype");
    if (scenes[s].params.flags != scenes[0].params.flags)
      return set_error(PSX_ERR_INVALID, "all scenes must share one semantics (params.flags)");
    if (scenes[s].n <= 0) return set_error(PSX_ERR_INVALID, "every scene needs at least one particle");
    if (scenes[s].mesh.n_triangles < 0 || (scenes[s].mesh.n_triangles > 0 && !scenes[s].mesh.vertices))
      return set_error(PSX_ERR_INVALID, "mesh vertices missing");
    total += scenes[s].n;
    tris += scenes[s].mesh.n_triangles;
  }
  // triangle indices are int32 on the device
  if (tris >= ((int64_t)1 << 31) - 1) return set_error(PSX_ERR_INVALID, "too many triangles (>= 2^31)");
  if (total >= ((int64_t)1 << 31) - 1) return set_error(PSX_ERR_INVALID, "too many particles (>= 2^31)");
  for (int64_t i = 0; i < total; ++i) {
    if (!(radius[i] > 0.0) || !std::isfinite(radius[i]))
      return set_error(PSX_ERR_INVALID, "radii must be positive and finite (particle " + std::to_string(i) + ")");
    if (!(mass[i] > 0.0) || !std::isfinite(mass[i]))
      return set_error(PSX_ERR_INVALID, "masses must be positive and finite (particle " + std::to_string(i) + ")");
  }
  PSX_TRY({
    auto impl = psx::make_sim(dtype);
    impl->n = total;
    impl->cuda = (scenes[0].params.flags & PSX_FLAG_SEMANTICS_CUDA) != 0;
    int64_t p0 = 0;
    for (int32_t s = 0; s < n_scenes; ++s) {
      psx::SceneHost sh;
      sh.prm = scenes[s].params;
      sh.n = scenes[s].n;
      sh.p_begin = p0;
      p0 += sh.n;
      if (scenes[s].mesh.n_triangles > 0) {
        psx::MeshHost m;
        m.nt = scenes[s].mesh.n_triangles;
        m.v.assign(scenes[s].mesh.vertices, scenes[s].mesh.vertices + 9 * m.nt);
        for (int a = 0; a < 3; ++a) {
          m.min[a] = m.v[a];
          m.max[a] = m.v[a];
        }
        // PolygonContainer::getMinCoords / getMaxCoords (PolygonContainer.cpp:288-318)
        for (int64_t k = 0; k < 3 * m.nt; ++k)
          for (int a = 0; a < 3; ++a) {
            const double x = m.v[(size_t)(3 * k + a)];
            if (!std::isfinite(x)) throw Error(PSX_ERR_INVALID, "mesh vertices must be finite");
            m.min[a] = (x < m.min[a]) ? x : m.min[a];
            m.max[a] = (m.max[a] < x) ? x : m.max[a];
          }
        sh.mesh = (int)impl->meshes.size();
        impl->meshes.push_back(std::move(m));
      } else {
        for (int a = 0; a < 3; ++a)
          if (!(sh.prm.box_max[a] > sh.prm.box_min[a]))
            throw Error(PSX_ERR_INVALID, "box_min must be < box_max on every axis");
      }
      impl->scenes.push_back(sh);
    }
    hipStream_t st;
    PSX_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    impl->own_stream = st;
    impl->stream = st;
    std::vector<double> hp(pos, pos + 3 * total), hv(vel, vel + 3 * total), hr(radius, radius + total),
        hm(mass, mass + total);
    impl->setup(hp, hv, hr, hm);
    auto* sim = new psx_sim;
    sim->impl = std::move(impl);
    *out = sim;
    return PSX_OK;
  })
}

extern "C" {

int psx_create(const psx_params* params, int64_t n, const double* pos, const double* vel,
               const double* radius, const double* mass, const psx_mesh* mesh, psx_sim** out) {
  if (!params) return set_error(PSX_ERR_INVALID, "params is NULL");
  if (n <= 0) return set_error(PSX_ERR_INVALID, "n must be positive");
  psx_scene s;
  std::memset(&s, 0, sizeof(s));
  s.params = *params;
  s.n = n;
  if (mesh) s.mesh = *mesh;
  return create_impl(1, &s, pos, vel, radius, mass, out);
}

int psx_batch_create(int32_t n_scenes, const psx_scene* scenes, const double* pos,
                     const double* vel, const double* radius, const double* mass, psx_sim** out) {
  return create_impl(n_scenes, scenes, pos, vel, radius, mass, out);
}

int psx_destroy(psx_sim* sim) {
  if (!sim) return PSX_OK;
  PSX_TRY({
    if (sim->impl) (void)hipStreamSynchronize(sim->impl->stream);
    delete sim;
    return PSX_OK;
  })
}

int psx_step(psx_sim* sim, int64_t nsteps) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  if (nsteps < 0) return set_error(PSX_ERR_INVALID, "nsteps must be >= 0");
  PSX_TRY({
    sim->impl->step(nsteps);
    return PSX_OK;
  })
}

int psx_prepare(psx_sim* sim, int64_t nsteps) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  if (nsteps < 0) return set_error(PSX_ERR_INVALID, "nsteps must be >= 0");
  PSX_TRY({
    sim->impl->prepare(nsteps);
    return PSX_OK;
  })
}

int psx_synchronize(psx_sim* sim) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  PSX_TRY({
    const uint32_t f = sim->impl->read_flags();
    if (f & psx::kFlagScanStall)
      return set_error(PSX_ERR_HIP, "cell-offset scan: a tile prefix was not published in time");
    if (f & psx::kFlagCountDrift)
      return set_error(PSX_ERR_HIP, "cell counts of a bucketing did not add up to the particle count");
    if (f & psx::kFlagNonFinite)
      return set_error(PSX_ERR_NUMERIC, "non-finite particle position or velocity produced by a step");
    return PSX_OK;
  })
}

int psx_set_stream(psx_sim* sim, void* hip_stream) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  PSX_TRY({
    sim->impl->set_stream(static_cast<hipStream_t>(hip_stream));
    return PSX_OK;
  })
}

int psx_set_tuning(psx_sim* sim, const psx_tuning* t) {
  if (!sim || !t) return set_error(PSX_ERR_INVALID, "NULL argument");
  psx_tuning c = *t;
  psx_tuning d = psx::default_tuning();
  if (c.block_integrate == 0) c.block_integrate = d.block_integrate;
  if (c.block_sort == 0) c.block_sort = d.block_sort;
  if (c.block_pairs == 0) c.block_pairs = d.block_pairs;
  for (int b : {c.block_integrate, c.block_sort, c.block_pairs})
    if (b < 64 || b > 1024 || (b % 64) != 0)
      return set_error(PSX_ERR_INVALID, "block sizes must be multiples of 64 in [64, 1024]");
  if (c.use_graph != 0 && c.use_graph != 1) return set_error(PSX_ERR_INVALID, "use_graph must be 0 or 1");
  if (sim->impl->dtype() == PSX_F64 && c.block_sort > psx::kMaxBlockSort<double>)
    return set_error(PSX_ERR_INVALID, "block_sort must be at most 512 for a binary64 simulation");
  if (c.sort_variant < 0 || c.sort_variant > 3)
    return set_error(PSX_ERR_INVALID, "sort_variant must be 0 (automatic), 1 (id blocks), 2 (dense lists) "
                                      "or 3 (incremental id blocks)");
  if (c.pair_variant != 0 && c.pair_variant != 1 && c.pair_variant != 2 && c.pair_variant != 4 &&
      c.pair_variant != 8 && c.pair_variant != 16 && c.pair_variant != psx::kPairStaged &&
      c.pair_variant != psx::kPairWave)
    return set_error(PSX_ERR_INVALID,
                     "pair_variant must be 0, 1, 2, 4, 8 or 16 (lanes per particle), 32 (LDS-staged "
                     "per workgroup) or 64 (LDS-staged per wave)");
  if (c.pair_chunk != 0 && c.pair_chunk != 2 && c.pair_chunk != 4)
    return set_error(PSX_ERR_INVALID, "pair_chunk must be 0 or 4, or 2");
  // the staged form takes kStageFactor positions per thread of LDS: at most 96 KiB
  const int vec_bytes = sim->impl->dtype() == PSX_F64 ? 32 : 16;
  if (c.pair_variant == psx::kPairStaged && psx::kStageFactor * c.block_pairs * vec_bytes > 96 * 1024)
    return set_error(PSX_ERR_INVALID, vec_bytes == 32
                                          ? "the LDS-staged pair kernel takes block_pairs up to 256 in binary64"
                                          : "the LDS-staged pair kernel takes block_pairs up to 512");
  // the wave-staged form takes two windows of 128 positions per wave of LDS: at most 64 KiB
  if (c.pair_variant == psx::kPairWave && (c.block_pairs / 64) * 2 * 128 * vec_bytes > 64 * 1024)
    return set_error(PSX_ERR_INVALID, "the wave-staged pair kernel takes block_pairs up to 512 in binary64");
  PSX_TRY({
    sim->impl->retune(c);
    return PSX_OK;
  })
}

int psx_get_tuning(const psx_sim* sim, psx_tuning* t) {
  if (!sim || !t) return set_error(PSX_ERR_INVALID, "NULL argument");
  *t = sim->impl->tuning;
  return PSX_OK;
}

int psx_get_state(psx_sim* sim, double* pos, double* vel) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  PSX_TRY({
    sim->impl->get_state(pos, vel);
    return PSX_OK;
  })
}

int psx_set_state(psx_sim* sim, const double* pos, const double* vel) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  if (!pos || !vel) return set_error(PSX_ERR_INVALID, "pos and vel must not be NULL");
  PSX_TRY({
    sim->impl->set_state(pos, vel);
    return PSX_OK;
  })
}

int psx_get_state_device(psx_sim* sim, void* pos_dev, void* vel_dev) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  PSX_TRY({
    sim->impl->get_state_device(pos_dev, vel_dev);
    return PSX_OK;
  })
}

int psx_set_state_device(psx_sim* sim, const void* pos_dev, const void* vel_dev) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  PSX_TRY({
    sim->impl->set_state_device(pos_dev, vel_dev);
    return PSX_OK;
  })
}

int64_t psx_num_particles(const psx_sim* sim) { return sim ? sim->impl->n : -1; }
int32_t psx_num_scenes(const psx_sim* sim) { return sim ? (int32_t)sim->impl->scenes.size() : -1; }
int64_t psx_num_cells(const psx_sim* sim) { return sim ? sim->impl->ncells : -1; }
int32_t psx_dtype_of(const psx_sim* sim) { return sim ? sim->impl->dtype() : -1; }

int psx_get_grid(const psx_sim* sim, int32_t scene, double* min3, double* h, int64_t* n3,
                 int64_t* cell_base) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  if (scene < 0 || scene >= (int32_t)sim->impl->scenes.size()) return set_error(PSX_ERR_INVALID, "scene out of range");
  const psx::SceneHost& s = sim->impl->scenes[(size_t)scene];
  for (int a = 0; a < 3; ++a) {
    if (min3) min3[a] = s.grid.min[a];
    if (n3) n3[a] = s.grid.n[a];
  }
  if (h) *h = s.grid.h;
  if (cell_base) *cell_base = s.cell_base;
  return PSX_OK;
}

int psx_debug_bucketing(psx_sim* sim, int32_t* cell, uint32_t* count, uint32_t* offset,
                        int32_t* perm) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  PSX_TRY({
    sim->impl->debug_bucketing(cell, count, offset, perm);
    return PSX_OK;
  })
}

int psx_set_params(psx_sim* sim, int32_t scene, const psx_params* params) {
  if (!sim || !params) return set_error(PSX_ERR_INVALID, "NULL argument");
  if (scene < 0 || scene >= (int32_t)sim->impl->scenes.size()) return set_error(PSX_ERR_INVALID, "scene out of range");
  int rc = check_params(params);
  if (rc) return rc;
  PSX_TRY({
    sim->impl->set_params(scene, *params);
    return PSX_OK;
  })
}

int psx_window_forward(psx_sim* sim, int64_t nsteps, int64_t checkpoint_every,
                       int64_t* window_id) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  if (nsteps <= 0 || checkpoint_every <= 0) return set_error(PSX_ERR_INVALID, "nsteps and checkpoint_every must be positive");
  PSX_TRY({
    const int64_t id = sim->impl->window_forward(nsteps, std::min(nsteps, checkpoint_every));
    if (window_id) *window_id = id;
    return PSX_OK;
  })
}

int psx_window_backward(psx_sim* sim, int64_t window_id, const void* dl_dpos, const void* dl_dvel,
                        void* dl_dpos0, void* dl_dvel0, double* dl_dparams) {
  if (!sim || !dl_dpos || !dl_dvel || !dl_dpos0 || !dl_dvel0)
    return set_error(PSX_ERR_INVALID, "NULL argument");
  PSX_TRY({
    sim->impl->window_backward(window_id, dl_dpos, dl_dvel, dl_dpos0, dl_dvel0, dl_dparams);
    return PSX_OK;
  })
}

int psx_window_stats(const psx_sim* sim, double* replay_ms, double* adjoint_ms) {
  if (!sim) return set_error(PSX_ERR_INVALID, "sim is NULL");
  sim->impl->window_stats(replay_ms, adjoint_ms);
  return PSX_OK;
}

int psx_statistics(psx_sim* sim, psx_stats* out) {
  if (!sim || !out) return set_error(PSX_ERR_INVALID, "NULL argument");
  PSX_TRY({
    sim->impl->statistics(out);
    return PSX_OK;
  })
}

int psx_statistics_noalloc(psx_sim* sim, psx_stats* out) {
  if (!sim || !out) return set_error(PSX_ERR_INVALID, "NULL argument");
  PSX_TRY({
    void* p = scratch_address();
    sim->impl->statistics_into(out, p, kScratchBytes);
    return PSX_OK;
  })
}

int psx_profile(psx_sim* sim, int64_t nsteps, double* kernel_ms, double* launches_per_step,
                int32_t max_kernels, int32_t* n_kernels) {
  if (!sim || !kernel_ms || !n_kernels) return set_error(PSX_ERR_INVALID, "NULL argument");
  if (nsteps <= 0 || nsteps > 100000) return set_error(PSX_ERR_INVALID, "nsteps must be in [1, 100000]");
  if (max_kernels < psx::kNumKernels) return set_error(PSX_ERR_INVALID, "max_kernels too small");
  PSX_TRY({
    sim->impl->profile(nsteps, kernel_ms, launches_per_step);
    *n_kernels = psx::kNumKernels;
    return PSX_OK;
  })
}

const char* psx_kernel_name(int32_t k) { return psx::kernel_name(k); }

const char* psx_last_error(void) { return psx::g_last_error.c_str(); }

int psx_abi_version(void) { return PSX_ABI_VERSION; }

}  // extern "C"
