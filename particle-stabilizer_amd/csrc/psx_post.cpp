// psx_post.cpp -- post-run outputs on the host: the overlap post-process and the sphere OBJ export.
//
// Both run once after a simulation (ParticleStabilizerMain.cpp:275-296), on the final state copied
// to the host; neither is on the per-step path.
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "psx.h"
#include "psx_host.hpp"

namespace {

using psx::set_error;

// Uniform grid of cell size h over a set of particles, for "everything within h of a point"
// queries: a pair closer than h lies in the 3x3x3 cells around either particle. Cell coordinates
// are taken relative to the lowest coordinate; callers make sure the extent stays below 2^30 cells.
struct CellGrid {
  double h = 1.0, org[3] = {0, 0, 0};
  std::unordered_map<uint64_t, std::vector<int64_t>> cells;

  static uint64_t key(int64_t a, int64_t b, int64_t c) {
    return (uint64_t)(a & 0x1fffff) | ((uint64_t)(b & 0x1fffff) << 21) | ((uint64_t)(c & 0x1fffff) << 42);
  }
  void coords(const double* x, int64_t* c) const {
    for (int a = 0; a < 3; ++a) c[a] = (int64_t)std::floor((x[a] - org[a]) / h);
  }
  void insert(const double* x, int64_t p) {
    int64_t c[3];
    coords(x, c);
    cells[key(c[0], c[1], c[2])].push_back(p);
  }
  template <class F>
  void near(const double* x, F&& f) const {
    int64_t c[3];
    coords(x, c);
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          auto it = cells.find(key(c[0] + dx, c[1] + dy, c[2] + dz));
          if (it == cells.end()) continue;
          for (int64_t q : it->second) f(q);
        }
  }
};

// Cell size for pairs of sum of radii at most 2 r_max, with room for rounding; false when the
// particles span too many cells for the grid (then the callers test every pair).
bool grid_for(int64_t n, const double* pos, double rmax, CellGrid& g) {
  if (!(rmax > 0.0)) return false;
  g.h = 2.0 * rmax * (1.0 + 1e-6);
  for (int a = 0; a < 3; ++a) {
    double lo = pos[a], hi = pos[a];
    for (int64_t i = 1; i < n; ++i) {
      lo = std::min(lo, pos[3 * i + a]);
      hi = std::max(hi, pos[3 * i + a]);
    }
    g.org[a] = lo;
    if (!((hi - lo) / g.h < 1073741824.0)) return false;
  }
  return true;
}

// Number of particles i with a partner j > i such that r_i + r_j - |x_i - x_j| > 0, and the
// largest such value (ParticleModel.cpp:227-243; |d| = sqrt((dx dx + dy dy) + dz dz)).
void overlap_check(int64_t n, const double* pos, const double* radius, double rmax,
                   int64_t* n_collided, double* max_overlap) {
  int64_t count = 0;
  double best = 0.0;
  auto test = [&](int64_t i, int64_t j, bool& hit) {
    const double d0 = pos[3 * i] - pos[3 * j], d1 = pos[3 * i + 1] - pos[3 * j + 1],
                 d2 = pos[3 * i + 2] - pos[3 * j + 2];
    const double cd = radius[i] + radius[j] - std::sqrt((d0 * d0 + d1 * d1) + d2 * d2);
    if (cd > 0.0) {
      hit = true;
      best = std::max(best, cd);
    }
  };
  CellGrid g;
  const bool use_grid = grid_for(n, pos, rmax, g);
  if (use_grid)
    for (int64_t i = 0; i < n; ++i) g.insert(pos + 3 * i, i);
  for (int64_t i = 0; i < n; ++i) {
    bool hit = false;
    if (use_grid) {
      g.near(pos + 3 * i, [&](int64_t j) {
        if (j > i) test(i, j, hit);
      });
    } else if (rmax > 0.0) {
      for (int64_t j = i + 1; j < n; ++j) test(i, j, hit);
    }
    if (hit) ++count;
  }
  *n_collided = count;
  *max_overlap = best;
}

// ParticleModel::resolveOverlaps (ParticleModel.cpp:154-225) with the same ranking, the same
// partner order and the same arithmetic, so that the lifted positions are bit for bit the
// reference's; it differs only in what it skips. The reference, for rank i, finds
// start = the first rank j < i with c_i - r_i <= c_j + r_j (c: coordinate along the axis; 0 when
// none) and tests every rank in [start, i) for overlap in rank order, lifting particle i on each
// overlapping one; a lift that moves c_i strictly between the coordinates of ranks k-1 and k
// (first such k > i) re-sorts ranks [i, k] and processes rank i again. Facts used here:
//  * ranks below i never change again (re-sorts and lifts touch rank i and above), so the
//    particles already processed form a fixed set. Their c_j + r_j sit in a max tree over ranks,
//    which gives start as "leftmost rank whose value is >= c_i - r_i" in O(log n); and they sit
//    in a grid, so that only those within r_i + r_max of particle i, the only ones that can
//    overlap it, are tested, in rank order: the reference's sequence of tests minus tests that
//    cannot succeed;
//  * ranks above i are in coordinate order (true at the start; a re-sort leaves [i, n) sorted),
//    so the re-sort bound k is an upper-bound search, and the re-sort itself, when the ranks it
//    covers hold no equal coordinates, has exactly one result: particle i moves to rank k - 1
//    (with equal coordinates std::sort is called on the range, as in the reference).
class OverlapResolver {
 public:
  OverlapResolver(int64_t n, double* pos, const double* radius, int axis, double rmax)
      : n_(n), pos_(pos), r_(radius), axis_(axis), rmax_(rmax), rank_((size_t)n), key_((size_t)n),
        rank_of_((size_t)n, -1) {
    leaves_ = 1;
    while (leaves_ < n) leaves_ *= 2;
    top_.assign(2 * (size_t)leaves_, -HUGE_VAL);
  }

  void run() {
    for (int64_t i = 0; i < n_; ++i) rank_[(size_t)i] = i;
    std::sort(rank_.begin(), rank_.end(), [&](int64_t a, int64_t b) { return c(a) < c(b); });
    for (int64_t k = 0; k < n_; ++k) key_[(size_t)k] = c(rank_[(size_t)k]);
    use_grid_ = grid_for(n_, pos_, rmax_, grid_);
    for (int64_t i = 0; i < n_; ++i) {
      if (process(i)) {
        --i;  // ParticleModel.cpp:218: rank i holds another particle now
        continue;
      }
      finalize(i);
    }
  }

 private:
  double c(int64_t p) const { return pos_[3 * p + axis_]; }

  // ParticleModel.cpp:178-185: the leftmost processed rank with c_j + r_j >= c_i - r_i, else 0.
  int64_t start_rank(int64_t ip) const {
    const double lo = c(ip) - r_[ip];
    if (!(lo <= top_[1])) return 0;
    size_t node = 1;
    while (node < (size_t)leaves_) node = (lo <= top_[2 * node]) ? 2 * node : 2 * node + 1;
    return (int64_t)(node - (size_t)leaves_);
  }

  // Ranks in [from, i) of processed particles that may overlap particle ip, ascending.
  void candidates(int64_t i, int64_t ip, int64_t from) {
    cand_.clear();
    if (use_grid_) {
      grid_.near(pos_ + 3 * ip, [&](int64_t q) {
        if (rank_of_[(size_t)q] >= from) cand_.push_back(rank_of_[(size_t)q]);
      });
      std::sort(cand_.begin(), cand_.end());
    } else {
      for (int64_t j = from; j < i; ++j) cand_.push_back(j);
    }
  }

  // ParticleModel.cpp:207-214 on a sorted [i + 1, n): the first k > i with
  // c(rank[k-1]) < x < c(rank[k]), or i when there is none. key_[k] = c(rank[k]) for k > i.
  int64_t resort_bound(int64_t i, double x) const {
    const int64_t m = std::upper_bound(key_.begin() + i + 1, key_.end(), x) - key_.begin();
    if (m >= i + 2 && m < n_ && key_[(size_t)(m - 1)] < x) return m;
    return i;
  }

  // ParticleModel.cpp:216-220
  void resort(int64_t i, int64_t upto) {
    key_[(size_t)i] = c(rank_[(size_t)i]);
    bool equal = false;
    for (int64_t k = i + 1; k + 1 < upto; ++k) equal |= key_[(size_t)k] == key_[(size_t)(k + 1)];
    if (equal) {
      std::sort(rank_.begin() + i, rank_.begin() + upto + 1, [&](int64_t a, int64_t b) { return c(a) < c(b); });
      for (int64_t k = i; k <= upto; ++k) key_[(size_t)k] = c(rank_[(size_t)k]);
    } else {
      std::rotate(rank_.begin() + i, rank_.begin() + i + 1, rank_.begin() + upto);
      std::rotate(key_.begin() + i, key_.begin() + i + 1, key_.begin() + upto);
    }
  }

  // ParticleModel.cpp:187-222 for rank i; true when the ranking was re-sorted.
  bool process(int64_t i) {
    const int64_t ip = rank_[(size_t)i];
    double* xi = pos_ + 3 * ip;
    const int o0 = axis_ == 0 ? 1 : 0, o1 = axis_ == 2 ? 1 : 2;
    int64_t from = start_rank(ip);
    for (;;) {
      candidates(i, ip, from);
      bool lifted = false;
      for (int64_t j : cand_) {
        const int64_t jp = rank_[(size_t)j];
        const double* xj = pos_ + 3 * jp;
        const double sr = r_[ip] + r_[jp];
        const double sr2 = sr * sr;
        const double dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
        if (!(sr2 - ((dx * dx + dy * dy) + dz * dz) > 0.0)) continue;
        const double d0 = xi[o0] - xj[o0], d1 = xi[o1] - xj[o1];
        xi[axis_] = xj[axis_] + std::sqrt(((sr * sr) - d0 * d0) - d1 * d1);
        // a lift past the extent the grid was sized for switches to testing every rank
        if (use_grid_ && !((xi[axis_] - grid_.org[axis_]) / grid_.h < 1073741824.0)) use_grid_ = false;
        const int64_t upto = resort_bound(i, xi[axis_]);
        if (upto != i) {
          resort(i, upto);
          return true;
        }
        from = j + 1;
        lifted = true;
        break;
      }
      if (!lifted) return false;
    }
  }

  void finalize(int64_t i) {
    const int64_t ip = rank_[(size_t)i];
    rank_of_[(size_t)ip] = i;
    key_[(size_t)i] = c(ip);
    grid_.insert(pos_ + 3 * ip, ip);
    size_t node = (size_t)leaves_ + (size_t)i;
    top_[node] = c(ip) + r_[ip];  // the right-hand side of ParticleModel.cpp:181
    for (node /= 2; node >= 1; node /= 2) top_[node] = std::max(top_[2 * node], top_[2 * node + 1]);
  }

  int64_t n_;
  double* pos_;
  const double* r_;
  int axis_;
  double rmax_;
  std::vector<int64_t> rank_;
  std::vector<double> key_;  // key_[k] = coordinate at rank k (kept for k > current rank)
  std::vector<int64_t> rank_of_, cand_;
  int64_t leaves_ = 1;
  std::vector<double> top_;  // max tree over ranks of c_j + r_j, -inf where not yet processed
  CellGrid grid_;
  bool use_grid_ = false;
};

bool make_parent_dirs(const std::string& path) {
  const size_t slash = path.find_last_of('/');
  if (slash == std::string::npos || slash == 0) return true;
  const std::string dir = path.substr(0, slash);
  std::string cur;
  for (size_t k = 0; k <= dir.size(); ++k) {
    if (k == dir.size() || dir[k] == '/') {
      if (!cur.empty() && ::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return false;
    }
    if (k < dir.size()) cur.push_back(dir[k]);
  }
  return true;
}

}  // namespace

extern "C" {

// ParticleModel::resolveOverlaps (ParticleModel.cpp:154-252); bit-identical lifted positions (see
// OverlapResolver), then the overlap check the reference logs. Positions and radii must be finite
// and radii non-negative (the reference's std::sort has no defined result with NaN coordinates).
int psx_resolve_overlaps(int64_t n, double* pos, const double* radius, int32_t axis,
                         int64_t* n_collided, double* max_overlap) {
  if (n < 0 || (n > 0 && (!pos || !radius))) return set_error(PSX_ERR_INVALID, "NULL argument");
  if (axis < 0 || axis > 2) return set_error(PSX_ERR_INVALID, "axis must be 0, 1 or 2");
  double rmax = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    if (!std::isfinite(pos[3 * i]) || !std::isfinite(pos[3 * i + 1]) || !std::isfinite(pos[3 * i + 2]) ||
        !std::isfinite(radius[i]) || radius[i] < 0.0)
      return set_error(PSX_ERR_INVALID, "positions and radii must be finite, radii non-negative");
    rmax = std::max(rmax, radius[i]);
  }
  try {
    if (rmax > 0.0) OverlapResolver(n, pos, radius, axis, rmax).run();
    int64_t nc = 0;
    double mo = 0.0;
    overlap_check(n, pos, radius, rmax, &nc, &mo);
    if (n_collided) *n_collided = nc;
    if (max_overlap) *max_overlap = mo;
    return PSX_OK;
  } catch (const std::bad_alloc&) {
    return set_error(PSX_ERR_ALLOC, "out of host memory");
  }
}

// ObjData::write (ObjData.cpp:8-140): every sphere as a UV sphere of 2 (ndivs-1)(2 ndivs-1)
// triangles, ndivs = max((int)sqrt(1e6 / (6 + n)) + 1, 6), in binary32 with the library sinf / cosf;
// "o", then all "v" (radius * unit vertex + centre), the shared "vn" of one sphere, then "f a//na
// b//nb c//nc" per triangle; numbers in the default stream format (%g, 6 significant digits).
int psx_write_obj_spheres(const char* path, int64_t n, const double* pos, const double* radius) {
  if (!path || n < 0 || (n > 0 && (!pos || !radius))) return set_error(PSX_ERR_INVALID, "NULL argument");
  if (!make_parent_dirs(path)) return set_error(PSX_ERR_IO, std::string("cannot create the directory of ") + path);
  const int ndivs = std::max((int)std::sqrt(1000000.0 / (6.0 + (double)n)) + 1, 6);
  const int nth = ndivs, nph = 2 * ndivs;
  const float dth = (float)(M_PI / (double)(float)(nth - 1));
  const float dph = (float)(2 * M_PI / (double)(float)(nph - 1));
  std::vector<float> u;  // 9 floats per triangle: unit vertices, also the normals
  for (int it = 0; it < nth - 1; ++it)
    for (int ip = 0; ip < nph - 1; ++ip) {
      const float th = (float)it * dth, th1 = th + dth;
      const float ph = (float)ip * dph, ph1 = ph + dph;
      const float q0[3] = {std::sin(th) * std::cos(ph), std::sin(th) * std::sin(ph), std::cos(th)};
      const float q1[3] = {std::sin(th1) * std::cos(ph), std::sin(th1) * std::sin(ph), std::cos(th1)};
      const float q2[3] = {std::sin(th1) * std::cos(ph1), std::sin(th1) * std::sin(ph1), std::cos(th1)};
      const float q3[3] = {std::sin(th) * std::cos(ph1), std::sin(th) * std::sin(ph1), std::cos(th)};
      for (const float* q : {q0, q1, q2, q2, q3, q0}) u.insert(u.end(), q, q + 3);
    }
  const size_t nf = u.size() / 9;
  FILE* f = std::fopen(path, "w");
  if (!f) return set_error(PSX_ERR_IO, std::string("cannot open ") + path);
  std::vector<char> buf(1 << 20);
  std::setvbuf(f, buf.data(), _IOFBF, buf.size());
  std::fputs("o Generated particles\n", f);
  for (int64_t p = 0; p < n; ++p) {
    const float cx = (float)pos[3 * p], cy = (float)pos[3 * p + 1], cz = (float)pos[3 * p + 2];
    const float r = (float)radius[p];
    for (size_t k = 0; k < 3 * nf; ++k) {
      const float* q = &u[3 * k];
      const float vx = r * q[0] + cx, vy = r * q[1] + cy, vz = r * q[2] + cz;
      std::fprintf(f, "v %g %g %g\n", (double)vx, (double)vy, (double)vz);
    }
  }
  for (size_t k = 0; k < 3 * nf; ++k)
    std::fprintf(f, "vn %g %g %g\n", (double)u[3 * k], (double)u[3 * k + 1], (double)u[3 * k + 2]);
  for (int64_t p = 0; p < n; ++p) {
    const size_t base = (size_t)p * nf * 3;
    for (size_t t = 0; t < nf; ++t) {
      const size_t fo = 3 * t, o = base + fo;
      std::fprintf(f, "f %zu//%zu %zu//%zu %zu//%zu\n", o + 1, fo + 1, o + 2, fo + 2, o + 3, fo + 3);
    }
  }
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok) return set_error(PSX_ERR_IO, std::string("write failed: ") + path);
  return PSX_OK;
}

}  // extern "C"
