// psx_adjoint.hip -- reverse-mode kernels of one step, for gradients over a window of steps.
//
// Forward step k (PhysicsEngine.cpp:38-48), with the contact sets of the step held fixed:
//   x' = x + v dt,  v' = v + g dt                         (integrate)
//   (x'', v'') = C(x', v')                                 (box walls or mesh, linear per hit set)
//   v''' = v'' + sum_{j in S(i)} f(x''_i, x''_j, v''_i, v''_j),  x''' = x''   (pairs)
// with f = -a e_s (u.d / d.d) d + k (R - |d|) d / |d|, d = x_i - x_j, u = v_i - v_j,
// a = 2 m_j / (m_i + m_j), R = r_i + r_j (ParticleModel.cpp:69-98), and S(i) the members of the 27
// cells around i's home cell (the cell of x''_i) that touch i.
//
// The adjoint of the pair term is gathered per particle: i receives the vector-Jacobian products of
// its own kicks (j in S(i)) and, with the opposite sign, of every kick f_j it contributes to
// (i in S(j)). The second set is {j : home(j) in N(bucket(i))}, read from a copy of the snapshot
// sorted by home cell, so the transpose is exact even where container pushes make home and bucket
// cells differ. No atomics: every particle writes only its own adjoint and parameter partials.
#include "psx_adjoint.hpp"
#include "psx_bucket.hpp"

namespace psx {

namespace {

__device__ inline int32_t xcd_tile_a(int32_t b, int32_t tiles) {
  const int32_t per = (tiles + 7) / 8;
  return (b % 8) * per + b / 8;
}

// Vector-Jacobian products of one kick f_i(j) with cotangent lam (the adjoint of v''' of i):
// gu = lam^T df/du, gd = lam^T df/dd, and lam . df/dk, lam . df/de_s.
template <class T>
struct PairVjp {
  T gu[3], gd[3];
  T gk, ge;
};

template <class T>
__device__ __forceinline__ void pair_vjp(const T d[3], const T u[3], T a, T R, T ks, T es,
                                         const T lam[3], PairVjp<T>& o) {
  const T D = dot3(d[0], d[1], d[2], d[0], d[1], d[2]);
  const T L = sqrt(D);
  const T w = dot3(u[0], u[1], u[2], d[0], d[1], d[2]);
  const T ld = dot3(lam[0], lam[1], lam[2], d[0], d[1], d[2]);
  const T wD = w / D;
  const T c_u = -a * es * ld / D;                  // gu = c_u d
  const T c_l = -a * es * wD + ks * (R / L - T(1));  // gd = c_l lam + c_v u + c_d d
  const T c_v = -a * es * ld / D;
  const T c_d = a * es * T(2) * w * ld / (D * D) - ks * R * ld / (L * L * L);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    o.gu[c] = c_u * d[c];
    o.gd[c] = c_l * lam[c] + c_v * u[c] + c_d * d[c];
  }
  o.gk = (R - L) * ld / L;
  o.ge = -a * wD * ld;
}

// The reference's contact test for the ordered pair (i <- j) on the snapshot.
template <class T>
__device__ __forceinline__ bool touching(const Vec4<T>& xi, const Vec4<T>& xj) {
  const T ex = xj.x - xi.x, ey = xj.y - xi.y, ez = xj.z - xi.z;
  const T dd = dot3(ex, ey, ez, ex, ey, ez);
  return sqrt(dd) < xi.w + xj.w;
}

__device__ inline void column_slices(const GridDev& g, int32_t hx, int32_t hy, int32_t hz,
                                     const uint32_t* __restrict__ offset, uint32_t* lo,
                                     uint32_t* hi) {
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

}  // namespace

// Adjoint of one step, one group of G lanes per particle of the bucket-sorted snapshot (posS,
// velS, idS, offS, cellS = bucket cell per slot; xpreS = the integrated, pre-container position
// per slot); posH/velH/idH/offH is the same snapshot sorted by home cell. lx/lv: adjoint of the
// post-step state (original order, 3 T per particle); writes the adjoint of the pre-step state to
// lx_o/lv_o and adds the parameter partials to gk/ge/gg (original order).
//
// Pair adjoint: the lanes of a group take every G-th candidate of each column slice, for i's own
// kicks (A) and for the kicks i contributes to (B); their partial sums are combined by a fixed
// xor butterfly, so the result is the same on every run and on every lane. Lane 0 then applies the
// container adjoint (box walls: J = diag; mesh: (I - n n^T) in reverse hit order for positions,
// (e / cnt) sum (I - 2 n n^T) for velocities) and the integrate adjoint.
template <class T, int G>
__global__ void __launch_bounds__(256)
    k_adj_step(int32_t ntab, const BlockInfo* __restrict__ btab,
               const SceneDev<T>* __restrict__ scenes, const Vec4<T>* __restrict__ posS,
               const Vec4<T>* __restrict__ velS, const int32_t* __restrict__ idS,
               const uint32_t* __restrict__ offS, const int32_t* __restrict__ cellS,
               const Vec4<T>* __restrict__ xpreS, const Vec4<T>* __restrict__ posH,
               const Vec4<T>* __restrict__ velH, const int32_t* __restrict__ idH,
               const uint32_t* __restrict__ offH, const TriGeo<T>* __restrict__ tgeo,
               const TriBox<T>* __restrict__ tbox, const int32_t* __restrict__ tcell_off,
               const T* __restrict__ lx, const T* __restrict__ lv, T* __restrict__ lx_o,
               T* __restrict__ lv_o, T* __restrict__ gk, T* __restrict__ ge,
               T* __restrict__ gg, const NextBucket hn, const Vec4<T>* __restrict__ hpos) {
  static_assert(G >= 1 && G <= 64 && (64 % G) == 0, "lanes per particle must divide 64");
  const int32_t tb = xcd_tile_a(blockIdx.x, ntab);
  if (tb >= ntab) return;
  const BlockInfo bi = btab[tb];
  const int32_t g = (int32_t)threadIdx.x % G;
  const int32_t p = bi.p_start + (int32_t)threadIdx.x / G;
  if (p >= bi.p_end) return;  // whole groups leave together
  const SceneDev<T>* __restrict__ s = scenes + bi.scene;
  const GridDev& grid = s->grid;
  const Vec4<T> xi = posS[p], vi = velS[p];
  const int64_t me = idS[p];
  const T li[3] = {lv[3 * me], lv[3 * me + 1], lv[3 * me + 2]};
  T ax[3] = {T(0), T(0), T(0)}, av[3] = {T(0), T(0), T(0)};
  T pk = T(0), pe = T(0);
  uint32_t lo[9], hi[9];
  // (A) i's own kicks: members of the cells around home(i), bucket-sorted
  {
    const int32_t hx = axis_index((double)xi.x, grid.min[0], grid.h, grid.n[0]);
    const int32_t hy = axis_index((double)xi.y, grid.min[1], grid.h, grid.n[1]);
    const int32_t hz = axis_index((double)xi.z, grid.min[2], grid.h, grid.n[2]);
    column_slices(grid, hx, hy, hz, offS, lo, hi);
    auto own = [&](const Vec4<T>& xj, const Vec4<T>& vj) {
      const T dv[3] = {xi.x - xj.x, xi.y - xj.y, xi.z - xj.z};
      const T uv[3] = {vi.x - vj.x, vi.y - vj.y, vi.z - vj.z};
      PairVjp<T> o;
      pair_vjp(dv, uv, T(2) * vj.w / (vi.w + vj.w), xi.w + xj.w, s->spring, s->es, li, o);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ax[c] += o.gd[c];
        av[c] += o.gu[c];
      }
      pk += o.gk;
      pe += o.ge;
    };
    // first round over all nine slices at once (lane g takes member g of each): independent loads
    // of the candidates, then of the touching ones' velocities, instead of nine round trips
    uint32_t hitA = 0;
    Vec4<T> xa[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const uint32_t q = lo[d] + (uint32_t)g;
      if (q < hi[d] && (int32_t)q != p) {
        xa[d] = posS[q];
        hitA |= 1u << d;
      }
    }
#pragma unroll
    for (int d = 0; d < 9; ++d)
      if ((hitA >> d & 1u) && !touching(xi, xa[d])) hitA &= ~(1u << d);
    Vec4<T> va[9];
#pragma unroll
    for (int d = 0; d < 9; ++d)
      if (hitA >> d & 1u) va[d] = velS[lo[d] + (uint32_t)g];
#pragma unroll
    for (int d = 0; d < 9; ++d)
      if (hitA >> d & 1u) own(xa[d], va[d]);
    // slices with more than G members: the rest in rounds of G
    for (int d = 0; d < 9; ++d)
      for (uint32_t q = lo[d] + (uint32_t)(g + G); q < hi[d]; q += G) {
        if ((int32_t)q == p) continue;
        const Vec4<T> xj = posS[q];
        if (!touching(xi, xj)) continue;
        own(xj, velS[q]);
      }
  }
  // (B) kicks i contributes to: particles whose home cell is around bucket(i), home-sorted
  {
    const int32_t lc = cellS[p] - grid.cell_base;
    const int32_t bz = lc % grid.n[2], rest = lc / grid.n[2];
    const int32_t by = rest % grid.n[1], bx = rest / grid.n[1];
    column_slices(grid, bx, by, bz, offH, lo, hi);
    auto other = [&](const Vec4<T>& xj, const Vec4<T>& vj, int64_t j) {
      const T lj[3] = {lv[3 * j], lv[3 * j + 1], lv[3 * j + 2]};
      const T dv[3] = {xj.x - xi.x, xj.y - xi.y, xj.z - xi.z};
      const T uv[3] = {vj.x - vi.x, vj.y - vi.y, vj.z - vi.z};
      PairVjp<T> o;
      pair_vjp(dv, uv, T(2) * vi.w / (vj.w + vi.w), xj.w + xi.w, s->spring, s->es, lj, o);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ax[c] -= o.gd[c];
        av[c] -= o.gu[c];
      }
    };
    // first round over all nine slices at once, as in (A)
    uint32_t hitB = 0;
    Vec4<T> xb[9];
    int32_t jb[9];
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const uint32_t q = lo[d] + (uint32_t)g;
      if (q < hi[d]) {
        jb[d] = idH[q];
        xb[d] = posH[q];
        hitB |= 1u << d;
      }
    }
#pragma unroll
    for (int d = 0; d < 9; ++d)
      if ((hitB >> d & 1u) && ((int64_t)jb[d] == me || !touching(xb[d], xi))) hitB &= ~(1u << d);
    Vec4<T> vb[9];
#pragma unroll
    for (int d = 0; d < 9; ++d)
      if (hitB >> d & 1u) vb[d] = velH[lo[d] + (uint32_t)g];
#pragma unroll
    for (int d = 0; d < 9; ++d)
      if (hitB >> d & 1u) other(xb[d], vb[d], jb[d]);
    for (int d = 0; d < 9; ++d)
      for (uint32_t q = lo[d] + (uint32_t)(g + G); q < hi[d]; q += G) {
        const int64_t j = idH[q];
        if (j == me) continue;
        const Vec4<T> xj = posH[q];
        if (!touching(xj, xi)) continue;
        other(xj, velH[q], j);
      }
  }
  // fixed-order combination of the group's partial sums (identical on every lane)
#pragma unroll
  for (int off = G / 2; off >= 1; off >>= 1) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ax[c] += __shfl_xor(ax[c], off, 64);
      av[c] += __shfl_xor(av[c], off, 64);
    }
    pk += __shfl_xor(pk, off, 64);
    pe += __shfl_xor(pe, off, 64);
  }
  if (g != 0) return;
  // adjoint of (x'', v''): the post-step cotangent plus the pair terms
  T lxs[3], lvs[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    lxs[c] = lx[3 * me + c] + ax[c];
    lvs[c] = li[c] + av[c];
  }
  // container adjoint at x' (the forward's arithmetic, so the hit sets are the forward's)
  const Vec4<T> x = xpreS[p];
  if (s->mesh < 0) {
    const T r = x.w;
    const T px[3] = {x.x, x.y, x.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const bool hmin = px[a] - r < s->bmin[a];
      const bool hmax = px[a] + r > s->bmax[a];
      if (hmin || hmax) lxs[a] = T(0);
      if (hmin) lvs[a] = -lvs[a] * s->emin[a];
      if (hmax) lvs[a] = -lvs[a] * s->emax[a];
    }
  } else {
    const int32_t ix = axis_index((double)x.x, grid.min[0], grid.h, grid.n[0]);
    const int32_t iy = axis_index((double)x.y, grid.min[1], grid.h, grid.n[1]);
    const int32_t iz = axis_index((double)x.z, grid.min[2], grid.h, grid.n[2]);
    const int32_t tc = s->tcell_base + (ix * grid.n[1] + iy) * grid.n[2] + iz;
    const int32_t k0 = tcell_off[tc], k1 = tcell_off[tc + 1];
    const Vec3<T> P = {x.x, x.y, x.z};
    const T r = x.w;
    // velocity: v'' = (e / cnt) sum_h (I - 2 n n^T) v'  ->  same map on the cotangent
    T sv[3] = {T(0), T(0), T(0)};
    int32_t cnt = 0;
    for (int32_t k = k0; k < k1; ++k) {
      const TriBox<T> bx = tbox[k];
      if (P.x - r > bx.hi[0] || P.x + r < bx.lo[0] || P.y - r > bx.hi[1] || P.y + r < bx.lo[1] ||
          P.z - r > bx.hi[2] || P.z + r < bx.lo[2])
        continue;
      const TriGeo<T> t = tgeo[bx.tri];
      if (!(dist_to(P, closest_on_triangle(t, P)) < r)) continue;
      const T ln = dot3(lvs[0], lvs[1], lvs[2], t.n[0], t.n[1], t.n[2]);
#pragma unroll
      for (int c = 0; c < 3; ++c) sv[c] += lvs[c] - T(2) * ln * t.n[c];
      ++cnt;
    }
    if (cnt > 0) {
      // position: y = P_m ... P_1 x' + const, P_h = I - n_h n_h^T  ->  apply P_m first
      for (int32_t k = k1 - 1; k >= k0; --k) {
        const TriBox<T> bx = tbox[k];
        if (P.x - r > bx.hi[0] || P.x + r < bx.lo[0] || P.y - r > bx.hi[1] ||
            P.y + r < bx.lo[1] || P.z - r > bx.hi[2] || P.z + r < bx.lo[2])
          continue;
        const TriGeo<T> t = tgeo[bx.tri];
        if (!(dist_to(P, closest_on_triangle(t, P)) < r)) continue;
        const T ln = dot3(lxs[0], lxs[1], lxs[2], t.n[0], t.n[1], t.n[2]);
#pragma unroll
        for (int c = 0; c < 3; ++c) lxs[c] -= ln * t.n[c];
      }
      const T f = s->ew / (T)cnt;
#pragma unroll
      for (int c = 0; c < 3; ++c) lvs[c] = f * sv[c];
    }
  }
  // integrate: lambda_x = lambda_x', lambda_v = lambda_v' + dt lambda_x'; d/dg_z += dt lambda_v'_z
  const T dt = s->dt;
  gk[me] += pk;
  ge[me] += pe;
  gg[me] += dt * lvs[2];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    lx_o[3 * me + c] = lxs[c];
    lv_o[3 * me + c] = lvs[c] + dt * lxs[c];
  }
  // the hash pass of the snapshot the next adjoint step buckets by home cell
  if (hn.count) {
    const int32_t c = bucket_entry(*s, hpos[p], hn.id[p], p, hn.blocks != 0, hn.cell_of, hn.count,
                                   hn.cell_ids, hn.id_stride, hn.ovf, hn.ovf_ctr);
    tile_count_wave(hn.tiles, c);
  }
}

// Deterministic sum of three per-particle partial arrays (fixed order: strided lane sums, then a
// fixed tree), into out[0..2] in double.
template <class T>
__global__ void __launch_bounds__(1024)
    k_reduce3(int32_t n, const T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ c,
              double* __restrict__ out) {
  __shared__ double sh[3][1024];
  double s0 = 0, s1 = 0, s2 = 0;
  for (int32_t i = threadIdx.x; i < n; i += 1024) {
    s0 += (double)a[i];
    s1 += (double)b[i];
    s2 += (double)c[i];
  }
  sh[0][threadIdx.x] = s0;
  sh[1][threadIdx.x] = s1;
  sh[2][threadIdx.x] = s2;
  __syncthreads();
  for (int w = 512; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) out[k] = sh[k][0];
}

template <class T>
__global__ void __launch_bounds__(256) k_fill(int64_t n, T* __restrict__ a, T value) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = value;
}

static inline int32_t xgrid(int32_t tiles) { return ((tiles + 7) / 8) * 8; }

template <class T>
hipError_t launch_adj_step(hipStream_t st, int ntab, const BlockInfo* btab,
                           const SceneDev<T>* scenes, const Vec4<T>* posS, const Vec4<T>* velS,
                           const int32_t* idS, const uint32_t* offS, const int32_t* cellS,
                           const Vec4<T>* xpreS, const Vec4<T>* posH, const Vec4<T>* velH,
                           const int32_t* idH, const uint32_t* offH, const TriGeo<T>* tgeo,
                           const TriBox<T>* tbox, const int32_t* tcell_off, const T* lx,
                           const T* lv, T* lx_o, T* lv_o, T* gk, T* ge, T* gg,
                           const NextBucket& hn, const Vec4<T>* hpos) {
  hipLaunchKernelGGL((k_adj_step<T, kAdjLanes>), dim3(xgrid(ntab)), dim3(256), 0, st, ntab, btab,
                     scenes, posS, velS, idS, offS, cellS, xpreS, posH, velH, idH, offH, tgeo,
                     tbox, tcell_off, lx, lv, lx_o, lv_o, gk, ge, gg, hn, hpos);
  return hipGetLastError();
}

template <class T>
__global__ void __launch_bounds__(256)
    k_copy2(int64_t n, const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ a_o,
            T* __restrict__ b_o) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    a_o[i] = a[i];
    b_o[i] = b[i];
  }
}

template <class T>
hipError_t launch_copy2(hipStream_t st, int64_t n, const T* a, const T* b, T* a_o, T* b_o) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy2<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, b,
                     a_o, b_o);
  return hipGetLastError();
}

template <class T>
hipError_t launch_reduce3(hipStream_t st, int32_t n, const T* a, const T* b, const T* c,
                          double* out) {
  hipLaunchKernelGGL(k_reduce3<T>, dim3(1), dim3(1024), 0, st, n, a, b, c, out);
  return hipGetLastError();
}

template <class T>
hipError_t launch_fill(hipStream_t st, int64_t n, T* a, T value) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, a, value);
  return hipGetLastError();
}

#define PSX_ADJ_INSTANTIATE(T)                                                                   \
  template hipError_t launch_adj_step<T>(hipStream_t, int, const BlockInfo*, const SceneDev<T>*, \
                                         const Vec4<T>*, const Vec4<T>*, const int32_t*,         \
                                         const uint32_t*, const int32_t*, const Vec4<T>*,        \
                                         const Vec4<T>*, const Vec4<T>*, const int32_t*,         \
                                         const uint32_t*, const TriGeo<T>*, const TriBox<T>*,    \
                                         const int32_t*, const T*, const T*, T*, T*, T*, T*, T*, \
                                         const NextBucket&, const Vec4<T>*); \
  template hipError_t launch_copy2<T>(hipStream_t, int64_t, const T*, const T*, T*, T*);         \
  template hipError_t launch_reduce3<T>(hipStream_t, int32_t, const T*, const T*, const T*,      \
                                        double*);                                                \
  template hipError_t launch_fill<T>(hipStream_t, int64_t, T*, T);

PSX_ADJ_INSTANTIATE(float)
PSX_ADJ_INSTANTIATE(double)

}  // namespace psx
