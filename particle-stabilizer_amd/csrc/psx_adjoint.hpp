// psx_adjoint.hpp -- launchers of the reverse-mode kernels (psx_adjoint.hip).
#pragma once

#include "psx_kernels.hpp"

namespace psx {

// Lanes per particle of the adjoint step kernel (its workgroups hold 256 / kAdjLanes particles).
// Measured on MI355X over the 16,384-sphere calibration window, per adjoint step: 1 lane 36.6 us,
// 2 lanes 27.6, 4 lanes 22.5, 8 lanes 23.6, 16 lanes 35.3.
constexpr int kAdjLanes = 4;
constexpr int kAdjPerBlock = 256 / kAdjLanes;

// Adjoint of one step (pair, container and integrate terms) for the bucket-sorted snapshot of the
// step: btab holds kAdjPerBlock particles per workgroup. xpreS: integrated, pre-container positions
// of the snapshot's slots. With hn.count set, the kernel also does the hash pass (cell, count, id
// block, tile count) of the snapshot hpos (ids hn.id) that the next adjoint step buckets by home
// cell.
template <class T>
hipError_t launch_adj_step(hipStream_t st, int ntab, const BlockInfo* btab,
                           const SceneDev<T>* scenes, const Vec4<T>* posS, const Vec4<T>* velS,
                           const int32_t* idS, const uint32_t* offS, const int32_t* cellS,
                           const Vec4<T>* xpreS, const Vec4<T>* posH, const Vec4<T>* velH,
                           const int32_t* idH, const uint32_t* offH, const TriGeo<T>* tgeo,
                           const TriBox<T>* tbox, const int32_t* tcell_off, const T* lx,
                           const T* lv, T* lx_o, T* lv_o, T* gk, T* ge, T* gg,
                           const NextBucket& hn, const Vec4<T>* hpos);

// a_o[i] = a[i], b_o[i] = b[i] for i < n.
template <class T>
hipError_t launch_copy2(hipStream_t st, int64_t n, const T* a, const T* b, T* a_o, T* b_o);

template <class T>
hipError_t launch_reduce3(hipStream_t st, int32_t n, const T* a, const T* b, const T* c,
                          double* out);

template <class T>
hipError_t launch_fill(hipStream_t st, int64_t n, T* a, T value);

}  // namespace psx
