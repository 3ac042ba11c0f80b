// psx_adjoint.hpp -- launchers of the reverse-mode kernels (psx_adjoint.hip).
#pragma once

#include "psx_device.hpp"

namespace psx {

template <class T>
hipError_t launch_adj_pairs(hipStream_t st, int ntab, const BlockInfo* btab,
                            const SceneDev<T>* scenes, const Vec4<T>* posS, const Vec4<T>* velS,
                            const int32_t* idS, const uint32_t* offS, const int32_t* cellS,
                            const Vec4<T>* posH, const Vec4<T>* velH, const int32_t* idH,
                            const uint32_t* offH, const T* lx, const T* lv, T* lx_o, T* lv_o,
                            T* gk, T* ge);

template <class T>
hipError_t launch_adj_container(hipStream_t st, int ntab, const BlockInfo* btab,
                                const SceneDev<T>* scenes, const Vec4<T>* pos, const Vec4<T>* vel,
                                const int32_t* id, const TriGeo<T>* tgeo, const TriBox<T>* tbox,
                                const int32_t* tcell_off, const T* lx, const T* lv, T* lx_o,
                                T* lv_o, T* gg);

template <class T>
hipError_t launch_reduce3(hipStream_t st, int32_t n, const T* a, const T* b, const T* c,
                          double* out);

template <class T>
hipError_t launch_fill(hipStream_t st, int64_t n, T* a, T value);

}  // namespace psx
