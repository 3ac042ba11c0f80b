// psx_kernels.hpp -- launchers of the per-step kernels (definitions in psx_kernels.hip).
#pragma once

#include "psx_device.hpp"

namespace psx {

enum : int { kModeIntegrate = 1, kModeContainer = 2, kModeCheck = 4 };
enum : uint32_t { kFlagNonFinite = 1u };

// Events that time the next kernel launch from its dispatch timestamps (consumed by the launch).
void set_launch_events(hipEvent_t start, hipEvent_t stop);

constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 16;
constexpr int kScanTile = kScanThreads * kScanPerThread;  // cells per scan tile

template <class T>
hipError_t launch_integrate(hipStream_t st, int ntab, int block, const BlockInfo* btab,
                            const SceneDev<T>* scenes, const Vec4<T>* pos, const Vec4<T>* vel,
                            Vec4<T>* pos_o, Vec4<T>* vel_o, int32_t* cell_of, uint32_t* slot,
                            uint32_t* count,
                            const TriGeo<T>* tgeo, const TriBox<T>* tbox,
                            const int32_t* tcell_off, uint32_t* flags, int mode);

hipError_t launch_scan_reduce(hipStream_t st, const uint32_t* count, uint32_t* tile_sums,
                              int32_t ntiles);
hipError_t launch_scan_tiles(hipStream_t st, uint32_t* tile_sums, uint32_t* tile_pre,
                             int32_t ntiles);
hipError_t launch_scan_down(hipStream_t st, uint32_t* count, uint32_t* offset,
                            const uint32_t* tile_pre, int32_t ntiles);

hipError_t launch_scatter_ids(hipStream_t st, int block, int32_t n, const int32_t* cell_of,
                              const uint32_t* slot, const uint32_t* offset, const int32_t* id,
                              int32_t* tmp);

template <class T>
hipError_t launch_rank_gather(hipStream_t st, int block, int32_t n, const int32_t* cell_of,
                              const uint32_t* offset, const int32_t* tmp, const Vec4<T>* pos,
                              const Vec4<T>* vel, const int32_t* id, Vec4<T>* pos_o,
                              Vec4<T>* vel_o, int32_t* id_o, int32_t* cell_o);

template <class T>
hipError_t launch_pairs(hipStream_t st, int ntab, int block, const BlockInfo* btab,
                        const SceneDev<T>* scenes, const Vec4<T>* pos, const Vec4<T>* vsn,
                        const uint32_t* offset, Vec4<T>* vout);

template <class T>
hipError_t launch_unsort(hipStream_t st, int32_t n, const Vec4<T>* pos, const Vec4<T>* vel,
                         const int32_t* id, T* pos3, T* vel3);

template <class T>
hipError_t launch_sort_in(hipStream_t st, int32_t n, Vec4<T>* pos, Vec4<T>* vel, const int32_t* id,
                          const T* pos3, const T* vel3);

}  // namespace psx
