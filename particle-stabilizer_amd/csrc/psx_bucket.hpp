// psx_bucket.hpp -- device functions of the hash pass shared by the kernels of the forward step
// (psx_kernels.hip) and of the adjoint (psx_adjoint.hip).
#pragma once

#include "psx_kernels.hpp"

namespace psx {

__device__ inline int lane_id() { return threadIdx.x & 63; }

// The grid cell of position x (cells numbered z-fastest).
template <class T>
__device__ __forceinline__ int32_t grid_cell(const SceneDev<T>& s, const Vec4<T>& x) {
  const GridDev& g = s.grid;
  const int32_t ix = axis_index((double)x.x, g.min[0], g.h, g.n[0]);
  const int32_t iy = axis_index((double)x.y, g.min[1], g.h, g.n[1]);
  const int32_t iz = axis_index((double)x.z, g.min[2], g.h, g.n[2]);
  return g.cell_base + (ix * g.n[1] + iy) * g.n[2] + iz;
}

// Files id me into cell c's id slots at arrival slot `slot` (or onto the cell's overflow chain).
__device__ __forceinline__ void file_id(int32_t me, int32_t c, uint32_t slot,
                                        int32_t* __restrict__ cell_ids, uint32_t stride,
                                        int2* __restrict__ ovf, uint32_t* __restrict__ ovf_ctr) {
  if (slot < (uint32_t)kCellIds) {
    cell_ids[(size_t)slot * stride + (uint32_t)c] = me;
  } else {  // at most n entries in all
    const uint32_t o = atomicAdd(ovf_ctr, 1u);
    const int32_t prev = atomicExch(cell_ids + (size_t)kCellIds * stride + (uint32_t)c, (int32_t)o);
    ovf[o] = make_int2(me, prev);
  }
}

// ParticleModel::updateBuckets (ParticleModel.cpp:49-62) for particle p (id me) at its
// pre-container position x: its cell and one more member counted there. With id blocks the
// particle's arrival slot (the counter's old value) places its id in the cell's id slots, slot k of
// cell c at cell_ids[k * stride + c] (slot-major: the first arrivals of consecutive cells, which
// consecutive particles are, lie side by side). An arrival past kCellIds takes an overflow entry
// (id, previous entry of its cell) and makes it the head of its cell's chain in the slot plane
// kCellIds: the cell's last count - kCellIds entries are exactly its overflow members, so a reader
// follows the chain that many steps from the head and never needs the chain cleared.
// For the dense lists the count alone is kept (an atomic whose result is not waited for) and the
// scatter kernel takes the slots by counting back down.
template <class T>
__device__ __forceinline__ int32_t bucket_entry(const SceneDev<T>& s, const Vec4<T>& x,
                                                int32_t me, int32_t p, bool blocks,
                                                int32_t* __restrict__ cell_of,
                                                uint32_t* __restrict__ count,
                                                int32_t* __restrict__ cell_ids,
                                                uint32_t stride, int2* __restrict__ ovf,
                                                uint32_t* __restrict__ ovf_ctr) {
  const int32_t c = grid_cell(s, x);
  // the two branches start differently so that the compiler keeps the dense form's atomic without
  // a return value
  if (blocks) {
    const uint32_t slot = atomicAdd(count + c, 1u);
    cell_of[p] = c;
    file_id(me, c, slot, cell_ids, stride, ovf, ovf_ctr);
  } else {
    cell_of[p] = c;
    atomicAdd(count + c, 1u);
  }
  return c;
}

// Incremental bucketing (NextBucket::arr): particle me leaves cell c_old (c_old < 0: none, the
// first pass) for cell c of the bucketing of generation `tag`. Both cells' arrival words are raised
// to the generation (which marks them changed for the rank kernel and restarts a word left by an
// earlier generation at zero arrivals), the counts move, and the particle takes the arrival slot
// that its addition to c's word returns. The raise and the addition are atomics on one word, so
// every addition of a generation lands on a word that already carries it. Both cells' marks take
// the generation's low byte (plain stores: every writer of a generation stores the same value). A
// particle that changes scan tile also moves one from its old tile's kept total to the new one's
// (TileTotals); after the first pass (c_old < 0) the totals are summed from the counts instead.
__device__ __forceinline__ void bucket_move(int32_t me, int32_t c, int32_t c_old, uint32_t tag,
                                            uint32_t* __restrict__ count,
                                            unsigned long long* __restrict__ arr,
                                            uint8_t* __restrict__ mark, const TileTotals& tt,
                                            int32_t* __restrict__ cell_ids, uint32_t stride,
                                            int2* __restrict__ ovf,
                                            uint32_t* __restrict__ ovf_ctr) {
  const unsigned long long t = (unsigned long long)tag << 32;
  if (c_old >= 0) {
    atomicSub(count + c_old, 1u);
    atomicMax(arr + c_old, t);
    mark[c_old] = (uint8_t)tag;
    const uint32_t to = (uint32_t)c_old >> tt.shift, tn = (uint32_t)c >> tt.shift;
    if (to != tn) {
      uint32_t* __restrict__ tot = tt.tot + (size_t)(*tt.gen & 1u) * (size_t)tt.ntiles;
      atomicSub(tot + to, 1u);
      atomicAdd(tot + tn, 1u);
    }
  }
  atomicAdd(count + c, 1u);
  atomicMax(arr + c, t);
  mark[c] = (uint8_t)tag;
  const uint32_t slot = (uint32_t)atomicAdd(arr + c, 1ull);
  file_id(me, c, slot, cell_ids, stride, ovf, ovf_ctr);
}

// Adds the calling lanes' particles (cells c) to their tiles' counts for the direct scan
// (TileCounts): one atomic for the wave when its particles share a tile, else one per particle.
__device__ __forceinline__ void tile_count_wave(const TileCounts& tiles, int32_t c) {
  if (!tiles.tc) return;
  uint32_t* __restrict__ tc =
      tiles.tc + (size_t)(*tiles.gen & 1u) * (size_t)tiles.ntiles * kTileStride;
  const int32_t t = c >> kDirectShift;
  const uint64_t act = __ballot(1);
  const int leader = __ffsll((unsigned long long)act) - 1;
  const int32_t t0 = __shfl(t, leader, 64);
  if (__ballot(t == t0) == act) {
    if (lane_id() == leader) atomicAdd(tc + (size_t)t0 * kTileStride, (uint32_t)__popcll(act));
  } else {
    atomicAdd(tc + (size_t)t * kTileStride, 1u);
  }
}

}  // namespace psx
