// psx_kernels.hpp -- launchers of the per-step kernels (definitions in psx_kernels.hip).
#pragma once

#include "psx.h"
#include "psx_device.hpp"

namespace psx {

enum : int { kModeIntegrate = 1, kModeContainer = 2, kModeCheck = 4 };
// kFlagCountDrift: the cell counts a scan read did not add up to the particle count (a bucketing
// state that no longer matches the particles; the offsets were clamped to it).
enum : uint32_t { kFlagNonFinite = 1u, kFlagScanStall = 2u, kFlagCountDrift = 4u };

// Events that time the next kernel launch from its dispatch timestamps (consumed by the launch).
void set_launch_events(hipEvent_t start, hipEvent_t stop);

// Look-back scan tiles: 16,384 cells (64 per thread). Measured on MI355X against 4,096-cell tiles:
// the 11.5M-cell hopper grid scans in 24.6 instead of 35.7 us and the 35M-cell batch in 70 instead
// of 98 us (fewer tiles, so shorter look-back chains and more loads in flight per workgroup).
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 64;
constexpr int kScanTile = kScanThreads * kScanPerThread;  // cells per scan tile
constexpr int kScanShift = 14;                            // log2(kScanTile)
static_assert(kScanTile == 1 << kScanShift, "kScanShift must match kScanTile");
// Cells per id-slot plane are rounded to this rather than to the scan tile, so the id slots take
// no more memory than the cells need.
constexpr int64_t kIdPlane = 4096;
constexpr int kCellIds = 8;  // id slots per cell; arrivals beyond go to the cell's overflow chain
constexpr uint32_t kHomeCull = 0x80000000u;  // home word: the home cell may cull its neighbours
// Home word of the incremental bucketing: the container moved the particle out of its bucket
// cell (the bucket cell is then in the cell_cur array at the particle's sorted slot). Grids hold at
// most 2^30 cells, so the cell index fits below both flag bits.
constexpr uint32_t kHomeMoved = 0x40000000u;
constexpr uint32_t kHomeCell = 0x3fffffffu;
// Per-cell flag byte (static per scene): the cell's 27 neighbours may be culled by distance in the
// pair kernel; the cell's triangle list is not empty (only such cells' lists are read).
constexpr uint8_t kCellCull = 1, kCellTris = 2;
// Largest workgroup of the rank/gather kernel: its binary64 form (rank, walls and the mesh contact
// in double) needs more than the 128 registers a 1024-wide workgroup leaves per lane.
template <class T>
constexpr int kMaxBlockSort = sizeof(T) == 4 ? 1024 : 512;

// Spare entries after the triangle-list entry array (the container kernel reads list entries four
// at a time).
constexpr int kTriPad = 4;

// Spare entries after the dense cell lists (the rank kernel reads a cell's first four at once).
constexpr int kListPad = 4;

// Spare entries after every sorted position and velocity array the pair kernel reads (its chunked
// loads may reach three entries past the last particle).
constexpr int kPairPad = 4;

// Spare entries around every cell-offset array the pair kernel reads: the binary32 two-wide walk
// loads the offsets of cells z - 1 .. z + 2 of a column with one 16-byte load, so it reads one entry
// before the first cell's and one past the final total. Arrays are allocated with kOffsetPad spare
// entries and passed to the kernels kOffsetLead entries in.
constexpr int kOffsetLead = 1;
constexpr int kOffsetPad = 3;

// The next step's hash pass (launch_integrate's work on a cell-sorted state), done by the pair kernel
// for each particle right after its new velocity: x + v_new dt, the cell of that position, its
// count and, with id blocks (blocks != 0), its entry in the cell's id block. The state the next
// step starts from is the one the pair kernel reads and writes (sorted snapshot positions, new
// velocities, sorted ids), so the result is what launch_integrate would produce from it.
// count == nullptr: no hash pass.
// Per-tile cell counts for the direct scan (grids of at most kDirectTiles tiles of kDirectTile
// cells): the hash pass adds each particle to its tile's count, one atomic per wave when a wave's
// particles share a tile, so that the scan finds every tile's prefix by summing the counts of the
// tiles before it instead of waiting on them. Two generations of counts alternate (tc + g * ntiles
// for generation parity g = *gen & 1): a scan reads its generation's and clears its own entry of
// the other's, which the next hash pass fills. Tile t's count is word t * kTileStride. *gen advances once per bucketing, in the rank/gather
// kernel. tc == nullptr: the look-back scan (larger grids).
constexpr int kDirectTile = 1024;
constexpr int kDirectShift = 10;
constexpr int kDirectTiles = 1024;
constexpr int kTileStride = 32;  // words between tile counts: one 128-byte line each, so that the
                                 // atomics of different tiles do not queue on one line
struct TileCounts {
  uint32_t* tc;
  uint32_t* gen;
  int32_t ntiles;
};

// Kept per-tile totals of the cell counts (incremental bucketing): two generations of ntiles words
// (tile t of 2^shift cells; generation parity *gen & 1, gen advancing once per bucketing in the
// rank/gather kernel). launch_tile_totals sums the counts of the first pass into the current
// generation's entries; later hash passes move one between two entries for each particle that
// changes tile; the scan sums them for its prefixes and writes the next generation's from the counts.
struct TileTotals {
  uint32_t* tot = nullptr;
  const uint32_t* gen = nullptr;
  int32_t ntiles = 0;
  int32_t shift = 0;  // log2 of the cells per tile: kDirectShift, kDirectShift + 2 or kScanShift
};

// Incremental bucketing (arr != nullptr; look-back-scan grids, id slots): the counts persist from
// one step to the next and only a particle whose next cell differs from its bucket cell touches
// them (one less in the old cell, one more in the new). arr holds one word per cell,
// (generation << 32) | arrivals: a particle that leaves or enters a cell raises the word to the
// generation of the coming bucketing (ovf_ctr[2] + 1; the scan advances ovf_ctr[2] to it), an
// entering one then adds itself to the arrivals and takes the id slot that the sum returns. The rank
// kernel finds a cell that nothing entered or left by its stale generation. Such particles also
// write the generation's low byte into both cells' entries of mark, which the rank kernel reads
// first (one byte per cell instead of eight): only a cell whose mark matches can have changed, and
// a stale match (a generation 256k earlier) is settled by the arrival word. The bucket cell of the
// particle at sorted slot p is its home cell (home[p] & kHomeCell) unless kHomeMoved is set, then
// cell_cur[p].
struct NextBucket {
  const int32_t* id;
  int32_t* cell_of;
  uint32_t* count;
  int32_t* cell_ids;
  uint32_t id_stride;  // cells per id slot (slot-major id blocks)
  int2* ovf;
  uint32_t* ovf_ctr;
  int32_t blocks;
  TileCounts tiles;
  unsigned long long* arr;  // incremental bucketing: per-cell arrival words (nullptr: full)
  uint8_t* mark;            // incremental bucketing: per-cell low byte of the last change's generation
  const uint32_t* home;     // incremental bucketing: home words of the pair kernel's state
  const int32_t* cell_cur;  // incremental bucketing: bucket cells where kHomeMoved is set
  TileTotals tt;            // incremental bucketing: kept tile totals
};

// Per-particle kernels: with `one` (a single scene, its host copy) the workgroups need no table
// (ntab = workgroups); otherwise btab holds one (scene, slice) entry per workgroup.
//
// Hash and count: the cell of each particle's position after the integration step
// (kModeIntegrate) or of its current position, counted in its cell (counts must be 0 on entry).
// blocks: the id is filed into the cell's id slots at its arrival slot (cell_ids: kCellIds slots
// per cell, slot-major, slot k of cell c at k * id_stride + c, id_stride = cells rounded up to
// kIdPlane, then one plane of per-cell overflow chain heads; ovf: n entries (id, previous entry of
// the cell) for arrivals past kCellIds; ovf_ctr[0] counts them and must be 0 on entry, k_scan
// resets it). Otherwise launch_scatter_ids builds dense cell lists once the offsets are known.
// arr != nullptr: the first pass of the incremental bucketing (NextBucket): every particle enters
// its cell as an arrival (counts zero on entry, blocks set).
template <class T>
hipError_t launch_integrate(hipStream_t st, int ntab, int block, const BlockInfo* btab,
                            const SceneDev<T>* scenes, const SceneDev<T>* one, bool blocks,
                            const Vec4<T>* pos, const Vec4<T>* vel, const int32_t* id,
                            int32_t* cell_of, uint32_t* count, int32_t* cell_ids,
                            uint32_t id_stride, int2* ovf, uint32_t* ovf_ctr,
                            const TileCounts& tiles, int mode, uint64_t* arr = nullptr,
                            uint8_t* mark = nullptr);

// The first kernel of a step with PSX_FLAG_SEMANTICS_CUDA: integration, the container (box walls or
// the CUDA backend's mesh contact), the state written back to pos/vel, then the hash pass of the
// post-container positions (cell, count, id slot; tile counts for the direct scan). cflags: the
// per-cell flag byte (kCellTris).
template <class T>
hipError_t launch_front(hipStream_t st, int ntab, int block, const BlockInfo* btab,
                        const SceneDev<T>* scenes, const SceneDev<T>* one, bool blocks,
                        Vec4<T>* pos, Vec4<T>* vel, const int32_t* id, const TriEntry<T>* tent,
                        const int32_t* tcell_off, const uint8_t* cflags, int32_t* cell_of,
                        uint32_t* count, int32_t* cell_ids, uint32_t id_stride, int2* ovf,
                        uint32_t* ovf_ctr, const TileCounts& tiles);

// p[0, n) = 0 (a kernel, so that it may sit in a captured stream).
hipError_t launch_zero_u32(hipStream_t st, uint32_t* p, int64_t n);
// p[0] += v on the device (one thread), in stream order.
hipError_t launch_add_u32(hipStream_t st, uint32_t* p, uint32_t v);

// list[offset[c] + slot] = id[p] with c = cell_of[p] and slot = --count[c]: unordered dense cell
// lists; the counts end at zero.
hipError_t launch_scatter_ids(hipStream_t st, int block, int32_t n, const int32_t* cell_of,
                              uint32_t* count, const uint32_t* offset, const int32_t* id,
                              int32_t* list);

// One-pass exclusive scan of count[0, ntiles * kScanTile) into offset[0 .. ntiles * kScanTile]
// (the last entry is the total); clears the counts when `clear`, moves ovf_ctr[0] to ovf_ctr[1],
// zeroes ovf_ctr[0] and advances the generation ovf_ctr[2] (ovf_ctr holds three words). Every
// offset is clamped to n (the particle count the counts must add up to; kFlagCountDrift when they
// do not), so that no index derived from the offsets can leave the particle arrays. tile_state (ntiles words) persists between launches and is zeroed once at
// allocation; launches on one handle must be serialized.
hipError_t launch_scan(hipStream_t st, uint32_t* count, uint32_t* offset, uint64_t* tile_state,
                       int32_t ntiles, uint32_t* ovf_ctr, uint32_t* flags, bool clear, uint32_t n);

// The kept tile totals of the incremental bucketing's first pass: the counts of each of tt.ntiles
// tiles of 2^tt.shift cells summed into the current generation's entry (TileTotals).
hipError_t launch_tile_totals(hipStream_t st, const uint32_t* count, const TileTotals& tt);

// The incremental bucketing's scan over ntiles tiles of 2^tt.shift cells (counts left in place):
// each tile's prefix is summed from the kept tile totals of the current generation (TileTotals), and
// each tile writes the total of its counts for the next; kFlagCountDrift where the two differ.
hipError_t launch_scan_tiles(hipStream_t st, const uint32_t* count, uint32_t* offset,
                             const TileTotals& tt, int32_t ntiles, uint32_t* ovf_ctr, uint32_t* flags,
                             uint32_t n);

// The same scan over tiles.ntiles tiles of kDirectTile cells, each tile's prefix summed from the
// per-tile counts of the hash pass (TileCounts); no tile waits for another.
hipError_t launch_scan_direct(hipStream_t st, uint32_t* count, uint32_t* offset,
                              const TileCounts& tiles, uint32_t* ovf_ctr, bool clear);

// The rank of the incremental bucketing (arr != nullptr; see NextBucket): off_old holds the offsets
// of the state being ranked (its sorted slot p lies in off_old[c] .. off_old[c + 1] of its bucket
// cell c). A particle whose cell nothing entered or left keeps its place in that range; in any other
// cell it counts the cell's remaining members with a smaller id (old members whose new cell is still
// the cell) and its arrivals with a smaller id. cell_cur_o receives the bucket cell of every sorted
// slot whose home cell differs from it (kHomeMoved). rank_o receives every sorted slot's rank (255
// for 255 and more) and rank_old holds those of the state being ranked, where the place of a
// sphere of an unchanged cell is read (one byte per sphere instead of the cell's old offset).
struct IncRank {
  const uint64_t* arr = nullptr;
  const uint8_t* mark = nullptr;
  const uint32_t* off_old = nullptr;
  int32_t* cell_cur_o = nullptr;
  const uint8_t* rank_old = nullptr;
  uint8_t* rank_o = nullptr;
};

// Stable rank among cell mates (from the id blocks, or from the dense lists when !blocks); the
// integration step (kModeIntegrate), container (kModeContainer:
// walls or mesh) and non-finite check (kModeCheck) on the state (pos, vel); gather into
// cell-sorted order, with the bucket cell (cell_o, optional) and the home word of the gathered
// position (home_o, for the pair kernel: home cell | kHomeCull when cull[home cell] is set) of
// every sorted slot. cull: one byte per cell, nonzero where the pair kernel may cull the
// neighbour cells by distance. xpre_o (optional): the integrated, pre-container position of every
// sorted slot (recorded for the adjoint). gen: the bucketing generation, advanced by one.
template <class T>
hipError_t launch_rank_gather(hipStream_t st, int ntab, int block, const BlockInfo* btab,
                              const SceneDev<T>* scenes, const SceneDev<T>* one, bool blocks,
                              int32_t n, const int32_t* cell_of, const uint32_t* offset,
                              const int32_t* cell_ids, uint32_t id_stride, const int2* ovf,
                              const uint32_t* ovf_ctr,
                              const int32_t* list, const Vec4<T>* pos, const Vec4<T>* vel,
                              const int32_t* id, const TriEntry<T>* tent,
                              const int32_t* tcell_off, uint32_t* flags, int mode,
                              const uint8_t* cull, Vec4<T>* pos_o, Vec4<T>* vel_o, int32_t* id_o,
                              int32_t* cell_o, uint32_t* home_o, Vec4<T>* xpre_o, uint32_t* gen,
                              const IncRank& inc = IncRank{});

// The pair kernel form that stages the workgroup's neighbour particles in LDS (pair_variant value;
// one lane per particle): at most kStageFactor sorted positions per workgroup thread.
constexpr int kPairStaged = 32;
// The one-lane pair kernel form that stages each wave's neighbour columns in LDS (pair_variant
// value): two windows of kWaveWin sorted positions per wave.
constexpr int kPairWave = 64;
constexpr int kStageFactor = 12;

// lanes: lanes per particle (1, 2, 4, 8 or 16), or kPairStaged; a workgroup holds block / lanes
// particles (block for kPairStaged). chunk: candidates per round of the one-lane kernel (4 or 2).
// home: the home word of every sorted slot (from the rank/gather pass that produced pos).
template <class T>
hipError_t launch_pairs(hipStream_t st, int ntab, int block, int lanes, int chunk,
                        const BlockInfo* btab, const SceneDev<T>* scenes, const SceneDev<T>* one,
                        const Vec4<T>* pos, const Vec4<T>* vsn, const uint32_t* offset,
                        const uint32_t* home, Vec4<T>* vout, const NextBucket& next);

// Per-workgroup partial statistics (combined on the host in workgroup order).
struct StatPart {
  double ke;
  double mn[3], mx[3];
  double max_overlap;
  uint32_t n_collided;
  uint32_t count;  // particles of the workgroup (0 for an empty one)
};

// Statistics of the sorted state (pos, vel, id) bucketed by its current positions (offset, home
// from that bucketing); ntab workgroups of 256 particles; part[ntab].
template <class T>
hipError_t launch_stats(hipStream_t st, int ntab, const BlockInfo* btab, const SceneDev<T>* scenes,
                        const SceneDev<T>* one, const Vec4<T>* pos, const Vec4<T>* vel,
                        const int32_t* id, const uint32_t* offset, const uint32_t* home,
                        StatPart* part);

// Per-scene statistics out[0 .. nscenes) from the partials part[0 .. ntab) of a launch_stats over
// the table btab (one entry per workgroup of 256 particles, scenes in ascending order), combined on
// the device in the order and with the operations of the host's combination.
hipError_t launch_stats_combine(hipStream_t st, int32_t nscenes, int32_t ntab,
                                const BlockInfo* btab, const StatPart* part, psx_stats* out);

template <class T>
hipError_t launch_unsort(hipStream_t st, int32_t n, const Vec4<T>* pos, const Vec4<T>* vel,
                         const int32_t* id, T* pos3, T* vel3);

template <class T>
hipError_t launch_sort_in(hipStream_t st, int32_t n, Vec4<T>* pos, Vec4<T>* vel, const int32_t* id,
                          const T* pos3, const T* vel3);

}  // namespace psx
