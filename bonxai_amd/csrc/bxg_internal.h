// bxg_internal.h -- device-side layout of a sparse voxel map on one MI355X
// and the structures shared by the kernels and the host runtime.
//
// One map (the counterpart of Bonxai::VoxelGrid<T>, bonxai.hpp:114-333):
//
//   root directory   open-addressing table of root keys, root_cap slots.
//                    A root key is (coord >> log2n) per axis, 21 bits each
//                    (biased by 2^20), packed into one 64-bit word so a slot
//                    is claimed with a single 64-bit atomicCAS.  The slot
//                    index IS the inner node: children[slot * inner_size + i]
//                    holds the leaf index of inner cell i, or EMPTY.
//   leaf pool        leaf_cap leaves handed out by an atomic bump counter:
//                    leaf_masks[leaf * leaf_words + w]  (bit n of a leaf is
//                    word n >> 6, bit n & 63, as mask.hpp:329-365)
//                    cells[(leaf * leaf_size + n) * value_bytes]
//                    leaf_origin[leaf] = min-corner coordinate of the leaf.
//                    touched[leaf * leaf_size + n] (probabilistic maps): one
//                    byte a cell, nonzero only while a scan runs, for the
//                    cells that scan updates in the leaves it creates.
//
// Invariant kept by every kernel: a cell whose mask bit is off holds all-zero
// bytes.  The reference never reads an off cell's storage (value() returns
// null, setValue/setCellOn/value(create) overwrite it, forEachCell and
// Serialize skip it), so the invariant is unobservable, and it lets the
// scan-integration kernels derive "cell was off" from "cell word is zero".
// Pool entries at or past leaf_count hold undefined bytes (a cleared or grown
// pool is not zeroed): whatever creates a leaf writes all its cells and mask
// words, as the reference's Grid constructor initialises a new leaf
// (bonxai.hpp:51-58): the scan's leaf assembly (k_scan_new_leaves), the bulk
// calls' k_zero_new_leaves, or k_write_leaves.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace bxg {

constexpr uint64_t kEmptyKey = ~0ull;
constexpr uint32_t kEmptyChild = 0xFFFFFFFFu;
constexpr uint32_t kPendingChild = 0xFFFFFFFEu;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int32_t kRootBias = 1 << 20;  // root key field bias (21-bit fields)

// device status bits (sticky until the host clears them)
enum : uint32_t {
  kStOutOfPool = 1u,
  kStRootFull = 2u,
  kStCoordRange = 4u,
  kStScratchFull = 8u,
};

// map flags
enum : uint32_t {
  kMapProbabilistic = 1u,
};

// Three 128-byte L2 lines: the bump counters, which the creating kernels
// update with atomics; the words every kernel reads while a call runs
// (status, cur, leaf_base: written only between kernels, or by a rare error
// flag); the statistics, which some kernels add to.  A load of a line that
// other lanes keep updating with atomics is served slowly, so the words the
// kernels read stay off the atomically updated lines.
struct alignas(128) MapCounters {
  uint32_t leaf_count;  // bump allocator (may run past leaf_cap on failure)
  uint32_t root_count;
  uint32_t pad0[30];
  uint32_t status;
  int32_t cur;  // ProbabilisticMap update_count, 1..3
  // leaf_count when the running scan or bulk call started: leaves at or
  // above it were created by that call, and it initialises them
  uint32_t leaf_base;
  uint32_t pad1;
  uint64_t scans_applied;
  uint64_t pad2[13];
  uint64_t stat_endpoint_updates;
  uint64_t stat_rays;
  uint64_t stat_ray_updates;
  uint64_t stat_ray_visits;  // counted runs: cells the scan's rays crossed
  uint64_t pad3[12];
};
static_assert(sizeof(MapCounters) == 384, "MapCounters is three cache lines");
static_assert(offsetof(MapCounters, status) == 128 && offsetof(MapCounters, stat_endpoint_updates) == 256,
              "MapCounters lines");

struct alignas(64) MapDev {
  double inv_res;
  double res;
  uint32_t inner_bits, leaf_bits, log2n, inner_size;
  uint32_t leaf_size, leaf_words, value_bytes, flags;
  uint64_t* root_keys;
  uint32_t* children;
  uint32_t root_mask;  // root_cap - 1
  uint32_t max_probe;
  uint64_t* leaf_masks;
  uint8_t* cells;
  int4* leaf_origin;
  uint8_t* touched;  // leaf_cap * leaf_size bytes, probabilistic maps only
  uint32_t leaf_cap;
  int32_t opt_miss, opt_hit, opt_cmin, opt_cmax, opt_thr;
  MapCounters* ctr;
  MapCounters* ctr_host;  // pinned host mirror, written at the end of every scan
};

// one scan of a (possibly batched) insertPointCloud call
struct alignas(16) ScanDesc {
  const MapDev* map;
  const void* points;  // xyz triples, float32 or float64 (see flags)
  uint64_t pt_off;     // first flattened point index of this scan
  uint32_t n;
  uint32_t flags;  // kScan* bits
  double fx, fy, fz;
  double max_range, r2;
  int32_t ox, oy, oz;  // origin voxel, posToCoord(origin)
  uint32_t tag;        // scan index inside the batch (endpoint table key)
  const int4* extra;   // endpoints added by addHit/MissPoint whose rays are due
  uint32_t nextra;
  uint32_t stride;     // bytes from one point to the next
  float tf[12];        // kScanTransform: rows 0..2 of the sensor-to-world matrix
  uint32_t reach;      // max_range in voxels, at most 2^24 (0: no range, as for
                       // addHitPoint / addMissPoint): ray length classes
};

enum : uint32_t {
  kScanF64 = 1u,        // points are float64
  kScanForceHit = 2u,   // addHitPoint: no max-range clip, kind hit
  kScanForceMiss = 4u,  // addMissPoint: no max-range clip, kind miss
  kScanNoRays = 8u,     // endpoint update only, keep update_count
  kScanTransform = 16u, // float points: drop non-finite, then x' = tf * (x, y, z, 1)
};

// size of a batch, read by the kernels from device memory so that a captured
// graph can replay any batch up to the capacity it was built for
struct alignas(16) BatchHdr {
  uint32_t nscans;
  uint32_t tag_bits;  // bits of the scan index in a winner's bucket (0..6)
  uint64_t total;
};

// Winners are bucketed by (scan index mod 64, direction of the ray) so that
// the rays one wavefront walks together touch the same cache lines; with
// ScanScratch::ray_order the direction bins are coarser and classes of ray
// length (equal parts of the scan's max range) sit between a direction
// sector and the direction inside it, so that a wavefront's rays also end at
// about one distance (k_scan_winners, ray_order).
constexpr uint32_t kDirBits = 12;  // direction and length-class bits of a bucket
constexpr uint32_t kMaxTagBits = 6;
constexpr uint32_t kMaxBuckets = 1u << (kDirBits + kMaxTagBits);
constexpr uint32_t kRegions = 8;   // winner-list regions, one per XCD
constexpr uint32_t kScanTile = 4096;  // buckets one block of the bucket scan covers
constexpr uint32_t kTiles = kMaxBuckets / kScanTile;
static_assert((1u << kDirBits) % kScanTile == 0 && kTiles <= 64, "bucket scan tiles");
// scan counter words: [0] winner count, [1, 1 + 6 * kRegions) region cursors
// and exhausted flags of the three walks (leaf pass, update pass, fresh
// pass), [kFreshWords, + 4) the fresh-pass split (below), [kTileWords, +
// kTiles) bucket-tile sums, [kClipWords, + 64) per scan (index mod 64) the
// points clipped at the max range among those of one wavefront in 16
constexpr uint32_t kTileWords = 64;
constexpr uint32_t kClipWords = kTileWords + kTiles;
constexpr uint32_t kCounterWords = kClipWords + 64;
// With at most 2^kMaxTagBits scans (one bucket tag each), the scans whose map
// is empty at the start take the highest tags, so their winners end the
// list: [kFreshWords, +1] the 64-bit mask of those scans, [+2] the first of
// their tags (kNone: no split, every pass walks the whole list and skips the
// other pipeline's rays), [+3] the index of their first winner.
constexpr uint32_t kFreshWords = 1 + 6 * kRegions;
static_assert(kFreshWords + 4 <= kTileWords, "counter words");

// scratch of one integration call, owned by the context
// One slot of the per-call endpoint table: key (scan tag, root slot, inner
// and leaf index of a voxel) and the lowest index of the points hitting it,
// side by side so a point's claim touches one cache line.  An empty slot is
// all ones (kEmptyKey, kNone).
struct alignas(16) EpSlot {
  uint64_t key;
  uint32_t first;
  uint32_t pad;
};
static_assert(kEmptyKey == ~0ull && kNone == 0xFFFFFFFFu, "an empty EpSlot is all ones");

struct ScanScratch {
  int4* rec;            // per flattened point: endpoint xyz, flags | scan << 16
  uint32_t* tpos;       // per point: endpoint-table slot of its endpoint
  EpSlot* ep;           // endpoint table
  uint32_t ep_mask;     // capacity - 1
  uint32_t chunk_max;   // most rays a wavefront takes per cursor atomic (64..1024)
  int32_t win_cells;    // lockstep window of the update pass, in cells (0: the default)
  int32_t win_leaves;   // lockstep window of the leaf pass, in leaves (0: the default)
  uint32_t fresh_pass;  // 1: the rays of maps empty at the scan's start take the fresh pass
  int32_t win_fresh;    // lockstep window of the fresh pass, in cells (0: the update pass's)
  uint32_t ray_order;   // winner order (0..5, k_scan_winners): 0 direction only
  uint32_t* winners;    // flattened point indices of per-voxel winners, bucketed
  uint32_t* wkey;       // per point: bucket of a winner, kNone otherwise
  uint32_t* hist;       // [kMaxBuckets] winners per bucket (zero between calls)
  uint32_t* offs;       // [kMaxBuckets] bucket write cursors
  uint32_t* counters;   // [0] winner count, [1..16] leaf-pass region cursors
                        // and exhausted flags, [17..32] the same for the ray
                        // pass, [33..48] for the fresh pass
};

// rec.w flag bits
enum : uint32_t {
  kRecHit = 1u,
  kRecMarked = 2u,
  kRecValid = 4u,
  kRecWinner = 8u,
};

// ------------------------------------------------------------ helpers

__host__ __device__ inline uint64_t mix64(uint64_t h) {
  h ^= h >> 33;
  h *= 0xFF51AFD7ED558CCDull;
  h ^= h >> 33;
  h *= 0xC4CEB9FE1A85EC53ull;
  h ^= h >> 33;
  return h;
}

// root key of coordinate (x, y, z); false when a field leaves 21 bits
__host__ __device__ inline bool pack_root(int32_t x, int32_t y, int32_t z, uint32_t log2n,
                                          uint64_t* key) {
  const int32_t rx = x >> log2n, ry = y >> log2n, rz = z >> log2n;
  const uint32_t ux = (uint32_t)(rx + kRootBias), uy = (uint32_t)(ry + kRootBias),
                 uz = (uint32_t)(rz + kRootBias);
  if ((ux | uy | uz) >= (1u << 21)) return false;
  *key = (uint64_t)ux | ((uint64_t)uy << 21) | ((uint64_t)uz << 42);
  return true;
}

__host__ __device__ inline void unpack_root(uint64_t key, uint32_t log2n, int32_t* x, int32_t* y,
                                            int32_t* z) {
  *x = ((int32_t)(key & 0x1FFFFF) - kRootBias) * (1 << log2n);
  *y = ((int32_t)((key >> 21) & 0x1FFFFF) - kRootBias) * (1 << log2n);
  *z = ((int32_t)((key >> 42) & 0x1FFFFF) - kRootBias) * (1 << log2n);
}

// getInnerIndex, bonxai.hpp:431-438
__host__ __device__ inline uint32_t inner_index(int32_t x, int32_t y, int32_t z, uint32_t lb,
                                                uint32_t ib) {
  const uint32_t im = (1u << ib) - 1;
  return ((uint32_t)(x >> lb) & im) | (((uint32_t)(y >> lb) & im) << ib) |
         (((uint32_t)(z >> lb) & im) << (2 * ib));
}

// getLeafIndex, bonxai.hpp:440-447
__host__ __device__ inline uint32_t leaf_index(int32_t x, int32_t y, int32_t z, uint32_t lb) {
  const uint32_t lm = (1u << lb) - 1;
  return ((uint32_t)x & lm) | (((uint32_t)y & lm) << lb) | (((uint32_t)z & lm) << (2 * lb));
}

// the cell hash of the order-free digest (same definition as the oracle's)
__host__ __device__ inline uint64_t cell_hash(int32_t x, int32_t y, int32_t z, uint64_t word) {
  const uint64_t a = (uint64_t)(uint32_t)x | ((uint64_t)(uint32_t)y << 32);
  const uint64_t b = (uint64_t)(uint32_t)z ^ (word * 0x9E3779B97F4A7C15ull);
  return mix64(a ^ mix64(b + 0x632BE59BD9B4E019ull));
}

}  // namespace bxg
