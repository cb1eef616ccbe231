// bxg_kernels.hip -- gfx950 kernels of the sparse voxel map.
//
// Built with -ffp-contract=off: the scan front end (insertPointCloud's
// max-range clip and posToCoord) must round every FP64 operation exactly as
// the reference's x86-64 double code, which has no fused multiply-add.
//
// Memory-model notes.  Cross-lane races are resolved with agent-scope atomics
// (root-slot claims, leaf claims, endpoint first-index claims, mask bits).
// Reads that only decide whether an atomic is needed are plain loads: a root
// key or a leaf index never changes once published inside a kernel, so a
// stale read can only show the EMPTY word, which sends the lane to the CAS.
// The ray phase writes cell words with plain stores: during that phase every
// update of a cell is the same function of the same prior word (the miss
// update guarded by update_id), so concurrent writers store identical
// values, and a stale read can only return the prior word or the updated
// word, either of which leads to the same final value.
//
// Pointers read out of a map header are cast to the global address space
// (BXG_G) so loads compile to global_load rather than flat_load.
#include "bxg_internal.h"
#include "bxg_walk.h"

#define BXG_G __attribute__((address_space(1)))

namespace bxg {

template <class T>
__device__ __forceinline__ T BXG_G* gp(T* p) {
  return (T BXG_G*)p;
}

// --------------------------------------------------------------- atomics

__device__ __forceinline__ uint32_t ld32(const uint32_t* p) {
  return __hip_atomic_load(gp(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st32(uint32_t* p, uint32_t v) {
  __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t cas32(uint32_t* p, uint32_t expect, uint32_t desired) {
  __hip_atomic_compare_exchange_strong(gp(p), &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expect;
}
__device__ __forceinline__ uint64_t cas64(uint64_t* p, uint64_t expect, uint64_t desired) {
  __hip_atomic_compare_exchange_strong(gp(p), &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expect;
}
__device__ __forceinline__ uint32_t add32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t add64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_add(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t or32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_or(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t or64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_or(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t and64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_and(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t xor64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_xor(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t min32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_min(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t max32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_max(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

// Minimum over the 64 lanes of a wavefront, every lane active: DPP moves
// within quads (quad_perm [1,0,3,2], [2,3,0,1]) and rows (row_ror 4, 8), then
// row_bcast 15 into rows 1 and 3 and row_bcast 31 into rows 2 and 3, so that
// lane 63 holds the minimum; no LDS round trips.
__device__ __forceinline__ int32_t wave_min(int32_t v) {
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x124, 0xF, 0xF, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x128, 0xF, 0xF, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));
  v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));
  return __builtin_amdgcn_readlane(v, 63);
}

// Append from the active lanes of a wavefront with one atomic per wavefront;
// returns this lane's slot (valid only where pred is true).
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  const uint64_t b = __ballot(pred);
  if (b == 0) return 0;
  const int leader = __ffsll((unsigned long long)b) - 1;
  uint32_t base = 0;
  if ((int)lane_id() == leader) base = add32(counter, (uint32_t)__popcll(b));
  base = __shfl(base, leader);
  return base + (uint32_t)__popcll(b & lanes_below());
}

// 64-bit form of wave_append, for counts that may pass 2^32.
__device__ __forceinline__ uint64_t wave_append64(uint64_t* counter, bool pred) {
  const uint64_t b = __ballot(pred);
  if (b == 0) return 0;
  const int leader = __ffsll((unsigned long long)b) - 1;
  uint64_t base = 0;
  if ((int)lane_id() == leader) base = add64(counter, (uint64_t)__popcll(b));
  base = __shfl(base, leader);
  return base + (uint64_t)__popcll(b & lanes_below());
}

// Bump allocation from per-map counters, aggregated over a wavefront: the
// active lanes with pred set are grouped by counter address (lanes of one
// wavefront usually serve one map), each group takes one atomic add of its
// size, and each lane gets base + its rank in the group.  Every active lane
// must call it (it is a wavefront collective); the result is meaningful only
// where pred is set.
__device__ __forceinline__ uint32_t wave_bump(uint32_t* counter, bool pred) {
  uint64_t pending = __ballot(pred);
  uint32_t mine = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const unsigned long long lc = __shfl((unsigned long long)counter, leader);
    const uint64_t same = __ballot(pred && (unsigned long long)counter == lc);
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = add32((uint32_t*)lc, (uint32_t)__popcll(same));
    base = __shfl(base, leader);
    if ((same >> lane_id()) & 1ull) mine = base + (uint32_t)__popcll(same & lanes_below());
    pending &= ~same;
  }
  return mine;
}

// The lowest lane of this lane's run: the maximal block of consecutive lanes
// that all have `on` set and equal (k1, k2).  Lanes of a wavefront that share
// a table entry and sit side by side let that lane do the entry's atomics
// for all of them.  Every lane must call it; the result means something only
// where `on` is set.
__device__ __forceinline__ int run_leader(bool on, uint64_t k1, uint32_t k2) {
  const uint32_t lane = lane_id();
  // the previous lane's keys: DPP wave_shr:1 (lane i reads lane i - 1; lane
  // 0's value is not used), no LDS round trip
  const uint32_t k1lo = (uint32_t)k1, k1hi = (uint32_t)(k1 >> 32);
  const uint32_t p1lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k1lo, 0x138, 0xF, 0xF, false);
  const uint32_t p1hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k1hi, 0x138, 0xF, 0xF, false);
  const uint64_t p1 = ((uint64_t)p1hi << 32) | p1lo;
  const uint32_t p2 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)k2, 0x138, 0xF, 0xF, false);
  const uint64_t onmask = __ballot(on);
  const bool prev_on = lane > 0 && ((onmask >> (lane - 1)) & 1ull);
  const bool start = on && !(prev_on && p1 == k1 && p2 == k2);
  const uint64_t upto = __ballot(start) & (~0ull >> (63u - lane));
  return upto ? 63 - __clzll(upto) : (int)lane;
}

// ------------------------------------------------------------ map view

// A register copy of a map header with global-address-space pointers.
struct MapView {
  uint64_t BXG_G* root_keys;
  uint32_t BXG_G* children;
  uint64_t BXG_G* masks;
  uint8_t BXG_G* cells;
  uint8_t BXG_G* touched;
  int4* origin;  // HIP vector types keep the generic address space
  MapCounters* ctr;
  double inv_res;
  uint32_t ib, lb, log2n, isz, lsz, lw, vb, flags;
  uint32_t root_mask, max_probe, leaf_cap;
  int32_t miss, hit, cmin, cmax, thr;
};

__device__ __forceinline__ MapView load_view(const MapDev* p) {
  const MapDev BXG_G* q = (const MapDev BXG_G*)p;
  MapView v;
  v.root_keys = gp(q->root_keys);
  v.children = gp(q->children);
  v.masks = gp(q->leaf_masks);
  v.cells = gp(q->cells);
  v.touched = gp(q->touched);
  v.origin = q->leaf_origin;
  v.ctr = q->ctr;
  v.inv_res = q->inv_res;
  v.ib = q->inner_bits;
  v.lb = q->leaf_bits;
  v.log2n = q->log2n;
  v.isz = q->inner_size;
  v.lsz = q->leaf_size;
  v.lw = q->leaf_words;
  v.vb = q->value_bytes;
  v.flags = q->flags;
  v.root_mask = q->root_mask;
  v.max_probe = q->max_probe;
  v.leaf_cap = q->leaf_cap;
  v.miss = q->opt_miss;
  v.hit = q->opt_hit;
  v.cmin = q->opt_cmin;
  v.cmax = q->opt_cmax;
  v.thr = q->opt_thr;
  return v;
}

// counter words read at kernel start (written by earlier kernels)
__device__ __forceinline__ uint32_t map_status(const MapView& m) { return *gp(&m.ctr->status); }
__device__ __forceinline__ int32_t map_cur(const MapView& m) { return *gp(&m.ctr->cur); }

constexpr uint32_t kFatal = kStOutOfPool | kStRootFull | kStCoordRange | kStScratchFull;

// ---------------------------------------------------------- directory

__device__ __forceinline__ uint32_t root_hash(uint64_t key, uint32_t mask) {
  return (uint32_t)mix64(key) & mask;
}

// Find-or-insert of a root key (lock-free CAS insert, linear probing).
// Returns the slot, or -1 (and sets kStRootFull) after max_probe probes;
// `inserted` tells whether this lane's CAS created it (the caller counts it).
__device__ int32_t root_probe(const MapView& m, uint64_t key, bool& inserted) {
  uint32_t h = root_hash(key, m.root_mask);
  inserted = false;
  for (uint32_t p = 0; p < m.max_probe; ++p) {
    uint64_t k = m.root_keys[h];
    if (k == kEmptyKey) {
      k = cas64((uint64_t*)&m.root_keys[h], kEmptyKey, key);
      if (k == kEmptyKey) {
        inserted = true;
        return (int32_t)h;
      }
    }
    if (k == key) return (int32_t)h;
    h = (h + 1) & m.root_mask;
  }
  or32(&m.ctr->status, kStRootFull);
  return -1;
}

// root_probe with the roots a wavefront inserts counted by one atomic per
// map (wave_bump), so every active lane must call it.
__device__ int32_t root_insert(const MapView& m, uint64_t key) {
  bool inserted;
  const int32_t slot = root_probe(m, key, inserted);
  wave_bump(&m.ctr->root_count, inserted);
  return slot;
}

// Lookup of a root key in a phase where no insert runs concurrently.
__device__ __forceinline__ int32_t root_find(const MapView& m, uint64_t key) {
  uint32_t h = root_hash(key, m.root_mask);
  for (uint32_t p = 0; p <= m.root_mask; ++p) {
    const uint64_t k = m.root_keys[h];
    if (k == key) return (int32_t)h;
    if (k == kEmptyKey) return -1;
    h = (h + 1) & m.root_mask;
  }
  return -1;
}

// Claim of the inner-grid entry *cp for a new leaf: true if this lane moved
// it from EMPTY to PENDING; otherwise `seen` is the entry as found (a leaf
// index, or PENDING while another lane creates the leaf).
__device__ __forceinline__ bool leaf_claim(uint32_t* cp, uint32_t& seen) {
  uint32_t v = *gp(cp);
  if (v == kEmptyChild) v = cas32(cp, kEmptyChild, kPendingChild);
  seen = v;
  return v == kEmptyChild;
}

// leaf_claim for a caller that has just loaded the entry and found it EMPTY:
// straight to the CAS (no second load round trip)
__device__ __forceinline__ bool leaf_claim_empty(uint32_t* cp, uint32_t& seen) {
  seen = cas32(cp, kEmptyChild, kPendingChild);
  return seen == kEmptyChild;
}

// Leaf indices for the lanes that claimed an entry (allocateLeafGrid and
// GridBlockAllocator::allocateBlock, bonxai.hpp:572-586 and
// grid_allocator.hpp:94-123): one wavefront-aggregated bump of the map's
// leaf counter per wavefront and map, then each claimer writes its leaf's
// origin and publishes the index in the entry.  A claimer past the pool's
// capacity puts the entry back to EMPTY and flags the map (the host grows the
// pool and runs the call again).  Every active lane must call it; returns the
// new index, or EMPTY on failure, where `claimed` is set.
__device__ __forceinline__ uint32_t leaf_publish(const MapView& m, bool claimed, uint32_t* cp,
                                                 uint32_t slot, int32_t lx, int32_t ly, int32_t lz) {
  const uint32_t idx = wave_bump(&m.ctr->leaf_count, claimed);
  if (!claimed) return kEmptyChild;
  if (idx >= m.leaf_cap) {
    st32(cp, kEmptyChild);
    or32(&m.ctr->status, kStOutOfPool);
    return kEmptyChild;
  }
  m.origin[idx] = make_int4(lx, ly, lz, (int32_t)slot);
  st32(cp, idx);
  return idx;
}

// Make sure leaf `ci` of root `slot` exists (a wavefront collective, through
// leaf_publish).  Returns the child word as seen: a leaf index, PENDING while
// another lane creates it, or EMPTY when the pool ran out.
__device__ __forceinline__ uint32_t leaf_request(const MapView& m, uint32_t slot, uint32_t ci,
                                                 int32_t lx, int32_t ly, int32_t lz) {
  uint32_t* cp = (uint32_t*)&m.children[(uint64_t)slot * m.isz + ci];
  uint32_t seen;
  const bool won = leaf_claim(cp, seen);
  const uint32_t idx = leaf_publish(m, won, cp, slot, lx, ly, lz);
  return won ? idx : seen;
}

// leaf index holding coordinate c, or kNone (read-only phases)
__device__ __forceinline__ uint32_t find_leaf(const MapView& m, int32_t x, int32_t y, int32_t z) {
  uint64_t key;
  if (!pack_root(x, y, z, m.log2n, &key)) return kNone;
  const int32_t s = root_find(m, key);
  if (s < 0) return kNone;
  const uint32_t c = m.children[(uint64_t)s * m.isz + inner_index(x, y, z, m.lb, m.ib)];
  return c >= kPendingChild ? kNone : c;
}

// The root slot of the last root a walk visited.
struct RootCache {
  int32_t rx, ry, rz, slot;
  bool valid;
  // fields set, so that the cache test below may read them without branches
  __device__ __forceinline__ void reset() {
    rx = ry = rz = 0;
    slot = -1;
    valid = false;
  }
};

// slot of the root holding (x, y, z): looked up (kInsert false) or created
// (kInsert true), cached while the walk stays in the root
template <bool kInsert>
__device__ __forceinline__ int32_t root_slot(const MapView& m, RootCache& rc, int32_t x, int32_t y,
                                             int32_t z) {
  const int32_t rmask = ~((1 << m.log2n) - 1);
  const int32_t rx = x & rmask, ry = y & rmask, rz = z & rmask;
  if (!rc.valid | (((rx ^ rc.rx) | (ry ^ rc.ry) | (rz ^ rc.rz)) != 0)) {
    uint64_t key;
    int32_t s = -1;
    if (pack_root(x, y, z, m.log2n, &key)) s = kInsert ? root_insert(m, key) : root_find(m, key);
    rc.rx = rx;
    rc.ry = ry;
    rc.rz = rz;
    rc.slot = s;
    rc.valid = true;
  }
  return rc.slot;
}

// ------------------------------------------------------- cell update

// signed 4-bit update_id and 28-bit log-odds of a CellT word
__device__ __forceinline__ int32_t cell_id(uint32_t w) { return ((int32_t)(w << 28)) >> 28; }
__device__ __forceinline__ int32_t cell_p(uint32_t w) { return ((int32_t)w) >> 4; }

// probability_log = min(p + hit, clamp_max) or max(p + miss, clamp_min),
// update_id = cur: addHitPoint, addMissPoint and the clearPoint lambda
// (probabilistic_map.cpp:30-41, 43-54, 81-89); the sum wraps like the
// reference's int arithmetic and the store keeps the low 28 bits like its
// 28-bit bit-field.
__device__ __forceinline__ uint32_t cell_hit(uint32_t w, int32_t hit, int32_t cmax, int32_t cur) {
  int32_t p = (int32_t)((uint32_t)cell_p(w) + (uint32_t)hit);
  p = p < cmax ? p : cmax;
  return ((uint32_t)p << 4) | ((uint32_t)cur & 0xFu);
}
__device__ __forceinline__ uint32_t cell_miss(uint32_t w, int32_t miss, int32_t cmin, int32_t cur) {
  int32_t p = (int32_t)((uint32_t)cell_p(w) + (uint32_t)miss);
  p = p > cmin ? p : cmin;
  return ((uint32_t)p << 4) | ((uint32_t)cur & 0xFu);
}

// ------------------------------------------------------------- rays

// The cell walk of a ray (RayWalk, collect_cells) is in bxg_walk.h, shared
// with the host check of the walk.
__device__ __forceinline__ uint32_t collect_run(const MapView& m, RayWalk& w, uint32_t li[kRun]) {
  return collect_cells(w, m.lb, li);
}

// The leaves a ray's cells [k, kend) lie in, one leaf per step, without
// visiting the cells.  Along an axis with |d| > 0 the ray takes its n-th step
// at cell ceil((2n - 1) m / (2|d|)) (the inverse of the closed form), so the
// cell at which it enters the leaf N axis steps away is that value for N;
// the next boundary is L steps further, which adds floor(L m / |d|) and a
// remainder to the quotient, so after the set-up (a few divisions a ray)
// each leaf costs additions and compares only.  Axes crossing at the same
// cell move the ray diagonally, into one leaf.  Same leaves, in the same
// order, as the cell walk.
struct LeafWalk {
  int32_t lx, ly, lz;  // origin of the current leaf
  int32_t stx, sty, stz;  // +-L per leaf along each axis
  int32_t kx, ky, kz;  // cell at which the ray enters the next leaf along each axis
  uint32_t qx, qy, qz, rx, ry, rz;  // kx = qx + (rx > 0): (2N - 1) m = qx * 2|dx| + rx
  uint32_t Qx, Qy, Qz, Rx, Ry, Rz;  // per further leaf: floor(L m / |d|), 2 (L m mod |d|)
  uint32_t bx, by, bz;  // 2|d|
  int32_t k, kend;
  static constexpr int32_t kNever = 0x7fffffff;
  __device__ __forceinline__ static void axis(int32_t c, int32_t s, int32_t d, int32_t n, int32_t m,
                                              uint32_t L, int32_t& lc, int32_t& st, int32_t& kc,
                                              uint32_t& q, uint32_t& r, uint32_t& Q, uint32_t& R,
                                              uint32_t& b) {
    const int32_t lm = (int32_t)L - 1;
    lc = c & ~lm;
    st = s * (int32_t)L;
    b = 2u * (uint32_t)(d > 0 ? d : 1);
    if (d == 0) {
      kc = kNever;
      q = r = Q = R = 0;
      return;
    }
    // axis steps from the origin to the first cell of the next leaf
    const int64_t N = (int64_t)n + (s > 0 ? (int64_t)L - (c & lm) : (int64_t)(c & lm) + 1);
    const uint64_t A = (uint64_t)(2 * N - 1) * (uint64_t)m;
    const uint64_t q64 = A / b;
    const uint64_t LM = (uint64_t)L * (uint64_t)m;
    const uint64_t Q64 = LM / (uint64_t)d;
    // quotients past 2^30 are beyond any ray (kend < 2^30): such an axis
    // never crosses again, and the sums below cannot overflow
    q = (uint32_t)(q64 < (1ull << 30) ? q64 : (1ull << 30));
    r = (uint32_t)(A - q64 * b);
    Q = (uint32_t)(Q64 < (1ull << 30) ? Q64 : (1ull << 30));
    R = 2u * (uint32_t)(LM - Q64 * (uint64_t)d);
    kc = (int32_t)(q + (r > 0 ? 1u : 0u));
  }
  // the leaf of cell w.k, walking cells [w.k, w.kend) of w's ray
  __device__ __forceinline__ void init(const RayWalk& w, uint32_t lb) {
    const uint32_t L = 1u << lb;
    k = w.k;
    kend = w.kend;
    const int64_t m2 = 2 * (int64_t)w.mx;
    // axis steps already taken at cell k (none at k = 0)
    const int32_t nx = k ? (int32_t)((2 * (int64_t)k * w.dx + w.mx) / m2) : 0;
    const int32_t ny = k ? (int32_t)((2 * (int64_t)k * w.dy + w.mx) / m2) : 0;
    const int32_t nz = k ? (int32_t)((2 * (int64_t)k * w.dz + w.mx) / m2) : 0;
    axis(w.cx, w.sx, w.dx, nx, w.mx, L, lx, stx, kx, qx, rx, Qx, Rx, bx);
    axis(w.cy, w.sy, w.dy, ny, w.mx, L, ly, sty, ky, qy, ry, Qy, Ry, by);
    axis(w.cz, w.sz, w.dz, nz, w.mx, L, lz, stz, kz, qz, rz, Qz, Rz, bz);
  }
  __device__ __forceinline__ bool done() const { return k >= kend; }
  __device__ __forceinline__ static void cross(int32_t& lc, int32_t st, int32_t& kc, uint32_t& q,
                                               uint32_t& r, uint32_t Q, uint32_t R, uint32_t b) {
    lc += st;
    q += Q;
    r += R;
    if (r >= b) {
      ++q;
      r -= b;
    }
    kc = q < (1u << 30) ? (int32_t)(q + (r > 0 ? 1u : 0u)) : kNever;
  }
  // move to the next leaf (or past kend)
  __device__ __forceinline__ void next() {
    int32_t kn = kx < ky ? kx : ky;
    kn = kn < kz ? kn : kz;
    if (kn >= kend) {
      k = kend;
      return;
    }
    if (kx == kn) cross(lx, stx, kx, qx, rx, Qx, Rx, bx);
    if (ky == kn) cross(ly, sty, ky, qy, ry, Qy, Ry, by);
    if (kz == kn) cross(lz, stz, kz, qz, rz, Qz, Rz, bz);
    k = kn;
  }
};

// Allocation pass, one step of a ray: the root of its current leaf (created
// if missing) and up to kLeafRun of its next leaves in that root; their
// inner-grid entries are loaded together (one memory round trip, not one a
// leaf), and only those still empty take the creating path.
constexpr int kLeafRun = 2;
__device__ __forceinline__ void leaves_step(const MapView& m, LeafWalk& lw, RootCache& rc) {
  const int32_t slot = root_slot<true>(m, rc, lw.lx, lw.ly, lw.lz);
  const int32_t rmask = ~((1 << m.log2n) - 1);
  int32_t lx[kLeafRun] = {}, ly[kLeafRun] = {}, lz[kLeafRun] = {};
  int n = 0;
#pragma unroll
  for (int j = 0; j < kLeafRun; ++j) {
    if (j == 0 || (n == j && !lw.done() && (lw.lx & rmask) == rc.rx && (lw.ly & rmask) == rc.ry &&
                   (lw.lz & rmask) == rc.rz)) {
      lx[j] = lw.lx;
      ly[j] = lw.ly;
      lz[j] = lw.lz;
      n = j + 1;
      lw.next();
    }
  }
  if (slot < 0) return;  // root directory full: the call fails
  uint32_t v[kLeafRun];
  uint32_t* cp[kLeafRun];
#pragma unroll
  for (int j = 0; j < kLeafRun; ++j) {
    cp[j] = (uint32_t*)&m.children[(uint64_t)slot * m.isz + inner_index(lx[j], ly[j], lz[j], m.lb, m.ib)];
    v[j] = j < n ? *gp(cp[j]) : 0u;
  }
  // claims first, then one aggregated allocation per run slot for the
  // whole wavefront (every active lane takes part, claimed or not)
  bool won[kLeafRun];
#pragma unroll
  for (int j = 0; j < kLeafRun; ++j) {
    // neighbouring rays in neighbouring lanes often reach the same missing
    // leaf at the same step: the lowest lane of each such run claims it
    const bool missing = j < n && v[j] == kEmptyChild;
    const int lead = run_leader(missing, (uint64_t)cp[j], 0u);
    uint32_t seen;
    won[j] = (missing && lead == (int)lane_id()) ? leaf_claim_empty(cp[j], seen) : false;
  }
#pragma unroll
  for (int j = 0; j < kLeafRun; ++j) leaf_publish(m, won[j], cp[j], (uint32_t)slot, lx[j], ly[j], lz[j]);
}

// One leaf of a ray, update pass: the cells of the ray in the current leaf
// (up to kRun) are loaded together, then each takes the miss update unless
// its update_id already equals cur.  Returns the cells written (counted
// form: exact, by CAS).
// leaf index of the walk's current cell (read-only phases)
__device__ __forceinline__ uint32_t walk_leaf(const MapView& m, const RayWalk& w, RootCache& rc) {
  const int32_t slot = root_slot<false>(m, rc, w.cx, w.cy, w.cz);
  if (slot < 0) return kNone;
  const uint32_t c = m.children[(uint64_t)slot * m.isz + inner_index(w.cx, w.cy, w.cz, m.lb, m.ib)];
  return c >= kPendingChild ? kNone : c;
}

// `leaf` is the leaf of the walk's current cell on entry; on return it is the
// leaf of the next run (looked up before this run's cells are loaded, so the
// two memory round trips overlap).
//
// Leaves at or above `lbase` were created by the running scan: each of their
// cells was off (zero) before it, and the ray update of such a cell gives the
// same word whichever ray gets there first, cell_miss(0).  So a run in a new
// leaf only marks its cells in the leaf's touched bytes (plain byte stores of
// 1, every writer agreeing; no read, no atomic), and k_scan_new_leaves writes
// the leaf's cells and mask words from those bytes after the scan's endpoints
// are marked.  In older leaves each cell is read, updated unless its
// update_id is cur, and a cell turned on sets its mask bit with an atomicOr.
template <bool kCount>
__device__ __forceinline__ uint32_t apply_step(const MapView& m, RayWalk& w, RootCache& rc,
                                               int32_t cur, uint32_t lbase, uint32_t& next_leaf) {
  const uint32_t leaf = next_leaf;
  uint32_t li[kRun];
  const uint32_t n = collect_run(m, w, li);
  next_leaf = w.done() ? kNone : walk_leaf(m, w, rc);
  if (leaf == kNone) return 0;
  if (leaf >= lbase) {
    uint8_t BXG_G* tb = m.touched + (uint64_t)leaf * m.lsz;
    uint32_t nupd = 0;
#pragma unroll
    for (int j = 0; j < kRun; ++j) {
      if ((uint32_t)j < n) {
        if (kCount) {
          // exact first marks: an atomic or on the byte's aligned word
          const uint32_t sh = (li[j] & 3u) * 8u;
          nupd += ((or32((uint32_t*)(tb + (li[j] & ~3u)), 1u << sh) >> sh) & 0xFFu) == 0u ? 1u : 0u;
        } else {
          tb[li[j]] = 1;
        }
      }
    }
    return nupd;
  }
  uint32_t BXG_G* base = (uint32_t BXG_G*)m.cells + (uint64_t)leaf * m.lsz;
  uint32_t v[kRun];
#pragma unroll
  for (int j = 0; j < kRun; ++j)
    v[j] = base[li[j]];  // slots j >= n re-read the run's first cell
  uint32_t nupd = 0;
#pragma unroll
  for (int j = 0; j < kRun; ++j) {
    const bool upd = ((uint32_t)j < n) & (cell_id(v[j]) != cur);  // one branch, not two
    if (upd) {
      uint32_t old = v[j];
      if (kCount) {
        for (;;) {
          const uint32_t seen = cas32((uint32_t*)&base[li[j]], old, cell_miss(old, m.miss, m.cmin, cur));
          if (seen == old) {
            ++nupd;
            break;
          }
          old = seen;
          if (cell_id(old) == cur) break;
        }
      } else {
        base[li[j]] = cell_miss(old, m.miss, m.cmin, cur);
      }
      if (old == 0)
        or64((uint64_t*)&m.masks[(uint64_t)leaf * m.lw + (li[j] >> 6)], 1ull << (li[j] & 63));
    }
  }
  return nupd;
}

// The fresh pass: rays of a map that was empty when the scan started, so
// every leaf they cross is new.  One walk creates the leaves (as the leaf
// pass) and marks the ray's cells (as the update pass does in new leaves),
// where the two-pass pipeline walks each ray twice.  Nothing it writes is
// visible in the map's cells until k_scan_new_leaves, and a failed call's
// marks are dropped there, so the pass may stop at an overflow like the leaf
// pass.  Per lane: the root slot and inner-grid entry of the current cell's
// leaf, loaded one step ahead.

// root slot (created if missing) and inner-grid entry of the walk's current
// cell (a plain load: a stale EMPTY only sends the lane to the claim's CAS, a
// stale PENDING to an atomic re-read)
__device__ __forceinline__ void fresh_entry(const MapView& m, const RayWalk& w, RootCache& rc,
                                            int32_t& slot, uint32_t& entry) {
  slot = root_slot<true>(m, rc, w.cx, w.cy, w.cz);
  entry = slot >= 0
              ? *gp(&m.children[(uint64_t)slot * m.isz + inner_index(w.cx, w.cy, w.cz, m.lb, m.ib)])
              : kEmptyChild;
}

// One leaf of a fresh-pass ray: the leaf is created if missing (the lowest
// lane of each run of lanes at one entry claims it, one aggregated bump per
// wavefront), then the run's cells before the endpoint are marked in the
// leaf's touched bytes.  A leaf another wavefront is still creating
// (PENDING) is re-read and the step taken again on the lane's next turn; its
// creator publishes it within its own step.  A full root directory or pool
// (the call fails) stops the ray.  The run that holds the endpoint's cell
// stores its leaf index in *ep_leaf.  Returns the cells marked first (counted
// form, exact, by atomics).
template <bool kCount>
__device__ __forceinline__ uint32_t fresh_step(const MapView& m, RayWalk& w, RootCache& rc,
                                               int32_t& slot, uint32_t& entry, bool& stopped,
                                               uint32_t* ep_leaf) {
  uint32_t* cp = slot >= 0 ? (uint32_t*)&m.children[(uint64_t)slot * m.isz +
                                                    inner_index(w.cx, w.cy, w.cz, m.lb, m.ib)]
                           : nullptr;
  const bool missing = slot >= 0 && entry == kEmptyChild;
  const int lead = run_leader(missing, (uint64_t)cp, 0u);
  bool won = false;
  uint32_t seen = entry;
  if (missing && lead == (int)lane_id()) won = leaf_claim_empty(cp, seen);
  const int32_t lmask = ~((1 << m.lb) - 1);
  const uint32_t idx = leaf_publish(m, won, cp, (uint32_t)slot, w.cx & lmask, w.cy & lmask, w.cz & lmask);
  if (won) seen = idx;
  // the other lanes of a leader's run take its outcome
  const uint32_t led = (uint32_t)__shfl((int)seen, lead);
  if (missing) seen = led;
  if (slot < 0 || seen == kEmptyChild) {
    stopped = true;
    return 0;
  }
  if (seen == kPendingChild) {
    entry = ld32(cp);
    return 0;
  }
  const int32_t k0 = w.k;
  uint32_t li[kRun];
  const uint32_t n = collect_run(m, w, li);
  // cells k0 .. k0 + n - 1 of the walk; the ray's are those before mx
  const int32_t rest = w.mx - k0;
  const uint32_t nray = rest <= 0 ? 0u : (n < (uint32_t)rest ? n : (uint32_t)rest);
  uint8_t BXG_G* tb = m.touched + (uint64_t)seen * m.lsz;
  uint32_t nupd = 0;
#pragma unroll
  for (int j = 0; j < kRun; ++j) {
    if ((uint32_t)j < nray) {
      if (kCount) {
        const uint32_t sh = (li[j] & 3u) * 8u;
        nupd += ((or32((uint32_t*)(tb + (li[j] & ~3u)), 1u << sh) >> sh) & 0xFFu) == 0u ? 1u : 0u;
      } else {
        tb[li[j]] = 1;
      }
    }
  }
  // the run holding the endpoint's cell (mx) leaves its leaf for
  // k_scan_new_endpoints, which then needs no lookup
  if (k0 <= w.mx && w.mx < k0 + (int32_t)n) *ep_leaf = seen;
  if (!w.done()) fresh_entry(m, w, rc, slot, entry);
  return nupd;
}

// --------------------------------------------------------- scan phases

__device__ __forceinline__ uint32_t find_scan(const ScanDesc* __restrict__ sd, uint32_t nscans,
                                              uint64_t gi) {
  uint32_t lo = 0, hi = nscans;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sd[mid].pt_off <= gi)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint64_t ep_key(uint32_t tag, uint32_t slot, uint32_t ci, uint32_t li) {
  return ((uint64_t)tag << 48) | ((uint64_t)slot << 24) | ((uint64_t)ci << 12) | (uint64_t)li;
}

// Phase 1 (per point): max-range clip and posToCoord in FP64
// (probabilistic_map.hpp:144-157, bonxai.hpp:404-410), root and leaf of the
// endpoint created, pre-scan update_id test, and the first-index claim of
// the endpoint voxel among the points whose voxel is to be updated.
__global__ void __launch_bounds__(256) k_scan_classify(const BatchHdr* __restrict__ hdr,
                                                       const ScanDesc* __restrict__ sd,
                                                       ScanScratch sc) {
  const uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = hdr->total;
  if (gi - lane_id() >= total) return;  // the whole wavefront lies past the last point
  // Every lane of a live wavefront stays to the end: the root, leaf and
  // endpoint claims below are wavefront collectives.
  const bool live = gi < total;
  // the scan of the wavefront's first point, searched once with uniform
  // (scalar) loads; lanes past a scan boundary step forward from it
  const uint32_t nscans = hdr->nscans;
  const uint64_t g0 = (uint64_t)blockIdx.x * blockDim.x +
                      ((uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) << 6);
  uint32_t s = find_scan(sd, nscans, g0);
  if (live)
    while (s + 1 < nscans && sd[s + 1].pt_off <= gi) ++s;
  else
    s = 0u;
  const ScanDesc& d = sd[s];
  const MapView m = load_view(d.map);
  int4 rec = make_int4(0, 0, 0, (int32_t)(s << 16));
  bool valid = false;
  bool clipped = false;
  int32_t ex = 0, ey = 0, ez = 0;
  uint32_t flags = kRecValid;
  if (live) {
    const uint64_t i = gi - d.pt_off;
    double tx, ty, tz;
    bool finite = true;
    const uint8_t BXG_G* pb = gp((const uint8_t*)d.points) + i * d.stride;
    if (d.flags & kScanF64) {
      const double BXG_G* p = (const double BXG_G*)pb;
      tx = p[0];
      ty = p[1];
      tz = p[2];
    } else {
      const float BXG_G* p = (const float BXG_G*)pb;
      float x = p[0], y = p[1], z = p[2];
      if (d.flags & kScanTransform) {
        // bonxai_server.cpp:147-153 drops points with a non-finite coordinate
        finite = __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z);
        // pcl::transformPointCloud in float (bonxai_server.cpp:166-170), no contraction
        const float* t = d.tf;
        const float nx = t[0] * x + (t[1] * y + (t[2] * z + t[3]));
        const float ny = t[4] * x + (t[5] * y + (t[6] * z + t[7]));
        const float nz = t[8] * x + (t[9] * y + (t[10] * z + t[11]));
        x = nx;
        y = ny;
        z = nz;
      }
      tx = (double)x;
      ty = (double)y;
      tz = (double)z;
    }
    if (finite) {
      const double vx = tx - d.fx, vy = ty - d.fy, vz = tz - d.fz;
      const double sq = (vx * vx + vy * vy) + vz * vz;
      double qx = tx, qy = ty, qz = tz;
      if (d.flags & (kScanForceHit | kScanForceMiss)) {
        if (d.flags & kScanForceHit) flags |= kRecHit;
      } else if (sq >= d.r2) {
        clipped = true;
        const double sr = __builtin_sqrt(sq);
        qx = d.fx + ((vx / sr) * d.max_range);
        qy = d.fy + ((vy / sr) * d.max_range);
        qz = d.fz + ((vz / sr) * d.max_range);
      } else {
        flags |= kRecHit;
      }
      const double ax = __builtin_floor(qx * m.inv_res);
      const double ay = __builtin_floor(qy * m.inv_res);
      const double az = __builtin_floor(qz * m.inv_res);
      const double lim = (double)(1 << (20 + m.log2n));
      if (ax >= -lim && ax < lim && ay >= -lim && ay < lim && az >= -lim && az < lim) {
        valid = true;
        ex = (int32_t)ax;
        ey = (int32_t)ay;
        ez = (int32_t)az;
      } else {
        or32(&m.ctr->status, kStCoordRange);
      }
    }
  }
  // The endpoint's root (created if missing), then its leaf (looked up),
  // then its first-index claim.  Neighbouring points of a scan usually share
  // a root and often a voxel, and neighbouring points sit in neighbouring
  // lanes: each run of lanes with the same key does the table work once, on
  // its lowest lane (run_leader), and the others take the result from it.
  uint64_t rkey = 0;
  if (valid) pack_root(ex, ey, ez, m.log2n, &rkey);
  const int rl = run_leader(valid, rkey, s);
  int32_t slot = -1;
  bool rnew = false;
  if (valid && rl == (int)lane_id()) slot = root_probe(m, rkey, rnew);
  wave_bump(&m.ctr->root_count, rnew);
  slot = __shfl(slot, rl);
  if (valid && slot < 0) {
    rec = make_int4(ex, ey, ez, (int32_t)(s << 16));
    valid = false;
  }
  uint32_t ci = 0, li = 0;
  if (valid) {
    ci = inner_index(ex, ey, ez, m.lb, m.ib);
    li = leaf_index(ex, ey, ez, m.lb);
  }
  // the endpoint's leaf as it was before the scan (the leaf pass creates the
  // winners' endpoint leaves; nothing creates leaves in this kernel): a
  // missing leaf means a cell with update_id 0, never cur
  const uint32_t leaf = valid ? *gp(&m.children[(uint64_t)slot * m.isz + ci]) : kEmptyChild;
  bool marked = valid;
  if (valid && leaf < kPendingChild) {
    const uint32_t w = ((const uint32_t BXG_G*)m.cells)[(uint64_t)leaf * m.lsz + li];
    marked = cell_id(w) != map_cur(m);
  }
  // first point of each voxel: lanes are in point order, so a run's lowest
  // lane holds its smallest index, the one the endpoint table keeps
  const uint64_t key = marked ? ep_key(d.tag, (uint32_t)slot, ci, li) : 0ull;
  const int el = run_leader(marked, key, 0u);
  uint32_t tp = kNone;
  if (marked && el == (int)lane_id()) {
    uint32_t h = (uint32_t)mix64(key) & sc.ep_mask;
    for (uint32_t p = 0; p <= sc.ep_mask; ++p) {
      uint64_t k = gp(&sc.ep[h])->key;
      if (k == kEmptyKey) {
        k = cas64(&sc.ep[h].key, kEmptyKey, key);
        if (k == kEmptyKey) k = key;
      }
      if (k == key) {
        min32(&sc.ep[h].first, (uint32_t)gi);
        tp = h;
        break;
      }
      h = (h + 1) & sc.ep_mask;
    }
  }
  tp = __shfl(tp, el);
  if (marked) {
    if (tp == kNone) {
      or32(&m.ctr->status, kStScratchFull);
    } else {
      flags |= kRecMarked;
      sc.tpos[gi] = tp;
    }
  }
  if (valid) rec = make_int4(ex, ey, ez, (int32_t)(flags | (s << 16)));
  if (live) sc.rec[gi] = rec;
  // points clipped at the max range, per scan, counted on one wavefront in
  // 16 (one add per scan of the wavefront; the words share two cache lines,
  // and an add from every wavefront measured 0.5 ms a 6.9 M-point frame):
  // k_scan_winners orders the rays by length only in scans that mix such
  // rays with returns
  uint64_t pend = ((gi >> 6) & 15u) == 0 ? __ballot(clipped) : 0ull;
  while (pend) {
    const int leader = __ffsll((unsigned long long)pend) - 1;
    const uint32_t ls = (uint32_t)__shfl((int)s, leader);
    const uint64_t same = __ballot(clipped && s == ls);
    if ((int)lane_id() == leader) add32(&sc.counters[kClipWords + (ls & 63u)], (uint32_t)__popcll(same));
    pend &= ~same;
  }
}

// Lanes of a wavefront that share a key form one group: one atomic add of
// the group size per group; returns this lane's rank-offset (base + rank).
__device__ __forceinline__ uint32_t wave_add_by_key(uint32_t* table, uint32_t key, bool pred) {
  uint64_t pending = __ballot(pred);
  uint32_t pos = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t lk = (uint32_t)__shfl((int)key, leader);
    const uint64_t same = __ballot(pred && key == lk);
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = add32(&table[lk], (uint32_t)__popcll(same));
    base = (uint32_t)__shfl((int)base, leader);
    if ((same >> lane_id()) & 1ull) pos = base + (uint32_t)__popcll(same & lanes_below());
    pending &= ~same;
  }
  return pos;
}

// octahedral direction bin of (dx, dy, dz), 2^b x 2^b bins in Morton order
// (b = 3..6)
__device__ __forceinline__ uint32_t dir_bin(int32_t dx, int32_t dy, int32_t dz, uint32_t b) {
  const float ax = fabsf((float)dx), ay = fabsf((float)dy), az = fabsf((float)dz);
  const float s = ax + ay + az;
  if (s == 0.f) return 0;
  float u = (float)dx / s, v = (float)dy / s;
  if (dz < 0) {
    const float uu = (1.f - fabsf(v)) * (u >= 0.f ? 1.f : -1.f);
    const float vv = (1.f - fabsf(u)) * (v >= 0.f ? 1.f : -1.f);
    u = uu;
    v = vv;
  }
  const float half = (float)(1u << (b - 1u));
  const uint32_t top = (1u << b) - 1u;
  uint32_t iu = (uint32_t)((u + 1.f) * half), iv = (uint32_t)((v + 1.f) * half);
  iu = iu > top ? top : iu;
  iv = iv > top ? top : iv;
  uint32_t key = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) key |= (((iu >> k) & 1u) << (2 * k)) | (((iv >> k) & 1u) << (2 * k + 1));
  return key;
}

// Winner order of ScanScratch::ray_order: 2^b x 2^b direction bins, 2^cb
// length classes, the top hb bits of the direction (Morton) above the class
// (2b + cb = kDirBits)
struct RayOrder {
  uint32_t b, cb, hb;
};
__device__ __forceinline__ RayOrder ray_order(uint32_t k) {
  switch (k) {
    case 1: return {5u, 2u, 0u};
    case 2: return {4u, 4u, 0u};
    case 3: return {4u, 4u, 3u};
    case 4: return {4u, 4u, 4u};
    case 5: return {5u, 2u, 3u};
    default: return {6u, 0u, 0u};
  }
}

// Phase 2 (per point): the lowest-index point of each claimed voxel is its
// winner (probabilistic_map.cpp:30-54 lets only the first point of a voxel
// through the update_id test); winners are counted per bucket.
__global__ void __launch_bounds__(256) k_scan_winners(const BatchHdr* __restrict__ hdr,
                                                      const ScanDesc* __restrict__ sd,
                                                      ScanScratch sc) {
  const uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool win = false;
  uint32_t key = 0;
  if (gi < hdr->total) {
    const int4 r = sc.rec[gi];
    if ((r.w & kRecMarked) && gp(&sc.ep[sc.tpos[gi]])->first == (uint32_t)gi) {
      win = true;
      const uint32_t s = (uint32_t)r.w >> 16;
      const ScanDesc& d = sd[s];
      uint32_t tag = s & ((1u << hdr->tag_bits) - 1u);
      const uint32_t split = sc.counters[kFreshWords + 2];
      if (split != kNone) {
        // the other scans keep their order in tags [0, split), the fresh
        // scans theirs in [split, nscans)
        const uint64_t fresh = (uint64_t)sc.counters[kFreshWords] |
                               ((uint64_t)sc.counters[kFreshWords + 1] << 32);
        const uint64_t below = (1ull << s) - 1ull;
        tag = ((fresh >> s) & 1ull) ? split + (uint32_t)__popcll(fresh & below)
                                    : (uint32_t)__popcll(~fresh & below);
      }
      const int32_t dx = r.x - d.ox, dy = r.y - d.oy, dz = r.z - d.oz;
      // The ray's length class (its cell count mx against the max range, in
      // 2^cb equal parts) sits between a coarse direction sector (hb bits)
      // and the direction inside the sector.  The rays of one wavefront then
      // end at about one distance (a batch waits for its longest ray) and,
      // ending together, cross the same leaves: short rays of one direction
      // hit near objects side by side, long ones sweep the far field.  The
      // sectors keep each region of a one-scan list (one per XCD) on one part
      // of the sphere, with a share of every length.
      // A scan with under 1/8 of its points clipped at the max range (as
      // estimated by the wavefronts that counted them), a scan of a room for
      // one, keeps the direction order: there neighbouring directions end at
      // about one distance already, and the classes would only split its
      // sweep of directions.
      const bool mixed = (uint64_t)sc.counters[kClipWords + (s & 63u)] * 128u >= (uint64_t)d.n;
      const RayOrder o = ray_order(mixed ? sc.ray_order : 0u);
      const uint32_t db = 2u * o.b - o.hb;  // direction bits below the sector
      const uint32_t mx = (uint32_t)max(abs(dx), max(abs(dy), abs(dz)));
      const uint32_t lc =
          d.reach ? min((1u << o.cb) - 1u, (uint32_t)(((uint64_t)mx << o.cb) / (d.reach + 1u))) : 0u;
      const uint32_t dir = dir_bin(dx, dy, dz, o.b);
      key = (tag << kDirBits) | ((dir >> db) << (o.cb + db)) | (lc << db) | (dir & ((1u << db) - 1u));
    }
    sc.wkey[gi] = win ? key : kNone;
  }
  wave_add_by_key(sc.hist, key, win);
}

// inclusive prefix sum over the 64 lanes of a wavefront
__device__ __forceinline__ uint32_t wave_scan(uint32_t v) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (uint32_t o = 1; o < 64u; o <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// Phase 2s (per bucket tile): sum of the winners of each tile of kScanTile
// buckets, four buckets a thread.
__global__ void __launch_bounds__(1024) k_scan_bucket_reduce(const BatchHdr* __restrict__ hdr,
                                                             ScanScratch sc) {
  __shared__ uint32_t wsum[16];
  const uint32_t nb = 1u << (kDirBits + hdr->tag_bits);
  const uint32_t base = blockIdx.x * kScanTile;
  if (base >= nb) return;
  const uint4 v = reinterpret_cast<const uint4*>(sc.hist + base)[threadIdx.x];
  const uint32_t s = wave_scan(v.x + v.y + v.z + v.w);
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  if (lane == 63) wsum[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < 16; ++w) t += wsum[w];
    sc.counters[kTileWords + blockIdx.x] = t;
  }
}

// Phase 2s' (per bucket tile): exclusive scan of the bucket counts into the
// bucket write cursors, carried over the earlier tiles; the counts are zeroed
// for the next call and the last tile writes the winner count to counters[0].
__global__ void __launch_bounds__(1024) k_scan_bucket_scan(const BatchHdr* __restrict__ hdr,
                                                           ScanScratch sc) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry;
  const uint32_t nb = 1u << (kDirBits + hdr->tag_bits);
  const uint32_t base = blockIdx.x * kScanTile;
  if (base >= nb) return;
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  if (wid == 0) {
    // blockIdx.x < kTiles <= 64: one lane per earlier tile
    const uint32_t t = lane < blockIdx.x ? sc.counters[kTileWords + lane] : 0u;
    const uint32_t tot = (uint32_t)__shfl((int)wave_scan(t), 63);
    if (lane == 0) carry = tot;
  }
  uint4* h4 = reinterpret_cast<uint4*>(sc.hist + base);
  const uint4 v = h4[threadIdx.x];
  const uint32_t s = v.x + v.y + v.z + v.w;
  const uint32_t incl = wave_scan(s);
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  if (wid == 0) {
    const uint32_t w = lane < 16u ? wsum[lane] : 0u;
    const uint32_t wi = wave_scan(w);
    if (lane < 16u) wsum[lane] = wi;
  }
  __syncthreads();
  uint32_t run = carry + (wid ? wsum[wid - 1] : 0u) + incl - s;
  uint4 o;
  o.x = run; run += v.x;
  o.y = run; run += v.y;
  o.z = run; run += v.z;
  o.w = run;
  reinterpret_cast<uint4*>(sc.offs + base)[threadIdx.x] = o;
  h4[threadIdx.x] = make_uint4(0u, 0u, 0u, 0u);
  // the fresh scans' winners start at bucket split << kDirBits (a tile's
  // first bucket: kScanTile divides 2^kDirBits); with no fresh scan, at the end
  const uint32_t split = sc.counters[kFreshWords + 2];
  if (split != kNone && (split << kDirBits) == base && threadIdx.x == 0) sc.counters[kFreshWords + 3] = o.x;
  if (base + kScanTile >= nb && threadIdx.x == 1023) {
    sc.counters[0] = carry + wsum[15];
    if (split != kNone && (split << kDirBits) >= nb) sc.counters[kFreshWords + 3] = carry + wsum[15];
  }
}

// Phase 2t (per point): winners written into their buckets.
__global__ void __launch_bounds__(256) k_scan_scatter(const BatchHdr* __restrict__ hdr,
                                                      ScanScratch sc) {
  const uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t key = kNone;
  if (gi < hdr->total) key = sc.wkey[gi];
  const bool win = key != kNone;
  const uint32_t pos = wave_add_by_key(sc.offs, win ? key : 0u, win);
  if (win) sc.winners[pos] = (uint32_t)gi;
}

// The passes over the winners' rays (walk_winners, start_ray).
enum : int { kPassLeaves = 0, kPassUpdate = 1, kPassFresh = 2 };

// Start the ray of winner `idx`; false if it has no cells to walk (origin
// voxel endpoint, endpoint-only call, or a map whose call failed), or if the
// ray belongs to the other pipeline: with sc.fresh_pass set, the rays of a
// map that was empty when the scan started (leaf_base 0: every leaf the scan
// touches is new) take only the fresh pass, and the other maps' rays only the
// leaf and update passes.
// With lg > 0, work unit `unit` is segment unit & (2^lg - 1) of winner
// unit >> lg: segment s of n = 2^lg walks cells [ceil(s mx / n), ceil((s + 1)
// mx / n)), so segment 0 holds cell 0 and the segments cover the ray once.
// `first` tells the caller the unit holds the ray's cell 0.
__device__ __forceinline__ bool start_ray(const ScanDesc* __restrict__ sd, const ScanScratch& sc,
                                          uint32_t unit, uint32_t lg, MapView& m, RayWalk& w,
                                          RootCache& rc, int32_t& cur, uint32_t& lbase,
                                          int pass, bool& first, uint32_t& win) {
  // the leaf and fresh passes walk the endpoint's cell too (they create
  // leaves); the update pass only the ray's cells
  const bool leaf_pass = pass != kPassUpdate;
  const uint32_t idx = unit >> lg, seg = unit & ((1u << lg) - 1u);
  first = seg == 0;
  const uint32_t gi = sc.winners[idx];
  win = gi;
  const int4 r = sc.rec[gi];
  const ScanDesc& d = sd[(uint32_t)r.w >> 16];
  const bool no_rays = (d.flags & kScanNoRays) != 0;
  const bool at_origin = r.x == d.ox && r.y == d.oy && r.z == d.oz;
  // the update pass walks cells [0, mx) of the ray (none for an endpoint in
  // the origin's voxel or a scan without rays); the leaf pass also the
  // endpoint's cell mx, so that it creates the winner's endpoint leaf with
  // the ray's (only that cell for a scan without rays)
  if (!leaf_pass && (no_rays || at_origin)) return false;
  if (leaf_pass && no_rays && seg != 0) return false;
  m = load_view(d.map);
  if (pass != kPassLeaves && (map_status(m) & kFatal)) return false;
  cur = map_cur(m);
  lbase = *gp(&m.ctr->leaf_base);
  if (sc.fresh_pass ? ((pass == kPassFresh) != (lbase == 0)) : pass == kPassFresh) return false;
  w.init(d.ox, d.oy, d.oz, r.x, r.y, r.z);
  const int32_t last = w.mx + (leaf_pass ? 1 : 0);
  int32_t k0 = 0, k1 = last;
  if (no_rays) {
    k0 = w.mx;
  } else if (lg) {
    const int64_t n = 1ll << lg;
    k0 = (int32_t)(((int64_t)last * seg + n - 1) >> lg);
    k1 = (int32_t)(((int64_t)last * (seg + 1) + n - 1) >> lg);
  }
  if (k0 >= k1) return false;
  if (k0 == 0)
    w.kend = k1;  // from the origin (and mx may be 0: no division)
  else
    w.seek(k0, k1);
  rc.reset();
  return true;
}

constexpr uint32_t kChunk = 256;  // default of ScanScratch::chunk_max
// cells a lane may run ahead of its wavefront's slowest ray in lockstep
// batches: update pass and leaf pass (defaults; the tuning keys ray_window
// and leaf_window set them per context.  Measured on 64 scans re-inserted
// into filled maps: 8 beats 4 and 16 for the update pass, 32 beats 16 for
// the leaf pass)
constexpr int32_t kSyncWindowCells = 8;
constexpr int32_t kSyncWindowLeaves = 32;
constexpr uint32_t kLockstepRays = 1;  // rays a lane from which batches run in lockstep
constexpr uint32_t kMaxSegLog = 3;     // at most 8 segments a ray when rays are fewer than lanes

// Work-stealing walk over the winner list.  The list is split into kRegions
// ranges; a wavefront starts in the range of its XCD (workgroups are
// dispatched round-robin over the XCDs) and takes up to chunk_max rays per
// atomic on the range's cursor.  Every lane works on its own ray, one leaf
// per step.  With at least kLockstepRays rays a lane, the wavefront takes its
// next 64 rays only when all 64 are done (lockstep batches): consecutive
// winners are rays of one map in neighbouring directions (and, with the ray
// order of k_scan_winners, of about one length), so a batch's rays stay side
// by side and cross the same cache lines at about the same time, where lanes
// refilled one by one drift apart (measured on 64 scans a call: the update
// pass 23 % and the leaf pass 18 % faster).  With fewer rays (one scan a
// call) lanes refill one by one: there, a batch waiting for its longest ray
// leaves the GPU idle (measured again with length-ordered rays: the LiDAR
// sequence 26 % slower in lockstep).  A range found empty
// is flagged so that idle wavefronts stop touching its cursor.  kMode 0: leaf
// allocation pass; 1: update pass; 2: counted update; 3: fresh pass (maps
// empty at the scan's start: create and mark in one walk); 4: counted fresh
// pass.
// cursors[0..kRegions) are the cursors, cursors[kRegions..2*kRegions) the
// exhausted flags.
template <int kMode>
__device__ void walk_winners(const ScanDesc* __restrict__ sd, const ScanScratch& sc,
                             uint32_t* cursors) {
  if (kMode >= 3 && !sc.fresh_pass) return;
  // the pass's part of the winner list: with a split, the fresh scans'
  // winners end the list
  const uint32_t total = sc.counters[0];
  const bool split = sc.counters[kFreshWords + 2] != kNone;
  const uint32_t at = split ? sc.counters[kFreshWords + 3] : 0u;
  const uint32_t w0 = (split && kMode >= 3) ? at : 0u;
  const uint32_t nwin = split ? (kMode >= 3 ? total - at : at) : total;
  const uint32_t lane = lane_id();
  const uint32_t lanes = gridDim.x * blockDim.x;
  const bool lockstep = nwin >= kLockstepRays * lanes;
  // fewer rays than lanes (one scan a call): each ray is walked as 2^lg
  // segments by different lanes, so the short list still fills the GPU
  uint32_t lg = 0;
  while (!lockstep && lg < kMaxSegLog && ((uint64_t)nwin << (lg + 1)) <= lanes) ++lg;
  const uint32_t nunits = nwin << lg;
  uint32_t chunk_pos = 0, chunk_end = 0;
  // small lists are one region: splitting them only multiplies cursor traffic
  const uint32_t cmax = sc.chunk_max ? sc.chunk_max : kChunk;
  constexpr bool kCounted = kMode == 2 || kMode == 4;
  constexpr bool kFreshPass = kMode >= 3;
  constexpr int kPass = kMode == 0 ? kPassLeaves : (kFreshPass ? kPassFresh : kPassUpdate);
  const int32_t wcells = sc.win_cells > 0 ? sc.win_cells : kSyncWindowCells;
  const int32_t win = kMode == 0 ? (sc.win_leaves > 0 ? sc.win_leaves : kSyncWindowLeaves)
                                 : (kFreshPass && sc.win_fresh > 0 ? sc.win_fresh : wcells);
  const uint32_t nreg = nunits >= kRegions * 8 * cmax ? kRegions : 1u;
  // chunk size: kChunk units when there is work for every wavefront, down
  // to one unit a lane on small lists
  uint32_t chunk = (nunits / (gridDim.x * (blockDim.x / 64u))) & ~63u;
  chunk = chunk < 64u ? 64u : (chunk > cmax ? cmax : chunk);
  uint32_t region = blockIdx.x % nreg, tried = 0;
  bool first = false;
  bool exhausted = false;
  bool have = false;
  MapView m;
  RayWalk w;
  LeafWalk lw;  // the leaf pass walks leaves, not cells
  RootCache rc;
  int32_t cur = 0;
  uint32_t lbase = 0;
  uint32_t nupd = 0;
  uint32_t next_leaf = kNone;
  // fresh pass: root slot and inner-grid entry of the current cell's leaf
  int32_t slot = -1;
  uint32_t entry = kEmptyChild;
  bool stopped = false;
  bool ray_first = false;
  uint32_t wgi = 0;  // the ray's winner (its point index)
  // counted form: per-lane totals, flushed when the lane's map changes
  MapCounters* cmap = nullptr;
  uint64_t c_upd = 0, c_rays = 0, c_vis = 0;
  for (;;) {
    // lockstep: a new batch once every lane's ray is done (wave-uniform
    // control flow); rays are handed out until every lane has one or the
    // list is exhausted
    const bool refill = !lockstep || __ballot(have) == 0;
    for (;;) {
      const uint64_t need = __ballot(!have);
      if (!refill || need == 0 || exhausted) break;
      if (chunk_pos == chunk_end) {
        const uint32_t lo = (uint32_t)((uint64_t)nunits * region / nreg);
        const uint32_t hi = (uint32_t)((uint64_t)nunits * (region + 1) / nreg);
        uint32_t base = hi - lo;
        // plain reads first: an exhausted range costs no atomic
        if (lane == 0 && !*gp(&cursors[kRegions + region]) && *gp(&cursors[region]) < hi - lo)
          base = add32(&cursors[region], chunk);
        base = (uint32_t)__shfl((int)base, 0);
        if (base >= hi - lo) {
          if (lane == 0) *gp(&cursors[kRegions + region]) = 1u;
          if (++tried == nreg) {
            exhausted = true;
            break;
          }
          region = (region + 1) % nreg;
          continue;
        }
        chunk_pos = lo + base;
        chunk_end = chunk_pos + chunk < hi ? chunk_pos + chunk : hi;
        continue;
      }
      const uint32_t avail = chunk_end - chunk_pos;
      const uint32_t want = (uint32_t)__popcll(need);
      const uint32_t take = want < avail ? want : avail;
      const uint32_t rank = (uint32_t)__popcll(need & lanes_below());
      if (!have && rank < take) {
        have = start_ray(sd, sc, chunk_pos + rank + (w0 << lg), lg, m, w, rc, cur, lbase, kPass, first, wgi);
        nupd = 0;
        stopped = false;
        // the unit holds cell 0 of a ray with cells (the fresh pass also
        // walks origin-voxel endpoints and endpoint-only calls: no ray)
        ray_first = have && first && w.k < w.mx;
        if (kMode == 0 && have) lw.init(w, m.lb);
        if ((kMode == 1 || kMode == 2) && have) next_leaf = walk_leaf(m, w, rc);
        if (kFreshPass && have) fresh_entry(m, w, rc, slot, entry);
        if (kCounted && have && m.ctr != cmap) {
          if (cmap) {
            add64(&cmap->stat_ray_updates, c_upd);
            add64(&cmap->stat_rays, c_rays);
            add64(&cmap->stat_ray_visits, c_vis);
          }
          cmap = m.ctr;
          c_upd = c_rays = c_vis = 0;
        }
        if (kMode == 2 && have) c_vis += (uint32_t)(w.kend - w.k);
        // the fresh pass also walks the endpoint's cell mx, not a ray cell
        if (kMode == 4 && have) c_vis += (uint32_t)max(0, min(w.kend, w.mx) - w.k);
      }
      chunk_pos += take;
    }
    if (__ballot(have) == 0) break;
    // lockstep: a lane more than the window's cells ahead of the wavefront's
    // slowest ray waits a step, so the batch stays at about one distance
    // from the sensor (runs end at leaf borders, so rays drift apart
    // otherwise; measured: the update pass about 8 % faster)
    const int32_t kk = kMode == 0 ? lw.k : w.k;
    int32_t kmin = have ? kk : 0x7fffffff;
    bool go = have;
    if (lockstep && kMode == 0) {
      // the leaf pass: DPP minimum (it also frees the registers that keep
      // the pass at 5 waves a SIMD; the update pass measured no gain); a
      // wavefront with no lane free to move lets every lane with a ray move
      kmin = wave_min(kmin);
      go = have && kk < kmin + win;
      if (__ballot(go) == 0) go = have;
    } else if (lockstep) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const int32_t t = __shfl_xor(kmin, o);
        kmin = t < kmin ? t : kmin;
      }
      go = have && kk < kmin + win;
    }
    if (go) {
      if (kMode == 0) {
        leaves_step(m, lw, rc);
      } else if (kFreshPass) {
        nupd += fresh_step<kCounted>(m, w, rc, slot, entry, stopped, &sc.tpos[wgi]);
      } else {
        nupd += apply_step<kCounted>(m, w, rc, cur, lbase, next_leaf);
      }
      if (kMode == 0 ? lw.done() : (w.done() || stopped)) {
        have = false;
        if (kCounted) {
          c_upd += nupd;
          c_rays += ray_first ? 1u : 0u;
        }
      }
    }
  }
  if (kCounted && cmap) {
    add64(&cmap->stat_ray_updates, c_upd);
    add64(&cmap->stat_rays, c_rays);
    add64(&cmap->stat_ray_visits, c_vis);
  }
}

// Phase 2b: every leaf on every winner's ray is created before any cell is
// written (so a pool overflow leaves the map's cells untouched).
__global__ void __launch_bounds__(256) k_scan_ray_leaves(const ScanDesc* __restrict__ sd,
                                                         ScanScratch sc) {
  walk_winners<0>(sd, sc, &sc.counters[1]);
}

// Phase 2c (one block per scan): leaves of the rays still due from
// addHitPoint / addMissPoint calls made since the map's last insert.
__global__ void __launch_bounds__(256) k_scan_extra_leaves(const BatchHdr* __restrict__ hdr,
                                                           const ScanDesc* __restrict__ sd) {
  if (blockIdx.x >= hdr->nscans) return;
  const ScanDesc& d = sd[blockIdx.x];
  if (d.flags & kScanNoRays) return;
  const MapView m = load_view(d.map);
  for (uint32_t j = threadIdx.x; j < d.nextra; j += blockDim.x) {
    const int4 e = d.extra[j];
    if (e.x == d.ox && e.y == d.oy && e.z == d.oz) continue;
    RayWalk w;
    RootCache rc;
    rc.reset();
    w.init(d.ox, d.oy, d.oz, e.x, e.y, e.z);
    LeafWalk lw;
    lw.init(w, m.lb);
    while (!lw.done()) leaves_step(m, lw, rc);
  }
}

constexpr uint32_t kLeafLater = 0xFFFFFFFEu;  // tpos: look the endpoint's leaf up after the fresh pass

// Phase 3 (per winner): endpoint hit/miss update (probabilistic_map.cpp:30-54)
// of the endpoints in leaves that existed before the scan, and release of
// the endpoint-table slot.  One writer per endpoint cell.  An endpoint in a
// leaf the scan created is left for k_scan_new_endpoints (after the ray
// pass): its leaf index goes to tpos.
__global__ void __launch_bounds__(256) k_scan_endpoints(const ScanDesc* __restrict__ sd,
                                                        ScanScratch sc) {
  const uint32_t nwin = sc.counters[0];
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t start = blockIdx.x * blockDim.x + threadIdx.x;
  // every lane runs the same number of iterations so the count reduction
  // below sees a converged wavefront
  for (uint32_t base = start - (start & 63u); base < nwin; base += stride) {
    const uint32_t idx = base + (start & 63u);
    uint32_t updated = 0;
    MapCounters* ctr = nullptr;
    if (idx < nwin) {
      const uint32_t gi = sc.winners[idx];
      const int4 r = sc.rec[gi];
      const uint32_t t = sc.tpos[gi];
      EpSlot BXG_G* e = gp(&sc.ep[t]);
      e->key = kEmptyKey;
      e->first = kNone;
      const ScanDesc& d = sd[(uint32_t)r.w >> 16];
      const MapView m = load_view(d.map);
      const bool failed = (map_status(m) & kFatal) != 0;
      const uint32_t lbase = *gp(&m.ctr->leaf_base);
      // in a map empty at the scan's start every endpoint leaf is new: the
      // fresh pass, which runs after this kernel, creates it and stores its
      // index here (k_scan_new_endpoints looks up any it did not), so there
      // is nothing to look up now
      const uint32_t leaf = (failed || lbase == 0) ? kNone : find_leaf(m, r.x, r.y, r.z);
      const bool fresh = leaf != kNone && leaf >= lbase;
      sc.tpos[gi] = fresh ? leaf : (lbase == 0 && !failed ? kLeafLater : kNone);
      if (leaf != kNone && !fresh) {
        const uint32_t li = leaf_index(r.x, r.y, r.z, m.lb);
        uint32_t BXG_G* cell = (uint32_t BXG_G*)m.cells + (uint64_t)leaf * m.lsz + li;
        const uint32_t w = *cell;
        const int32_t cur = map_cur(m);
        *cell = (r.w & kRecHit) ? cell_hit(w, m.hit, m.cmax, cur) : cell_miss(w, m.miss, m.cmin, cur);
        if (w == 0) or64((uint64_t*)&m.masks[(uint64_t)leaf * m.lw + (li >> 6)], 1ull << (li & 63));
        updated = 1;
        ctr = m.ctr;
      }
    }
    // one counter update per distinct map among the wavefront's lanes
    uint64_t pending = __ballot(updated);
    while (pending) {
      const int leader = __ffsll((unsigned long long)pending) - 1;
      MapCounters* lc = (MapCounters*)__shfl((unsigned long long)ctr, leader);
      const uint64_t same = __ballot(updated && ctr == lc);
      if ((int)lane_id() == leader) add64(&lc->stat_endpoint_updates, (uint64_t)__popcll(same));
      pending &= ~same;
    }
  }
}

// Phase 4d (per winner): the endpoint update of the endpoints in leaves the
// scan created (k_scan_endpoints left their leaf index in tpos, or
// kLeafLater where the fresh pass was still to create it).  Such a
// cell held zero before the scan, so its endpoint word is cell_hit(0) or
// cell_miss(0), whatever ray marked it meanwhile (the reference updates
// endpoints first and its rays skip them, probabilistic_map.cpp:30-54,
// 81-89): its touched byte becomes 2 (hit) or 1 (miss), one writer per
// voxel, and k_scan_new_leaves writes the word.  The counted form takes back
// the ray update the ray pass counted for a cell that turns out to be an
// endpoint.
template <bool kCount>
__global__ void __launch_bounds__(256) k_scan_new_endpoints(const ScanDesc* __restrict__ sd,
                                                            ScanScratch sc) {
  const uint32_t nwin = sc.counters[0];
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t start = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t base = start - (start & 63u); base < nwin; base += stride) {
    const uint32_t idx = base + (start & 63u);
    uint32_t updated = 0, retracted = 0;
    MapCounters* ctr = nullptr;
    if (idx < nwin) {
      const uint32_t gi = sc.winners[idx];
      uint32_t leaf = sc.tpos[gi];
      const int4 r = sc.rec[gi];
      const MapView m = load_view(sd[(uint32_t)r.w >> 16].map);
      if (leaf == kLeafLater) leaf = (map_status(m) & kFatal) ? kNone : find_leaf(m, r.x, r.y, r.z);
      if (leaf != kNone) {
        uint8_t BXG_G* tb = m.touched + (uint64_t)leaf * m.lsz + leaf_index(r.x, r.y, r.z, m.lb);
        if (kCount) retracted = *tb != 0u ? 1u : 0u;
        *tb = (r.w & kRecHit) ? 2 : 1;
        updated = 1;
        ctr = m.ctr;
      }
    }
    uint64_t pending = __ballot(updated);
    while (pending) {
      const int leader = __ffsll((unsigned long long)pending) - 1;
      MapCounters* lc = (MapCounters*)__shfl((unsigned long long)ctr, leader);
      const uint64_t same = __ballot(updated && ctr == lc);
      const uint64_t back = __ballot(retracted && ctr == lc);
      if ((int)lane_id() == leader) {
        add64(&lc->stat_endpoint_updates, (uint64_t)__popcll(same));
        if (kCount && back) add64(&lc->stat_ray_updates, (uint64_t)0 - (uint64_t)__popcll(back));
      }
      pending &= ~same;
    }
  }
}

// Phase 4: the ray clear of updateFreeCells (probabilistic_map.cpp:77-101):
// every cell of Ray(origin, endpoint) takes the miss update unless its
// update_id already equals the scan's id.
template <bool kCount>
__global__ void __launch_bounds__(256) k_scan_rays(const ScanDesc* __restrict__ sd,
                                                   ScanScratch sc) {
  walk_winners<kCount ? 2 : 1>(sd, sc, &sc.counters[1 + 2 * kRegions]);
}

// Phase 4' (per winner of a map empty at the scan's start, with
// ScanScratch::fresh_pass): leaf creation and the ray clear in one walk (the
// leaf pass and the update pass skip these rays).
template <bool kCount>
__global__ void __launch_bounds__(256) k_scan_fresh_rays(const ScanDesc* __restrict__ sd,
                                                         ScanScratch sc) {
  walk_winners<kCount ? 4 : 3>(sd, sc, &sc.counters[1 + 4 * kRegions]);
}

// Phase 4b (one block per scan): rays due from addHitPoint / addMissPoint.
template <bool kCount>
__global__ void __launch_bounds__(256) k_scan_extra_rays(const BatchHdr* __restrict__ hdr,
                                                         const ScanDesc* __restrict__ sd) {
  if (blockIdx.x >= hdr->nscans) return;
  const ScanDesc& d = sd[blockIdx.x];
  if (d.flags & kScanNoRays) return;
  const MapView m = load_view(d.map);
  if (map_status(m) & kFatal) return;
  const int32_t cur = map_cur(m);
  const uint32_t lbase = *gp(&m.ctr->leaf_base);
  for (uint32_t j = threadIdx.x; j < d.nextra; j += blockDim.x) {
    const int4 e = d.extra[j];
    if (e.x == d.ox && e.y == d.oy && e.z == d.oz) continue;
    RayWalk w;
    RootCache rc;
    rc.reset();
    w.init(d.ox, d.oy, d.oz, e.x, e.y, e.z);
    uint32_t n = 0;
    uint32_t next_leaf = walk_leaf(m, w, rc);
    while (!w.done()) n += apply_step<kCount>(m, w, rc, cur, lbase, next_leaf);
    if (kCount) {
      add64(&m.ctr->stat_ray_updates, n);
      add64(&m.ctr->stat_rays, 1);
      add64(&m.ctr->stat_ray_visits, (uint32_t)w.mx);
    }
  }
}

// Phase 4c: the leaves the scan created, written whole.  Their cells were
// all off before the scan (a created leaf is initialised, as the reference's
// Grid constructor does, bonxai.hpp:51-58), and the scan left in each
// cell's touched byte what it did there: 0 nothing, 1 a miss (a ray, or a
// miss endpoint: cell_miss(0)), 2 a hit endpoint (cell_hit(0)).  So every
// cell word follows from its byte, and the mask is exactly the nonzero bytes
// (Accessor::value(create) sets the bit, bonxai.hpp:485-490): one pass per
// new leaf writes its cells and mask words and clears the bytes it read,
// with no read of the cell pool and no atomic.  The new leaves of each
// scan's map are [leaf_base, leaf_count); a map whose call failed wrote no
// byte, and its new leaves are written off (zero).
constexpr uint32_t kNewLeafRun = 4;  // leaves a wavefront of k_scan_new_leaves has in flight
__global__ void __launch_bounds__(256) k_scan_new_leaves(const BatchHdr* __restrict__ hdr,
                                                         const ScanDesc* __restrict__ sd) {
  const uint32_t lane = lane_id();
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t nscans = hdr->nscans;
  for (uint32_t s = 0; s < nscans; ++s) {
    const MapView m = load_view(sd[s].map);
    const MapCounters BXG_G* ctr = gp(m.ctr);
    const uint32_t lo = ctr->leaf_base;
    const uint32_t hi = min(ctr->leaf_count, m.leaf_cap);
    if (lo >= hi) continue;
    const int32_t cur = ctr->cur;
    // a failed call's leaves stay off: its marks (the fresh pass may stop
    // part-way) are cleared and give no cell
    const bool failed = (ctr->status & kFatal) != 0;
    const uint32_t wmiss = failed ? 0u : cell_miss(0u, m.miss, m.cmin, cur);
    const uint32_t whit = failed ? 0u : cell_hit(0u, m.hit, m.cmax, cur);
    if (m.lsz == 512) {
      // 8^3 leaves (the default leaf_bits 3): lane l has cells 4l..4l+3 and
      // 256+4l..256+4l+3, so each store instruction writes 1 KB in one piece
      // (measured: 2.98 ms a fresh frame's new leaves, against 3.63 ms with
      // 8 consecutive cells a lane).  The cell stores are nontemporal: a
      // frame's new leaves (8.4 GB) would only push its other data out of
      // the caches (2.52 ms against 2.98 ms).
      typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
      for (uint32_t leaf0 = lo + wave; leaf0 < hi; leaf0 += kNewLeafRun * nwaves) {
        uint32_t t[kNewLeafRun][2];
#pragma unroll
        for (uint32_t k = 0; k < kNewLeafRun; ++k) {
          const uint32_t leaf = leaf0 + k * nwaves;
          const uint32_t BXG_G* tp = (const uint32_t BXG_G*)(m.touched + (uint64_t)leaf * 512u) + lane;
          t[k][0] = leaf < hi ? tp[0] : 0u;
          t[k][1] = leaf < hi ? tp[64] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kNewLeafRun; ++k) {
          const uint32_t leaf = leaf0 + k * nwaves;
          if (leaf >= hi) break;  // uniform across the wavefront
#pragma unroll
          for (uint32_t h = 0; h < 2; ++h) {
            uint32_t v[4];
            uint32_t nib = 0;
#pragma unroll
            for (uint32_t b = 0; b < 4; ++b) {
              const uint32_t cls = (t[k][h] >> (8 * b)) & 0xFFu;
              v[b] = cls == 0u ? 0u : (cls == 1u ? wmiss : whit);
              nib |= (cls != 0u && !failed ? 1u : 0u) << b;
            }
            u32x4 BXG_G* cp = (u32x4 BXG_G*)(m.cells + (uint64_t)leaf * 2048u + 1024u * h) + lane;
            __builtin_nontemporal_store(u32x4{v[0], v[1], v[2], v[3]}, cp);
            if (t[k][h]) ((uint32_t BXG_G*)(m.touched + (uint64_t)leaf * 512u + 256u * h))[lane] = 0u;
            // mask word 4h + r: the nibbles of lanes 16r..16r+15, lane 16r + i
            // at bits 4i..4i+3; lanes 16r..16r+7 fill the low half, 16r+8..
            // 16r+15 the high half.  OR over each half-row of 8 lanes (DPP
            // quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror), then lane 16r
            // takes the high half from lane 16r+8 (row_ror:8).
            const uint32_t sh = 4u * (lane & 7u);
            uint32_t mlo = (lane & 8u) ? 0u : nib << sh;
            uint32_t mhi = (lane & 8u) ? nib << sh : 0u;
            mlo |= (uint32_t)__builtin_amdgcn_update_dpp((int)mlo, (int)mlo, 0xB1, 0xF, 0xF, false);
            mhi |= (uint32_t)__builtin_amdgcn_update_dpp((int)mhi, (int)mhi, 0xB1, 0xF, 0xF, false);
            mlo |= (uint32_t)__builtin_amdgcn_update_dpp((int)mlo, (int)mlo, 0x4E, 0xF, 0xF, false);
            mhi |= (uint32_t)__builtin_amdgcn_update_dpp((int)mhi, (int)mhi, 0x4E, 0xF, 0xF, false);
            mlo |= (uint32_t)__builtin_amdgcn_update_dpp((int)mlo, (int)mlo, 0x141, 0xF, 0xF, false);
            mhi |= (uint32_t)__builtin_amdgcn_update_dpp((int)mhi, (int)mhi, 0x141, 0xF, 0xF, false);
            mhi = (uint32_t)__builtin_amdgcn_update_dpp((int)mhi, (int)mhi, 0x128, 0xF, 0xF, false);
            if ((lane & 15u) == 0u)
              m.masks[(uint64_t)leaf * 8u + 4u * h + (lane >> 4)] = ((uint64_t)mhi << 32) | mlo;
          }
        }
      }
    } else {
      for (uint32_t leaf = lo + wave; leaf < hi; leaf += nwaves) {
        uint8_t BXG_G* tb = m.touched + (uint64_t)leaf * m.lsz;
        uint32_t BXG_G* cells = (uint32_t BXG_G*)m.cells + (uint64_t)leaf * m.lsz;
        for (uint32_t w = 0; w < m.lw; ++w) {
          const uint32_t i = w * 64u + lane;
          const uint32_t cls = i < m.lsz ? tb[i] : 0u;
          if (i < m.lsz) {
            cells[i] = cls == 0u ? 0u : (cls == 1u ? wmiss : whit);
            if (cls) tb[i] = 0;
          }
          const uint64_t b = __ballot(cls != 0u && !failed);
          if (lane == 0) m.masks[(uint64_t)leaf * m.lw + w] = b;
        }
      }
    }
  }
}

// Allocation-free scans (bxg_insert_point_cloud_noalloc): the batch header
// and the one scan descriptor, passed by value, are written into the
// caller-side scratch; a counted call also clears the map's scan statistics.
__global__ void k_na_setup(BatchHdr hdr, ScanDesc desc, BatchHdr* __restrict__ hdr_out,
                           ScanDesc* __restrict__ desc_out, MapCounters* __restrict__ ctr,
                           uint32_t reset_stats) {
  if (threadIdx.x != 0) return;
  *hdr_out = hdr;
  *desc_out = desc;
  if (reset_stats) {
    ctr->stat_rays = 0;
    ctr->stat_ray_updates = 0;
    ctr->stat_ray_visits = 0;
  }
}

// After a scan that ran out of nodes (it wrote no cell): the overflow bits
// are cleared and the bump counter, which ran past the pool, comes back to
// the pool's size.  The leaves allocated below it stay, every cell off.
__global__ void k_na_recover(MapCounters* __restrict__ ctr, uint32_t leaf_cap) {
  if (threadIdx.x != 0) return;
  ctr->status &= ~(kStOutOfPool | kStRootFull);
  ctr->leaf_count = min(ctr->leaf_count, leaf_cap);
}

// Phase 0 (per scan): the leaf count at the start of the scan, so that later
// phases can tell the leaves the scan creates from older ones.
// With the fresh pass, also the scans whose map is empty (their winners get
// the highest bucket tags: the fresh pass walks the end of the list, the
// other passes the rest).
__global__ void k_scan_begin(const BatchHdr* __restrict__ hdr, const ScanDesc* __restrict__ sd,
                             ScanScratch sc) {
  const uint32_t nscans = hdr->nscans;
  for (uint32_t s = threadIdx.x; s < nscans; s += blockDim.x) {
    const MapView m = load_view(sd[s].map);
    MapCounters BXG_G* ctr = gp(m.ctr);
    const uint32_t base = min(ctr->leaf_count, m.leaf_cap);
    ctr->leaf_base = base;
    if (nscans <= 64u) {  // the first wavefront holds every scan
      const uint64_t fresh = __ballot(sc.fresh_pass && base == 0);
      if (threadIdx.x == 0) {
        sc.counters[kFreshWords] = (uint32_t)fresh;
        sc.counters[kFreshWords + 1] = (uint32_t)(fresh >> 32);
        sc.counters[kFreshWords + 2] = sc.fresh_pass ? nscans - (uint32_t)__popcll(fresh) : kNone;
      }
    }
  }
  if ((nscans > 64u || nscans == 0u) && threadIdx.x == 0) sc.counters[kFreshWords + 2] = kNone;
  if (threadIdx.x < 64u) sc.counters[kClipWords + threadIdx.x] = 0;
}

// Phase 5 (per scan): advance update_count 1 -> 2 -> 3 -> 1
// (probabilistic_map.cpp:103-105) unless the scan failed; reset the batch;
// copy every scanned map's counters to its pinned host mirror, so the host
// learns pool use and errors from one event after the call (no copy a map).
__global__ void k_scan_finalize(const BatchHdr* __restrict__ hdr, const ScanDesc* __restrict__ sd,
                                ScanScratch sc) {
  for (uint32_t s = threadIdx.x; s < hdr->nscans; s += blockDim.x) {
    const MapDev BXG_G* md = (const MapDev BXG_G*)sd[s].map;
    MapCounters BXG_G* ctr = gp(md->ctr);
    if (!(sd[s].flags & kScanNoRays) && !(ctr->status & kFatal)) {
      const int32_t c = ctr->cur + 1;
      ctr->cur = c == 4 ? 1 : c;
      ctr->scans_applied += 1;
    }
    const uint64_t BXG_G* src = (const uint64_t BXG_G*)ctr;
    uint64_t* dst = (uint64_t*)md->ctr_host;
    uint64_t v[sizeof(MapCounters) / 8];
#pragma unroll
    for (uint32_t k = 0; k < sizeof(MapCounters) / 8; ++k) v[k] = src[k];
#pragma unroll
    for (uint32_t k = 0; k < sizeof(MapCounters) / 8; ++k) dst[k] = v[k];
  }
  if (threadIdx.x <= 6 * kRegions) sc.counters[threadIdx.x] = 0;
}

// ------------------------------------------------------ bulk cell ops

// One slot of a bulk call's cell table: the cell's key and the first and
// last input index naming it, on one 16-byte line so that a claim touches
// one cache line.  An empty slot is (all ones, all ones, 0).
struct alignas(16) OpSlot {
  uint64_t key;
  uint32_t first;
  uint32_t last;
};

struct OpScratch {
  OpSlot* slots;     // the call's table (mask + 1 slots)
  uint32_t mask;
  uint32_t* tpos;    // per input: table slot (kNone if no cell)
  uint32_t* leaf;    // per input: its leaf index, or kNone where not known yet
  uint8_t* prebit;   // per input: mask bit before the call
};

// empty the first n slots of a call's table
__global__ void __launch_bounds__(256) k_op_clear(OpSlot* __restrict__ slots, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    OpSlot e;
    e.key = kEmptyKey;
    e.first = kNone;
    e.last = 0;
    slots[i] = e;
  }
}

// The leaf count when a bulk call starts: the leaves at or above it are the
// call's, created with undefined bytes.
__global__ void k_op_begin(const MapDev* __restrict__ mp) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const MapView m = load_view(mp);
  MapCounters BXG_G* ctr = gp(m.ctr);
  ctr->leaf_base = min(ctr->leaf_count, m.leaf_cap);
}

// The leaves a bulk call created, initialised before any cell of them is
// written: mask words and cell bytes zero (every cell off), as the
// reference's Grid constructor leaves a new leaf (bonxai.hpp:51-58).  Grid-
// stride over 8-byte words (a leaf's cells are a multiple of 8 bytes).
__global__ void __launch_bounds__(256) k_zero_new_leaves(const MapDev* __restrict__ mp) {
  const MapView m = load_view(mp);
  const MapCounters BXG_G* ctr = gp(m.ctr);
  const uint64_t lo = ctr->leaf_base;
  const uint64_t hi = min(ctr->leaf_count, m.leaf_cap);
  if (lo >= hi) return;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t start = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nm = (hi - lo) * m.lw;
  for (uint64_t i = start; i < nm; i += stride) m.masks[lo * m.lw + i] = 0;
  const uint64_t cw = (uint64_t)m.lsz * m.vb / 8;  // 8-byte words a leaf
  uint64_t BXG_G* cells = (uint64_t BXG_G*)m.cells + lo * cw;
  const uint64_t nc = (hi - lo) * cw;
  for (uint64_t i = start; i < nc; i += stride) cells[i] = 0;
}

// create the root and leaf of every coordinate (setValue / setCellOn path)
__global__ void __launch_bounds__(256) k_touch(const MapDev* __restrict__ mp,
                                               const int32_t* __restrict__ coords, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MapView m = load_view(mp);
  const int32_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  uint64_t key;
  if (!pack_root(x, y, z, m.log2n, &key)) {
    or32(&m.ctr->status, kStCoordRange);
    return;
  }
  const int32_t slot = root_insert(m, key);
  if (slot < 0) return;
  const int32_t lmask = ~((1 << m.lb) - 1);
  leaf_request(m, (uint32_t)slot, inner_index(x, y, z, m.lb, m.ib), x & lmask, y & lmask, z & lmask);
}

// First and last input index of every distinct cell, and its prior mask
// bit, in one pass with the creation of missing roots and leaves (kCreate:
// setValue / setCellOn; setCellOff creates nothing and skips absent cells).
// A cell's table key is its root slot, inner index and leaf index, known
// before its leaf is, so a lane never waits for another lane's leaf: a leaf
// still being created (or created in this call: index at or above the
// call's leaf_base, its bytes not yet initialised) has all mask bits off.
// kFirst: the call needs each cell's first input and prior mask bit (was_on
// is asked for, or setCellOn / setCellOff); a setValue batch without was_on
// needs only the last input of each cell (last write wins).
template <bool kCreate, bool kFirst>
__global__ void __launch_bounds__(256) k_op_resolve(const MapDev* __restrict__ mp,
                                                    const int32_t* __restrict__ coords,
                                                    uint64_t n, OpScratch os) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MapView m = load_view(mp);
  const int32_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  os.tpos[i] = kNone;
  if (kFirst) os.prebit[i] = 0;
  uint64_t rkey;
  if (!pack_root(x, y, z, m.log2n, &rkey)) {
    if (kCreate) or32(&m.ctr->status, kStCoordRange);
    return;
  }
  const int32_t slot = kCreate ? root_insert(m, rkey) : root_find(m, rkey);
  if (slot < 0) return;
  const uint32_t ci = inner_index(x, y, z, m.lb, m.ib);
  const uint32_t li = leaf_index(x, y, z, m.lb);
  uint32_t leaf;
  if (kCreate) {
    const int32_t lmask = ~((1 << m.lb) - 1);
    leaf = leaf_request(m, (uint32_t)slot, ci, x & lmask, y & lmask, z & lmask);
    if (leaf == kEmptyChild) return;  // pool exhausted: the call fails
  } else {
    leaf = m.children[(uint64_t)slot * m.isz + ci];
    if (leaf >= kPendingChild) return;
  }
  // a leaf still being created by another lane: its index is looked up by
  // k_op_apply, after this kernel
  os.leaf[i] = leaf < kPendingChild ? leaf : kNone;
  if (kFirst && leaf < *gp(&m.ctr->leaf_base))
    os.prebit[i] = (uint8_t)((m.masks[(uint64_t)leaf * m.lw + (li >> 6)] >> (li & 63)) & 1);
  const uint64_t key = ((uint64_t)(uint32_t)slot << 24) | ((uint64_t)ci << 12) | (uint64_t)li;
  uint32_t h = (uint32_t)mix64(key) & os.mask;
  for (uint32_t p = 0; p <= os.mask; ++p) {
    uint64_t k = gp(&os.slots[h])->key;
    if (k == kEmptyKey) {
      k = cas64(&os.slots[h].key, kEmptyKey, key);
      if (k == kEmptyKey) k = key;
    }
    if (k == key) {
      if (kFirst) min32(&os.slots[h].first, (uint32_t)i);
      max32(&os.slots[h].last, (uint32_t)i);
      os.tpos[i] = h;
      return;
    }
    h = (h + 1) & os.mask;
  }
}

enum : uint32_t { kOpSetValue = 0, kOpSetOn = 1, kOpSetOff = 2 };

__device__ __forceinline__ void copy_bytes(uint8_t BXG_G* dst, const uint8_t BXG_G* src, uint32_t vb) {
  for (uint32_t b = 0; b < vb; ++b) dst[b] = src[b];
}

// Sequential-order semantics of a batch of setValue / setCellOn / setCellOff
// (bonxai.hpp:449-466, 537-569): was_on of an input is the prior bit for the
// first occurrence of its cell and "on" (off for setCellOff) for later ones;
// setValue keeps the last occurrence's value, setCellOn the first one's.
__global__ void __launch_bounds__(256) k_op_apply(const MapDev* __restrict__ mp,
                                                  const int32_t* __restrict__ coords,
                                                  const uint8_t* __restrict__ vals, uint64_t n,
                                                  uint32_t op, uint8_t* __restrict__ was_on,
                                                  OpScratch os) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MapView m = load_view(mp);
  const uint32_t t = os.tpos[i];
  if (t == kNone || (map_status(m) & kFatal)) {  // no cell, or the call failed
    if (was_on) was_on[i] = 0;
    return;
  }
  const uint32_t last = os.slots[t].last;
  // first and the prior bit exist only where k_op_resolve was asked for them
  const bool need_first = was_on != nullptr || op != kOpSetValue;
  const uint32_t first = need_first ? os.slots[t].first : 0u;
  const uint8_t pre = need_first ? os.prebit[i] : (uint8_t)0;
  const int32_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  const uint32_t known = os.leaf[i];
  const uint32_t leaf = known != kNone ? known : find_leaf(m, x, y, z);
  if (leaf == kNone) {
    if (was_on) was_on[i] = 0;
    return;
  }
  const uint32_t li = leaf_index(x, y, z, m.lb);
  const uint64_t g = (uint64_t)leaf * m.lsz + li;
  uint64_t* mw = (uint64_t*)&m.masks[(uint64_t)leaf * m.lw + (li >> 6)];
  const uint64_t bit = 1ull << (li & 63);
  const uint32_t vb = m.vb;
  if (op == kOpSetOff) {
    if (was_on) was_on[i] = (uint8_t)((uint32_t)i == first ? pre : 0);
    if ((uint32_t)i == last) {
      and64(mw, ~bit);
      for (uint32_t b = 0; b < vb; ++b) m.cells[g * vb + b] = 0;
    }
    return;
  }
  if (was_on) was_on[i] = (uint8_t)((uint32_t)i == first ? pre : 1);
  if (op == kOpSetValue) {
    if ((uint32_t)i == last) {
      if (vb == 4)
        ((uint32_t BXG_G*)m.cells)[g] = gp((const uint32_t*)vals)[i];
      else if (vb)
        copy_bytes(m.cells + g * vb, gp(vals) + i * vb, vb);
      or64(mw, bit);
    }
  } else {  // setCellOn: the first occurrence decides, and writes only if off
    if ((uint32_t)i == first) {
      if (!pre && vb && vals) copy_bytes(m.cells + g * vb, gp(vals) + i * vb, vb);
      or64(mw, bit);
    }
  }
}


// ConstAccessor::value / isCellOn (bonxai.hpp:496-534) for a batch.  Each
// thread carries kLookups lookups through the chain root key -> inner entry
// -> cell (-> mask word) together, so every stage issues kLookups
// independent loads (lookup j of a block is element base + j * blockDim.x:
// every stage's loads and stores stay coalesced across the wavefront).
constexpr uint32_t kLookups = 4;

__device__ __forceinline__ int32_t root_find_from(const MapView& m, uint64_t key, uint32_t h) {
  for (uint32_t p = 0; p <= m.root_mask; ++p) {
    const uint64_t k = m.root_keys[h];
    if (k == key) return (int32_t)h;
    if (k == kEmptyKey) return -1;
    h = (h + 1) & m.root_mask;
  }
  return -1;
}

__global__ void __launch_bounds__(256) k_lookup(const MapDev* __restrict__ mp,
                                                const int32_t* __restrict__ coords, uint64_t n,
                                                uint8_t* __restrict__ vals,
                                                uint8_t* __restrict__ found) {
  const MapView m = load_view(mp);
  const uint64_t base = (uint64_t)blockIdx.x * blockDim.x * kLookups + threadIdx.x;
  const int32_t BXG_G* cg = gp(coords);
  int32_t x[kLookups], y[kLookups], z[kLookups];
  uint64_t key[kLookups];
  uint32_t h[kLookups];
  bool live[kLookups];
#pragma unroll
  for (uint32_t j = 0; j < kLookups; ++j) {
    const uint64_t i = base + (uint64_t)j * blockDim.x;
    live[j] = i < n;
    x[j] = live[j] ? cg[3 * i] : 0;
    y[j] = live[j] ? cg[3 * i + 1] : 0;
    z[j] = live[j] ? cg[3 * i + 2] : 0;
  }
  uint64_t k0[kLookups];
#pragma unroll
  for (uint32_t j = 0; j < kLookups; ++j) {
    live[j] = live[j] && pack_root(x[j], y[j], z[j], m.log2n, &key[j]);
    h[j] = live[j] ? root_hash(key[j], m.root_mask) : 0u;
    k0[j] = live[j] ? m.root_keys[h[j]] : kEmptyKey;
  }
  uint32_t c[kLookups];
#pragma unroll
  for (uint32_t j = 0; j < kLookups; ++j) {
    int32_t slot = -1;
    if (k0[j] == key[j] && live[j]) {
      slot = (int32_t)h[j];
    } else if (live[j] && k0[j] != kEmptyKey) {
      slot = root_find_from(m, key[j], (h[j] + 1) & m.root_mask);  // a collision: probe on
    }
    c[j] = slot >= 0 ? m.children[(uint64_t)slot * m.isz + inner_index(x[j], y[j], z[j], m.lb, m.ib)]
                     : kEmptyChild;
  }
  uint32_t li[kLookups];
  uint64_t g[kLookups];
#pragma unroll
  for (uint32_t j = 0; j < kLookups; ++j) {
    li[j] = leaf_index(x[j], y[j], z[j], m.lb);
    g[j] = (uint64_t)(c[j] < kPendingChild ? c[j] : 0u) * m.lsz + li[j];
  }
  const uint32_t vb = m.vb;
  // 4- and 8-byte cells: the cell first.  Off cells hold zero bytes, so a
  // nonzero cell is on and its mask word is not read; only a zero cell (off,
  // or on with a zero value) costs the mask load.
  uint64_t cv[kLookups];
  const bool cell_first = vb == 4 || vb == 8;
#pragma unroll
  for (uint32_t j = 0; j < kLookups; ++j) {
    cv[j] = 0;
    if (cell_first && c[j] < kPendingChild)
      cv[j] = vb == 4 ? (uint64_t)((const uint32_t BXG_G*)m.cells)[g[j]]
                      : ((const uint64_t BXG_G*)m.cells)[g[j]];
  }
  bool on[kLookups];
#pragma unroll
  for (uint32_t j = 0; j < kLookups; ++j) {
    on[j] = cv[j] != 0;
    if (!on[j] && c[j] < kPendingChild)
      on[j] = (m.masks[(uint64_t)c[j] * m.lw + (li[j] >> 6)] >> (li[j] & 63)) & 1ull;
  }
#pragma unroll
  for (uint32_t j = 0; j < kLookups; ++j) {
    const uint64_t i = base + (uint64_t)j * blockDim.x;
    if (i >= n) continue;
    if (found) found[i] = on[j];
    if (!vals || !vb) continue;
    if (vb == 4) {
      ((uint32_t*)vals)[i] = (uint32_t)cv[j];
    } else if (vb == 8) {
      ((uint64_t*)vals)[i] = cv[j];
    } else {
      for (uint32_t b = 0; b < vb; ++b) vals[i * vb + b] = on[j] ? m.cells[g[j] * vb + b] : 0;
    }
  }
}

// isOccupied / isUnknown / isFree (probabilistic_map.cpp:56-75) for a batch:
// 1 occupied, -1 free, 0 unknown
__global__ void __launch_bounds__(256) k_classify_cells(const MapDev* __restrict__ mp,
                                                        const int32_t* __restrict__ coords,
                                                        uint64_t n, int8_t* __restrict__ state) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MapView m = load_view(mp);
  const int32_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  const uint32_t leaf = find_leaf(m, x, y, z);
  int8_t st = 0;
  if (leaf != kNone) {
    const uint32_t li = leaf_index(x, y, z, m.lb);
    if ((m.masks[(uint64_t)leaf * m.lw + (li >> 6)] >> (li & 63)) & 1) {
      const int32_t p = cell_p(((const uint32_t BXG_G*)m.cells)[(uint64_t)leaf * m.lsz + li]);
      st = p > m.thr ? 1 : (p < m.thr ? -1 : 0);
    }
  }
  state[i] = st;
}

// ------------------------------------------------------ whole-map passes

// forEachCell (bonxai.hpp:704-743) as a compaction: one thread per cell of
// every allocated leaf; filter 0 all, 1 p > threshold, 2 p < threshold.
__global__ void __launch_bounds__(256) k_export(const MapDev* __restrict__ mp, uint32_t nleaves,
                                                uint32_t filter, int32_t* __restrict__ coords,
                                                uint8_t* __restrict__ vals, uint64_t cap,
                                                uint64_t* __restrict__ count) {
  const MapView m = load_view(mp);
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t leaf = t >> (3 * m.lb);
  const uint32_t li = (uint32_t)(t & (m.lsz - 1));
  bool pass = false;
  if (leaf < nleaves) {
    pass = (m.masks[leaf * m.lw + (li >> 6)] >> (li & 63)) & 1;
    if (pass && filter) {
      const int32_t p = cell_p(((const uint32_t BXG_G*)m.cells)[leaf * m.lsz + li]);
      pass = filter == 1 ? p > m.thr : p < m.thr;
    }
  }
  const uint64_t pos = wave_append64(count, pass);
  if (!pass || pos >= cap) return;
  const int4 o = m.origin[leaf];
  const uint32_t lb = m.lb, lm = (1u << lb) - 1;
  if (coords) {
    coords[3 * pos] = o.x | (int32_t)(li & lm);
    coords[3 * pos + 1] = o.y | (int32_t)((li >> lb) & lm);
    coords[3 * pos + 2] = o.z | (int32_t)((li >> (2 * lb)) & lm);
  }
  const uint32_t vb = m.vb;
  if (vals && vb) copy_bytes(gp(vals) + pos * vb, m.cells + (leaf * m.lsz + li) * vb, vb);
}

// order-free digest of the on cells (count, sum, xor, sum of remix)
__global__ void __launch_bounds__(256) k_digest(const MapDev* __restrict__ mp, uint32_t nleaves,
                                                uint64_t* __restrict__ out) {
  const MapView m = load_view(mp);
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t leaf = t >> (3 * m.lb);
  const uint32_t li = (uint32_t)(t & (m.lsz - 1));
  uint64_t h = 0, h2 = 0, c = 0;
  if (leaf < nleaves && ((m.masks[leaf * m.lw + (li >> 6)] >> (li & 63)) & 1)) {
    const int4 o = m.origin[leaf];
    const uint32_t lb = m.lb, lm = (1u << lb) - 1, vb = m.vb;
    uint64_t word = 0;
    const uint8_t BXG_G* src = m.cells + (leaf * m.lsz + li) * vb;
    if (vb <= 8) {
      for (uint32_t b = 0; b < vb; ++b) word |= (uint64_t)src[b] << (8 * b);
    } else {
      for (uint32_t off = 0; off < vb; off += 8) {
        uint64_t part = 0;
        for (uint32_t b = 0; b < 8 && off + b < vb; ++b) part |= (uint64_t)src[off + b] << (8 * b);
        word = mix64(word ^ part);
      }
    }
    h = cell_hash(o.x | (int32_t)(li & lm), o.y | (int32_t)((li >> lb) & lm),
                  o.z | (int32_t)((li >> (2 * lb)) & lm), word);
    h2 = mix64(h ^ 0xA0761D6478BD642Full);
    c = 1;
  }
  // wavefront reduction, then one atomic per wavefront
  uint64_t x = h;
  for (int off = 32; off > 0; off >>= 1) {
    c += __shfl_xor(c, off);
    h += __shfl_xor(h, off);
    x ^= __shfl_xor(x, off);
    h2 += __shfl_xor(h2, off);
  }
  if (lane_id() == 0 && c) {
    add64(&out[0], c);
    add64(&out[1], h);
    xor64(&out[2], x);
    add64(&out[3], h2);
  }
}

// activeCellsCount (bonxai.hpp:689-701): popcount of every leaf mask word
__global__ void __launch_bounds__(256) k_count_on(const MapDev* __restrict__ mp, uint32_t nleaves,
                                                  uint64_t* __restrict__ out) {
  const MapView m = load_view(mp);
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t c = 0;
  if (t < (uint64_t)nleaves * m.lw) {
    const uint64_t w = m.masks[t];
    c = m.lsz < 64 ? (uint64_t)__popcll(w & ((1ull << m.lsz) - 1)) : (uint64_t)__popcll(w);
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if (lane_id() == 0 && c) add64(out, c);
}

// clear(SET_ALL_CELLS_OFF): every mask word and every cell byte to zero
__global__ void __launch_bounds__(256) k_cells_off(const MapDev* __restrict__ mp, uint32_t nleaves) {
  const MapView m = load_view(mp);
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nw = (uint64_t)nleaves * m.lw;
  for (uint64_t w = t; w < nw; w += (uint64_t)gridDim.x * blockDim.x) m.masks[w] = 0;
  const uint64_t nb = (uint64_t)nleaves * m.lsz * m.vb / 4;
  for (uint64_t b = t; b < nb; b += (uint64_t)gridDim.x * blockDim.x) ((uint32_t BXG_G*)m.cells)[b] = 0;
}

// re-insert every root of an old directory into the (empty) directory of m;
// with prune set, roots without any leaf are dropped
__global__ void __launch_bounds__(256) k_rehash(const MapDev* __restrict__ mp,
                                                const uint64_t* __restrict__ old_keys,
                                                const uint32_t* __restrict__ old_children,
                                                uint32_t old_cap, uint32_t prune) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= old_cap) return;
  const uint64_t key = old_keys[s];
  if (key == kEmptyKey) return;
  const MapView m = load_view(mp);
  if (prune) {
    bool any = false;
    for (uint32_t c = 0; c < m.isz && !any; ++c)
      any = old_children[(uint64_t)s * m.isz + c] < kPendingChild;
    if (!any) return;
  }
  const int32_t ns = root_insert(m, key);
  if (ns < 0) return;
  for (uint32_t c = 0; c < m.isz; ++c) {
    const uint32_t leaf = old_children[(uint64_t)s * m.isz + c];
    m.children[(uint64_t)ns * m.isz + c] = leaf;
    if (leaf < kPendingChild) m.origin[leaf].w = ns;
  }
}

// ------------------------------------------ releaseUnusedMemory compaction

// live[k] = 1 if leaf k has any cell on (bonxai.hpp:367-387 keeps exactly those)
__global__ void __launch_bounds__(256) k_leaf_live(const MapDev* __restrict__ mp, uint32_t nleaves,
                                                   uint32_t* __restrict__ live) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const MapView m = load_view(mp);
  uint64_t any = 0;
  for (uint32_t w = 0; w < m.lw; ++w) any |= m.masks[(uint64_t)k * m.lw + w];
  live[k] = any != 0;
}

// Exclusive prefix sum of n 32-bit values in tiles of kScanTile: per-tile
// sums, then a one-block scan of the tile sums, then each tile with its carry.
__global__ void __launch_bounds__(1024) k_tile_reduce(const uint32_t* __restrict__ in, uint64_t n,
                                                      uint32_t* __restrict__ tiles) {
  __shared__ uint32_t wsum[16];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + 4ull * threadIdx.x;
  uint32_t s = 0;
  for (uint32_t j = 0; j < 4; ++j) s += base + j < n ? in[base + j] : 0u;
  s = wave_scan(s);
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  if (lane == 63) wsum[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < 16; ++w) t += wsum[w];
    tiles[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(1024) k_tile_sums_scan(uint32_t* __restrict__ tiles,
                                                         uint32_t ntiles, uint32_t* __restrict__ total) {
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry;
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t b = 0; b < ntiles; b += 1024) {
    const uint32_t i = b + threadIdx.x;
    const uint32_t v = i < ntiles ? tiles[i] : 0u;
    const uint32_t incl = wave_scan(v);
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    if (wid == 0) {
      const uint32_t w = lane < 16u ? wsum[lane] : 0u;
      const uint32_t wi = wave_scan(w);
      if (lane < 16u) wsum[lane] = wi;
    }
    __syncthreads();
    const uint32_t excl = carry + (wid ? wsum[wid - 1] : 0u) + incl - v;
    if (i < ntiles) tiles[i] = excl;
    __syncthreads();
    if (threadIdx.x == 0) carry += wsum[15];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(1024) k_tile_scan(const uint32_t* __restrict__ in, uint64_t n,
                                                    const uint32_t* __restrict__ tiles,
                                                    uint32_t* __restrict__ out) {
  __shared__ uint32_t wsum[16];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + 4ull * threadIdx.x;
  uint32_t v[4], s = 0;
  for (uint32_t j = 0; j < 4; ++j) {
    v[j] = base + j < n ? in[base + j] : 0u;
    s += v[j];
  }
  const uint32_t incl = wave_scan(s);
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  if (wid == 0) {
    const uint32_t w = lane < 16u ? wsum[lane] : 0u;
    const uint32_t wi = wave_scan(w);
    if (lane < 16u) wsum[lane] = wi;
  }
  __syncthreads();
  uint32_t run = tiles[blockIdx.x] + (wid ? wsum[wid - 1] : 0u) + incl - s;
  for (uint32_t j = 0; j < 4; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
}

// Move every live leaf k to index newidx[k] of the new pool arrays and point
// its inner-grid entry there; the entries of dead leaves become empty.
// One block per old leaf.
__global__ void __launch_bounds__(256) k_compact_leaves(
    const MapDev* __restrict__ mp, uint32_t nleaves, const uint32_t* __restrict__ live,
    const uint32_t* __restrict__ newidx, uint64_t* __restrict__ dmasks, uint8_t* __restrict__ dcells,
    int4* __restrict__ dorigin) {
  const uint32_t k = blockIdx.x;
  if (k >= nleaves) return;
  const MapView m = load_view(mp);
  const int4 o = m.origin[k];
  const bool on = live[k] != 0;
  const uint32_t nk = newidx[k];
  if (threadIdx.x == 0) {
    const uint32_t ii = inner_index(o.x, o.y, o.z, m.lb, m.ib);
    m.children[(uint64_t)o.w * m.isz + ii] = on ? nk : kEmptyChild;
    if (on) dorigin[nk] = o;
  }
  if (!on) return;
  for (uint32_t w = threadIdx.x; w < m.lw; w += blockDim.x)
    dmasks[(uint64_t)nk * m.lw + w] = m.masks[(uint64_t)k * m.lw + w];
  const uint64_t nb = (uint64_t)m.lsz * m.vb;
  if (nb % 4 == 0) {
    const uint32_t BXG_G* src = (const uint32_t BXG_G*)(m.cells + (uint64_t)k * nb);
    uint32_t* dst = (uint32_t*)(dcells + (uint64_t)nk * nb);
    for (uint64_t b = threadIdx.x; b < nb / 4; b += blockDim.x) dst[b] = src[b];
  } else {
    for (uint64_t b = threadIdx.x; b < nb; b += blockDim.x)
      dcells[(uint64_t)nk * nb + b] = m.cells[(uint64_t)k * nb + b];
  }
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ----------------------------------------------------- leaf page I/O

// Copy whole leaves (mask words + cell bytes) named by their origin coords
// into flat buffers; exists[k] = 0 where the leaf is absent (zeros copied).
__global__ void __launch_bounds__(256) k_read_leaves(const MapDev* __restrict__ mp,
                                                     const int32_t* __restrict__ origins, uint32_t n,
                                                     uint64_t* __restrict__ masks,
                                                     uint8_t* __restrict__ cells,
                                                     uint8_t* __restrict__ exists) {
  const MapView m = load_view(mp);
  const uint32_t k = blockIdx.x;
  if (k >= n) return;
  const uint32_t leaf = find_leaf(m, origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]);
  if (threadIdx.x == 0) exists[k] = leaf != kNone;
  for (uint32_t w = threadIdx.x; w < m.lw; w += blockDim.x)
    masks[(uint64_t)k * m.lw + w] = leaf != kNone ? m.masks[(uint64_t)leaf * m.lw + w] : 0;
  const uint64_t nb = (uint64_t)m.lsz * m.vb;
  for (uint64_t b = threadIdx.x; b < nb; b += blockDim.x)
    cells[k * nb + b] = leaf != kNone ? m.cells[leaf * nb + b] : 0;
}

// Overwrite whole leaves (they must exist: run k_touch on the origins first).
// Cells whose bit is off are written as zero bytes (the off-cell invariant).
__global__ void __launch_bounds__(256) k_write_leaves(const MapDev* __restrict__ mp,
                                                      const int32_t* __restrict__ origins,
                                                      uint32_t n, const uint64_t* __restrict__ masks,
                                                      const uint8_t* __restrict__ cells) {
  const MapView m = load_view(mp);
  const uint32_t k = blockIdx.x;
  if (k >= n) return;
  const uint32_t leaf = find_leaf(m, origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]);
  if (leaf == kNone) return;
  for (uint32_t w = threadIdx.x; w < m.lw; w += blockDim.x)
    m.masks[(uint64_t)leaf * m.lw + w] = masks[(uint64_t)k * m.lw + w];
  const uint32_t vb = m.vb;
  const uint64_t nb = (uint64_t)m.lsz * vb;
  for (uint64_t b = threadIdx.x; b < nb; b += blockDim.x) {
    const uint32_t li = (uint32_t)(b / vb);
    const bool on = (masks[(uint64_t)k * m.lw + (li >> 6)] >> (li & 63)) & 1;
    m.cells[leaf * nb + b] = on ? cells[k * nb + b] : 0;
  }
}

// origins of all allocated leaves (for paging a whole map to the host)
__global__ void __launch_bounds__(256) k_leaf_origins(const MapDev* __restrict__ mp, uint32_t nleaves,
                                                      int32_t* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const MapView m = load_view(mp);
  const int4 o = m.origin[k];
  out[3 * k] = o.x;
  out[3 * k + 1] = o.y;
  out[3 * k + 2] = o.z;
}

}  // namespace bxg
