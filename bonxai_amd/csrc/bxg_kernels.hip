// bxg_kernels.hip -- gfx950 kernels of the sparse voxel map.
//
// Built with -ffp-contract=off: the scan front end (insertPointCloud's
// max-range clip and posToCoord) must round every FP64 operation exactly as
// the reference's x86-64 double code, which has no fused multiply-add.
//
// Memory-model notes.  Cross-lane races are resolved with agent-scope atomics
// (root-slot claims, leaf claims, endpoint first-index claims, mask bits).
// The ray phase writes cell words with plain stores: during that phase every
// update of a cell is the same function of the same prior word (the miss
// update guarded by update_id), so concurrent writers store identical
// values, and a stale read can only return the prior word or the updated
// word, either of which leads to the same final value.
#include "bxg_internal.h"

namespace bxg {

// --------------------------------------------------------------- atomics

__device__ __forceinline__ uint32_t ld32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st32(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t cas32(uint32_t* p, uint32_t expect, uint32_t desired) {
  __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expect;
}
__device__ __forceinline__ uint64_t cas64(uint64_t* p, uint64_t expect, uint64_t desired) {
  __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return expect;
}
__device__ __forceinline__ uint32_t add32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t add64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t or32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t or64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t and64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t xor64(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_xor(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t min32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t max32(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Append from the active lanes of a wavefront with one atomic per wavefront;
// returns this lane's slot (valid only where pred is true).
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
  const uint64_t b = __ballot(pred);
  if (b == 0) return 0;
  const int leader = __ffsll((unsigned long long)b) - 1;
  uint32_t base = 0;
  if ((int)lane_id() == leader) base = add32(counter, (uint32_t)__popcll(b));
  base = __shfl(base, leader);
  return base + (uint32_t)__popcll(b & ((1ull << lane_id()) - 1ull));
}

// ---------------------------------------------------------- directory

__device__ __forceinline__ uint32_t root_hash(uint64_t key, uint32_t mask) {
  return (uint32_t)mix64(key) & mask;
}

// Find-or-insert of a root key (lock-free CAS insert, linear probing).
// Returns the slot, or -1 (and sets kStRootFull) after max_probe probes.
__device__ int32_t root_insert(const MapDev& m, uint64_t key) {
  uint32_t h = root_hash(key, m.root_mask);
  for (uint32_t p = 0; p < m.max_probe; ++p) {
    uint64_t k = ld64(&m.root_keys[h]);
    if (k == kEmptyKey) {
      k = cas64(&m.root_keys[h], kEmptyKey, key);
      if (k == kEmptyKey) {
        add32(&m.ctr->root_count, 1u);
        return (int32_t)h;
      }
    }
    if (k == key) return (int32_t)h;
    h = (h + 1) & m.root_mask;
  }
  or32(&m.ctr->status, kStRootFull);
  return -1;
}

// Lookup of a root key in a phase where no insert runs concurrently.
__device__ __forceinline__ int32_t root_find(const MapDev& m, uint64_t key) {
  uint32_t h = root_hash(key, m.root_mask);
  for (uint32_t p = 0; p <= m.root_mask; ++p) {
    const uint64_t k = m.root_keys[h];
    if (k == key) return (int32_t)h;
    if (k == kEmptyKey) return -1;
    h = (h + 1) & m.root_mask;
  }
  return -1;
}

// Make sure leaf `ci` of root `slot` exists.  The lane that moves the child
// word from EMPTY to PENDING allocates the leaf and publishes its index; all
// others return at once (they do not need the index before the next kernel).
// Returns the child word as seen (a leaf index, PENDING, or EMPTY on failure).
__device__ uint32_t leaf_request(const MapDev& m, uint32_t slot, uint32_t ci, int32_t lx,
                                 int32_t ly, int32_t lz) {
  uint32_t* cp = &m.children[(uint64_t)slot * m.inner_size + ci];
  uint32_t v = ld32(cp);
  if (v != kEmptyChild) return v;
  v = cas32(cp, kEmptyChild, kPendingChild);
  if (v != kEmptyChild) return v;
  const uint32_t idx = add32(&m.ctr->leaf_count, 1u);
  if (idx >= m.leaf_cap) {
    st32(cp, kEmptyChild);
    or32(&m.ctr->status, kStOutOfPool);
    return kEmptyChild;
  }
  m.leaf_origin[idx] = make_int4(lx, ly, lz, (int32_t)slot);
  st32(cp, idx);
  return idx;
}

// leaf index holding coordinate c, or kNone (read-only phases)
__device__ __forceinline__ uint32_t find_leaf(const MapDev& m, int32_t x, int32_t y, int32_t z) {
  uint64_t key;
  if (!pack_root(x, y, z, m.log2n, &key)) return kNone;
  const int32_t s = root_find(m, key);
  if (s < 0) return kNone;
  const uint32_t c = m.children[(uint64_t)s * m.inner_size + inner_index(x, y, z, m.leaf_bits, m.inner_bits)];
  return c >= kPendingChild ? kNone : c;
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

// Integer 3-D Bresenham walk of RayIterator (probabilistic_map.hpp:162-203):
// cell 0 is the origin voxel, then max|d|-1 further cells; the end voxel is
// not part of the ray and a zero-length ray is empty.
struct RayWalk {
  int32_t cx, cy, cz;
  int32_t sx, sy, sz;
  int32_t dx, dy, dz;
  int32_t ex, ey, ez;
  int32_t mx;
  __device__ __forceinline__ void init(int32_t ox, int32_t oy, int32_t oz, int32_t tx, int32_t ty,
                                       int32_t tz) {
    cx = ox;
    cy = oy;
    cz = oz;
    dx = tx - ox;
    dy = ty - oy;
    dz = tz - oz;
    sx = dx < 0 ? -1 : 1;
    sy = dy < 0 ? -1 : 1;
    sz = dz < 0 ? -1 : 1;
    dx = dx < 0 ? -dx : dx;
    dy = dy < 0 ? -dy : dy;
    dz = dz < 0 ? -dz : dz;
    mx = dx > dy ? dx : dy;
    mx = mx > dz ? mx : dz;
    ex = ey = ez = 0;
  }
  __device__ __forceinline__ void step() {
    ex += dx;
    ey += dy;
    ez += dz;
    if (2 * ex >= mx) {
      cx += sx;
      ex -= mx;
    }
    if (2 * ey >= mx) {
      cy += sy;
      ey -= mx;
    }
    if (2 * ez >= mx) {
      cz += sz;
      ez -= mx;
    }
  }
};

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
  if (gi >= hdr->total) return;
  const uint32_t s = find_scan(sd, hdr->nscans, gi);
  const ScanDesc& d = sd[s];
  const MapDev& m = *d.map;
  const uint64_t i = gi - d.pt_off;
  double tx, ty, tz;
  if (d.flags & kScanF64) {
    const double* p = (const double*)d.points + 3 * i;
    tx = p[0];
    ty = p[1];
    tz = p[2];
  } else {
    const float* p = (const float*)d.points + 3 * i;
    tx = (double)p[0];
    ty = (double)p[1];
    tz = (double)p[2];
  }
  const double vx = tx - d.fx, vy = ty - d.fy, vz = tz - d.fz;
  const double sq = (vx * vx + vy * vy) + vz * vz;
  double qx = tx, qy = ty, qz = tz;
  uint32_t flags = kRecValid;
  if (d.flags & (kScanForceHit | kScanForceMiss)) {
    if (d.flags & kScanForceHit) flags |= kRecHit;
  } else if (sq >= d.r2) {
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
  if (!(ax >= -lim && ax < lim && ay >= -lim && ay < lim && az >= -lim && az < lim)) {
    or32(&m.ctr->status, kStCoordRange);
    sc.rec[gi] = make_int4(0, 0, 0, (int32_t)(s << 16));
    return;
  }
  const int32_t ex = (int32_t)ax, ey = (int32_t)ay, ez = (int32_t)az;
  uint64_t rkey;
  pack_root(ex, ey, ez, m.log2n, &rkey);
  const int32_t slot = root_insert(m, rkey);
  if (slot < 0) {
    sc.rec[gi] = make_int4(ex, ey, ez, (int32_t)(s << 16));
    return;
  }
  const uint32_t ci = inner_index(ex, ey, ez, m.leaf_bits, m.inner_bits);
  const uint32_t li = leaf_index(ex, ey, ez, m.leaf_bits);
  const int32_t lmask = ~((1 << m.leaf_bits) - 1);
  const uint32_t leaf = leaf_request(m, (uint32_t)slot, ci, ex & lmask, ey & lmask, ez & lmask);
  bool marked = true;  // a leaf created in this scan has all cells zero: id 0
  if (leaf < kPendingChild) {
    const uint32_t w = ld32((const uint32_t*)m.cells + (uint64_t)leaf * m.leaf_size + li);
    marked = cell_id(w) != m.ctr->cur;
  }
  if (marked) {
    const uint64_t key = ep_key(d.tag, (uint32_t)slot, ci, li);
    uint32_t h = (uint32_t)mix64(key) & sc.ep_mask;
    uint32_t tp = kNone;
    for (uint32_t p = 0; p <= sc.ep_mask; ++p) {
      uint64_t k = ld64(&sc.ep_keys[h]);
      if (k == kEmptyKey) {
        k = cas64(&sc.ep_keys[h], kEmptyKey, key);
        if (k == kEmptyKey) k = key;
      }
      if (k == key) {
        min32(&sc.ep_first[h], (uint32_t)gi);
        tp = h;
        break;
      }
      h = (h + 1) & sc.ep_mask;
    }
    if (tp == kNone) {
      or32(&m.ctr->status, kStScratchFull);
    } else {
      flags |= kRecMarked;
      sc.tpos[gi] = tp;
    }
  }
  sc.rec[gi] = make_int4(ex, ey, ez, (int32_t)(flags | (s << 16)));
}

// every leaf on Ray(o, e) created (no cell touched)
__device__ void ray_leaves(const MapDev& m, int32_t ox, int32_t oy, int32_t oz, int32_t tx,
                           int32_t ty, int32_t tz) {
  RayWalk w;
  w.init(ox, oy, oz, tx, ty, tz);
  const int32_t lmask = ~((1 << m.leaf_bits) - 1);
  const int32_t rmask = ~((1 << m.log2n) - 1);
  int32_t plx = w.cx ^ 1, ply = 0, plz = 0;  // no previous leaf
  int32_t prx = w.cx ^ 1, pry = 0, prz = 0;
  int32_t slot = -1;
  for (int32_t k = 0; k < w.mx; ++k) {
    if (k) w.step();
    const int32_t lx = w.cx & lmask, ly = w.cy & lmask, lz = w.cz & lmask;
    if (lx != plx || ly != ply || lz != plz) {
      plx = lx;
      ply = ly;
      plz = lz;
      const int32_t rx = w.cx & rmask, ry = w.cy & rmask, rz = w.cz & rmask;
      if (rx != prx || ry != pry || rz != prz) {
        prx = rx;
        pry = ry;
        prz = rz;
        uint64_t key;
        pack_root(w.cx, w.cy, w.cz, m.log2n, &key);
        slot = root_insert(m, key);
      }
      if (slot < 0) return;
      leaf_request(m, (uint32_t)slot, inner_index(w.cx, w.cy, w.cz, m.leaf_bits, m.inner_bits), lx,
                   ly, lz);
    }
  }
}

// The miss update of every cell of Ray(o, e) whose update_id differs from
// cur; returns the number of cells written when counting (exact CAS form).
template <bool kCount>
__device__ uint32_t ray_apply(const MapDev& m, int32_t ox, int32_t oy, int32_t oz, int32_t tx,
                              int32_t ty, int32_t tz, int32_t cur) {
  const int32_t miss = m.opt_miss, cmin = m.opt_cmin;
  uint32_t* cells = (uint32_t*)m.cells;
  const int32_t lmask = ~((1 << m.leaf_bits) - 1);
  RayWalk w;
  w.init(ox, oy, oz, tx, ty, tz);
  int32_t plx = w.cx ^ 1, ply = 0, plz = 0;
  uint32_t leaf = 0;
  uint32_t nupd = 0;
  for (int32_t k = 0; k < w.mx; ++k) {
    if (k) w.step();
    const int32_t lx = w.cx & lmask, ly = w.cy & lmask, lz = w.cz & lmask;
    if (lx != plx || ly != ply || lz != plz) {
      plx = lx;
      ply = ly;
      plz = lz;
      leaf = find_leaf(m, w.cx, w.cy, w.cz);
    }
    if (leaf == kNone) continue;
    const uint32_t li = leaf_index(w.cx, w.cy, w.cz, m.leaf_bits);
    uint32_t* cp = cells + (uint64_t)leaf * m.leaf_size + li;
    uint32_t v = *cp;
    if (cell_id(v) != cur) {
      if (kCount) {
        for (;;) {
          const uint32_t old = cas32(cp, v, cell_miss(v, miss, cmin, cur));
          if (old == v) {
            ++nupd;
            break;
          }
          v = old;
          if (cell_id(v) == cur) break;
        }
      } else {
        *cp = cell_miss(v, miss, cmin, cur);
      }
      if (v == 0) or64(&m.leaf_masks[(uint64_t)leaf * m.leaf_words + (li >> 6)], 1ull << (li & 63));
    }
  }
  return nupd;
}

// Phase 2 (per point): the lowest-index point of each claimed voxel is its
// winner (probabilistic_map.cpp:30-54 lets only the first point of a voxel
// through the update_id test); winners are listed, and every leaf on the
// winner's ray is created before any cell is written.
__global__ void __launch_bounds__(256) k_scan_winners(const BatchHdr* __restrict__ hdr,
                                                      const ScanDesc* __restrict__ sd,
                                                      ScanScratch sc) {
  const uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool win = false;
  int4 r = make_int4(0, 0, 0, 0);
  if (gi < hdr->total) {
    r = sc.rec[gi];
    if ((r.w & kRecMarked) && sc.ep_first[sc.tpos[gi]] == (uint32_t)gi) win = true;
  }
  const uint32_t pos = wave_append(&sc.counters[0], win);
  if (!win) return;
  sc.winners[pos] = (uint32_t)gi;
  const ScanDesc& d = sd[(uint32_t)r.w >> 16];
  if (d.flags & kScanNoRays) return;
  if (r.x == d.ox && r.y == d.oy && r.z == d.oz) return;
  ray_leaves(*d.map, d.ox, d.oy, d.oz, r.x, r.y, r.z);
}

// Phase 2b (one block per scan): leaves of the rays still due from
// addHitPoint / addMissPoint calls made since the map's last insert.
__global__ void __launch_bounds__(256) k_scan_extra_leaves(const BatchHdr* __restrict__ hdr,
                                                           const ScanDesc* __restrict__ sd) {
  if (blockIdx.x >= hdr->nscans) return;
  const ScanDesc& d = sd[blockIdx.x];
  if (d.flags & kScanNoRays) return;
  for (uint32_t j = threadIdx.x; j < d.nextra; j += blockDim.x) {
    const int4 e = d.extra[j];
    if (e.x == d.ox && e.y == d.oy && e.z == d.oz) continue;
    ray_leaves(*d.map, d.ox, d.oy, d.oz, e.x, e.y, e.z);
  }
}

constexpr uint32_t kFatal = kStOutOfPool | kStRootFull | kStCoordRange | kStScratchFull;

// Phase 3 (per winner): endpoint hit/miss update (probabilistic_map.cpp:30-54)
// and release of the endpoint-table slot.  One writer per endpoint cell.
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
    const MapDev* mp = nullptr;
    if (idx < nwin) {
      const uint32_t gi = sc.winners[idx];
      const int4 r = sc.rec[gi];
      const uint32_t t = sc.tpos[gi];
      sc.ep_keys[t] = kEmptyKey;
      sc.ep_first[t] = kNone;
      const ScanDesc& d = sd[(uint32_t)r.w >> 16];
      const MapDev& m = *d.map;
      const uint32_t leaf = (m.ctr->status & kFatal) ? kNone : find_leaf(m, r.x, r.y, r.z);
      if (leaf != kNone) {
        const uint32_t li = leaf_index(r.x, r.y, r.z, m.leaf_bits);
        uint32_t* cell = (uint32_t*)m.cells + (uint64_t)leaf * m.leaf_size + li;
        const uint32_t w = *cell;
        const int32_t cur = m.ctr->cur;
        *cell = (r.w & kRecHit) ? cell_hit(w, m.opt_hit, m.opt_cmax, cur)
                                : cell_miss(w, m.opt_miss, m.opt_cmin, cur);
        if (w == 0)
          or64(&m.leaf_masks[(uint64_t)leaf * m.leaf_words + (li >> 6)], 1ull << (li & 63));
        updated = 1;
        mp = &m;
      }
    }
    // one counter update per distinct map among the wavefront's lanes
    uint64_t pending = __ballot(updated);
    while (pending) {
      const int leader = __ffsll((unsigned long long)pending) - 1;
      const MapDev* lm = (const MapDev*)__shfl((unsigned long long)mp, leader);
      const uint64_t same = __ballot(updated && mp == lm);
      if ((int)lane_id() == leader) add64(&lm->ctr->stat_endpoint_updates, (uint64_t)__popcll(same));
      pending &= ~same;
    }
  }
}

// Phase 4 (per winner): the ray clear of updateFreeCells
// (probabilistic_map.cpp:77-101): every cell of Ray(origin, endpoint) takes
// the miss update unless its update_id already equals the scan's id.
template <bool kCount>
__global__ void __launch_bounds__(256) k_scan_rays(const ScanDesc* __restrict__ sd,
                                                   ScanScratch sc) {
  const uint32_t nwin = sc.counters[0];
  for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < nwin;
       idx += gridDim.x * blockDim.x) {
    const uint32_t gi = sc.winners[idx];
    const int4 r = sc.rec[gi];
    const ScanDesc& d = sd[(uint32_t)r.w >> 16];
    const MapDev& m = *d.map;
    if (d.flags & kScanNoRays) continue;
    if (m.ctr->status & kFatal) continue;
    if (r.x == d.ox && r.y == d.oy && r.z == d.oz) continue;
    const uint32_t n = ray_apply<kCount>(m, d.ox, d.oy, d.oz, r.x, r.y, r.z, m.ctr->cur);
    if (kCount) {
      add64(&m.ctr->stat_ray_updates, n);
      add64(&m.ctr->stat_rays, 1);
    }
  }
}

// Phase 4b (one block per scan): rays due from addHitPoint / addMissPoint.
template <bool kCount>
__global__ void __launch_bounds__(256) k_scan_extra_rays(const BatchHdr* __restrict__ hdr,
                                                         const ScanDesc* __restrict__ sd) {
  if (blockIdx.x >= hdr->nscans) return;
  const ScanDesc& d = sd[blockIdx.x];
  if (d.flags & kScanNoRays) return;
  const MapDev& m = *d.map;
  if (m.ctr->status & kFatal) return;
  for (uint32_t j = threadIdx.x; j < d.nextra; j += blockDim.x) {
    const int4 e = d.extra[j];
    if (e.x == d.ox && e.y == d.oy && e.z == d.oz) continue;
    const uint32_t n = ray_apply<kCount>(m, d.ox, d.oy, d.oz, e.x, e.y, e.z, m.ctr->cur);
    if (kCount) {
      add64(&m.ctr->stat_ray_updates, n);
      add64(&m.ctr->stat_rays, 1);
    }
  }
}

// Phase 5 (per scan): advance update_count 1 -> 2 -> 3 -> 1
// (probabilistic_map.cpp:103-105) unless the scan failed; reset the batch.
__global__ void k_scan_finalize(const BatchHdr* __restrict__ hdr, const ScanDesc* __restrict__ sd,
                                ScanScratch sc) {
  for (uint32_t s = threadIdx.x; s < hdr->nscans; s += blockDim.x) {
    if (sd[s].flags & kScanNoRays) continue;
    const MapDev& m = *sd[s].map;
    if (m.ctr->status & kFatal) continue;
    const int32_t c = m.ctr->cur + 1;
    m.ctr->cur = c == 4 ? 1 : c;
    m.ctr->scans_applied += 1;
  }
  if (threadIdx.x == 0) sc.counters[0] = 0;
}

// ------------------------------------------------------ bulk cell ops

struct OpScratch {
  uint64_t* keys;   // table keys (global cell index)
  uint32_t* first;  // min input index per key
  uint32_t* last;   // max input index per key
  uint32_t mask;
  uint32_t* tpos;    // per input: table slot (kNone if no cell)
  uint8_t* prebit;   // per input: mask bit before the call
};

// create the root and leaf of every coordinate (setValue / setCellOn path)
__global__ void __launch_bounds__(256) k_touch(const MapDev* __restrict__ mp,
                                               const int32_t* __restrict__ coords, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MapDev& m = *mp;
  const int32_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  uint64_t key;
  if (!pack_root(x, y, z, m.log2n, &key)) {
    or32(&m.ctr->status, kStCoordRange);
    return;
  }
  const int32_t slot = root_insert(m, key);
  if (slot < 0) return;
  const int32_t lmask = ~((1 << m.leaf_bits) - 1);
  leaf_request(m, (uint32_t)slot, inner_index(x, y, z, m.leaf_bits, m.inner_bits), x & lmask,
               y & lmask, z & lmask);
}

// first and last input index of every distinct cell, and its prior mask bit
__global__ void __launch_bounds__(256) k_op_resolve(const MapDev* __restrict__ mp,
                                                    const int32_t* __restrict__ coords,
                                                    uint64_t n, OpScratch os) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MapDev& m = *mp;
  const int32_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  const uint32_t leaf = find_leaf(m, x, y, z);
  if (leaf == kNone) {
    os.tpos[i] = kNone;
    os.prebit[i] = 0;
    return;
  }
  const uint32_t li = leaf_index(x, y, z, m.leaf_bits);
  const uint64_t key = (uint64_t)leaf * m.leaf_size + li;
  os.prebit[i] = (uint8_t)((m.leaf_masks[(uint64_t)leaf * m.leaf_words + (li >> 6)] >> (li & 63)) & 1);
  uint32_t h = (uint32_t)mix64(key) & os.mask;
  for (uint32_t p = 0; p <= os.mask; ++p) {
    uint64_t k = ld64(&os.keys[h]);
    if (k == kEmptyKey) {
      k = cas64(&os.keys[h], kEmptyKey, key);
      if (k == kEmptyKey) k = key;
    }
    if (k == key) {
      min32(&os.first[h], (uint32_t)i);
      max32(&os.last[h], (uint32_t)i);
      os.tpos[i] = h;
      return;
    }
    h = (h + 1) & os.mask;
  }
  os.tpos[i] = kNone;
}

enum : uint32_t { kOpSetValue = 0, kOpSetOn = 1, kOpSetOff = 2 };

__device__ __forceinline__ void copy_value(uint8_t* dst, const uint8_t* src, uint32_t vb) {
  if (vb == 4) {
    *(uint32_t*)dst = *(const uint32_t*)src;
  } else if (vb == 8) {
    *(uint64_t*)dst = *(const uint64_t*)src;
  } else {
    for (uint32_t b = 0; b < vb; ++b) dst[b] = src[b];
  }
}
__device__ __forceinline__ void copy_value_unaligned(uint8_t* dst, const uint8_t* src, uint32_t vb) {
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
  const MapDev& m = *mp;
  const uint32_t t = os.tpos[i];
  if (t == kNone) {
    if (was_on) was_on[i] = 0;
    return;
  }
  if (mp->ctr->status & kFatal) {  // the call failed: leave every cell as it was
    if (was_on) was_on[i] = 0;
    return;
  }
  const uint32_t first = os.first[t], last = os.last[t];
  const uint8_t pre = os.prebit[i];
  const int32_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  const uint32_t leaf = find_leaf(m, x, y, z);
  if (leaf == kNone) {
    if (was_on) was_on[i] = 0;
    return;
  }
  const uint32_t li = leaf_index(x, y, z, m.leaf_bits);
  const uint64_t g = (uint64_t)leaf * m.leaf_size + li;
  uint64_t* mw = &m.leaf_masks[(uint64_t)leaf * m.leaf_words + (li >> 6)];
  const uint64_t bit = 1ull << (li & 63);
  const uint32_t vb = m.value_bytes;
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
      if (vb) copy_value_unaligned(m.cells + g * vb, vals + i * vb, vb);
      or64(mw, bit);
    }
  } else {  // setCellOn: the first occurrence decides, and writes only if off
    if ((uint32_t)i == first) {
      if (!pre && vb && vals) copy_value_unaligned(m.cells + g * vb, vals + i * vb, vb);
      or64(mw, bit);
    }
  }
}

// return the table slots of the call to the empty state
__global__ void __launch_bounds__(256) k_op_release(uint64_t n, OpScratch os) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t t = os.tpos[i];
  if (t == kNone) return;
  if (os.last[t] == (uint32_t)i) {
    os.keys[t] = kEmptyKey;
    os.first[t] = kNone;
    os.last[t] = 0;
  }
}

// ConstAccessor::value / isCellOn (bonxai.hpp:496-534) for a batch
__global__ void __launch_bounds__(256) k_lookup(const MapDev* __restrict__ mp,
                                                const int32_t* __restrict__ coords, uint64_t n,
                                                uint8_t* __restrict__ vals,
                                                uint8_t* __restrict__ found) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const MapDev& m = *mp;
  const int32_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  const uint32_t leaf = find_leaf(m, x, y, z);
  const uint32_t vb = m.value_bytes;
  bool on = false;
  uint64_t g = 0;
  if (leaf != kNone) {
    const uint32_t li = leaf_index(x, y, z, m.leaf_bits);
    g = (uint64_t)leaf * m.leaf_size + li;
    on = (m.leaf_masks[(uint64_t)leaf * m.leaf_words + (li >> 6)] >> (li & 63)) & 1;
  }
  if (found) found[i] = on;
  if (vals && vb) {
    if (vb == 4) {
      ((uint32_t*)vals)[i] = on ? ((const uint32_t*)m.cells)[g] : 0u;
    } else if (vb == 8) {
      ((uint64_t*)vals)[i] = on ? ((const uint64_t*)m.cells)[g] : 0ull;
    } else {
      for (uint32_t b = 0; b < vb; ++b) vals[i * vb + b] = on ? m.cells[g * vb + b] : 0;
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
  const MapDev& m = *mp;
  const int32_t x = coords[3 * i], y = coords[3 * i + 1], z = coords[3 * i + 2];
  const uint32_t leaf = find_leaf(m, x, y, z);
  int8_t st = 0;
  if (leaf != kNone) {
    const uint32_t li = leaf_index(x, y, z, m.leaf_bits);
    if ((m.leaf_masks[(uint64_t)leaf * m.leaf_words + (li >> 6)] >> (li & 63)) & 1) {
      const int32_t p = cell_p(((const uint32_t*)m.cells)[(uint64_t)leaf * m.leaf_size + li]);
      st = p > m.opt_thr ? 1 : (p < m.opt_thr ? -1 : 0);
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
                                                uint32_t* __restrict__ count) {
  const MapDev& m = *mp;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t leaf = t >> (3 * m.leaf_bits);
  const uint32_t li = (uint32_t)(t & (m.leaf_size - 1));
  bool pass = false;
  if (leaf < nleaves) {
    pass = (m.leaf_masks[leaf * m.leaf_words + (li >> 6)] >> (li & 63)) & 1;
    if (pass && filter) {
      const int32_t p = cell_p(((const uint32_t*)m.cells)[leaf * m.leaf_size + li]);
      pass = filter == 1 ? p > m.opt_thr : p < m.opt_thr;
    }
  }
  const uint32_t pos = wave_append(count, pass);
  if (!pass || pos >= cap) return;
  const int4 o = m.leaf_origin[leaf];
  const uint32_t lb = m.leaf_bits, lm = (1u << lb) - 1;
  if (coords) {
    coords[3 * (uint64_t)pos] = o.x | (int32_t)(li & lm);
    coords[3 * (uint64_t)pos + 1] = o.y | (int32_t)((li >> lb) & lm);
    coords[3 * (uint64_t)pos + 2] = o.z | (int32_t)((li >> (2 * lb)) & lm);
  }
  const uint32_t vb = m.value_bytes;
  if (vals && vb) copy_value_unaligned(vals + (uint64_t)pos * vb, m.cells + (leaf * m.leaf_size + li) * vb, vb);
}

// order-free digest of the on cells (count, sum, xor, sum of remix)
__global__ void __launch_bounds__(256) k_digest(const MapDev* __restrict__ mp, uint32_t nleaves,
                                                uint64_t* __restrict__ out) {
  const MapDev& m = *mp;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t leaf = t >> (3 * m.leaf_bits);
  const uint32_t li = (uint32_t)(t & (m.leaf_size - 1));
  uint64_t h = 0, h2 = 0, c = 0;
  if (leaf < nleaves && ((m.leaf_masks[leaf * m.leaf_words + (li >> 6)] >> (li & 63)) & 1)) {
    const int4 o = m.leaf_origin[leaf];
    const uint32_t lb = m.leaf_bits, lm = (1u << lb) - 1, vb = m.value_bytes;
    uint64_t word = 0;
    const uint8_t* src = m.cells + (leaf * m.leaf_size + li) * vb;
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
  const MapDev& m = *mp;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t c = 0;
  if (t < (uint64_t)nleaves * m.leaf_words) {
    const uint64_t w = m.leaf_masks[t];
    c = m.leaf_size < 64 ? (uint64_t)__popcll(w & ((1ull << m.leaf_size) - 1)) : (uint64_t)__popcll(w);
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if (lane_id() == 0 && c) add64(out, c);
}

// clear(SET_ALL_CELLS_OFF): every mask word and every cell byte to zero
__global__ void __launch_bounds__(256) k_cells_off(const MapDev* __restrict__ mp, uint32_t nleaves) {
  const MapDev& m = *mp;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t nw = (uint64_t)nleaves * m.leaf_words;
  for (uint64_t w = t; w < nw; w += (uint64_t)gridDim.x * blockDim.x) m.leaf_masks[w] = 0;
  const uint64_t nb = (uint64_t)nleaves * m.leaf_size * m.value_bytes / 4;
  for (uint64_t b = t; b < nb; b += (uint64_t)gridDim.x * blockDim.x) ((uint32_t*)m.cells)[b] = 0;
}

// re-insert every root of an old directory into the (empty) directory of m
__global__ void __launch_bounds__(256) k_rehash(const MapDev* __restrict__ mp,
                                                const uint64_t* __restrict__ old_keys,
                                                const uint32_t* __restrict__ old_children,
                                                uint32_t old_cap) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= old_cap) return;
  const uint64_t key = old_keys[s];
  if (key == kEmptyKey) return;
  const MapDev& m = *mp;
  const int32_t ns = root_insert(m, key);
  if (ns < 0) return;
  for (uint32_t c = 0; c < m.inner_size; ++c) {
    const uint32_t leaf = old_children[(uint64_t)s * m.inner_size + c];
    m.children[(uint64_t)ns * m.inner_size + c] = leaf;
    if (leaf < kPendingChild) m.leaf_origin[leaf].w = ns;
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
  const MapDev& m = *mp;
  const uint32_t k = blockIdx.x;
  if (k >= n) return;
  const uint32_t leaf = find_leaf(m, origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]);
  if (threadIdx.x == 0) exists[k] = leaf != kNone;
  for (uint32_t w = threadIdx.x; w < m.leaf_words; w += blockDim.x)
    masks[(uint64_t)k * m.leaf_words + w] = leaf != kNone ? m.leaf_masks[(uint64_t)leaf * m.leaf_words + w] : 0;
  const uint64_t nb = (uint64_t)m.leaf_size * m.value_bytes;
  for (uint64_t b = threadIdx.x; b < nb; b += blockDim.x)
    cells[k * nb + b] = leaf != kNone ? m.cells[leaf * nb + b] : 0;
}

// Overwrite whole leaves (they must exist: run k_touch on the origins first).
// Cells whose bit is off are written as zero bytes (the off-cell invariant).
__global__ void __launch_bounds__(256) k_write_leaves(const MapDev* __restrict__ mp,
                                                      const int32_t* __restrict__ origins,
                                                      uint32_t n, const uint64_t* __restrict__ masks,
                                                      const uint8_t* __restrict__ cells) {
  const MapDev& m = *mp;
  const uint32_t k = blockIdx.x;
  if (k >= n) return;
  const uint32_t leaf = find_leaf(m, origins[3 * k], origins[3 * k + 1], origins[3 * k + 2]);
  if (leaf == kNone) return;
  for (uint32_t w = threadIdx.x; w < m.leaf_words; w += blockDim.x)
    m.leaf_masks[(uint64_t)leaf * m.leaf_words + w] = masks[(uint64_t)k * m.leaf_words + w];
  const uint32_t vb = m.value_bytes;
  const uint64_t nb = (uint64_t)m.leaf_size * vb;
  for (uint64_t b = threadIdx.x; b < nb; b += blockDim.x) {
    const uint32_t li = (uint32_t)(b / vb);
    const bool on = (masks[(uint64_t)k * m.leaf_words + (li >> 6)] >> (li & 63)) & 1;
    m.cells[leaf * nb + b] = on ? cells[k * nb + b] : 0;
  }
}

// origins of all allocated leaves (for paging a whole map to the host)
__global__ void __launch_bounds__(256) k_leaf_origins(const MapDev* __restrict__ mp, uint32_t nleaves,
                                                      int32_t* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nleaves) return;
  const int4 o = mp->leaf_origin[k];
  out[3 * k] = o.x;
  out[3 * k + 1] = o.y;
  out[3 * k + 2] = o.z;
}

// ------------------------------------------------ arithmetic self-check

// The FP64 operations of the scan front end, element-wise, for the
// correctly-rounded checks of sqrt, division, multiply-add order and floor.
__global__ void k_fp64_probe(const double* __restrict__ a, const double* __restrict__ b,
                             const double* __restrict__ c, uint64_t n, double* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[5 * i + 0] = __builtin_sqrt(a[i]);
  out[5 * i + 1] = a[i] / b[i];
  out[5 * i + 2] = c[i] + ((a[i] / b[i]) * c[i]);
  out[5 * i + 3] = __builtin_floor(a[i] * b[i]);
  out[5 * i + 4] = (a[i] * a[i] + b[i] * b[i]) + c[i] * c[i];
}

}  // namespace bxg
