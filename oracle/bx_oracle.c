/*
 * bx_oracle.c -- CPU restatement of the reference's sparse voxel grid and of
 * ProbabilisticMap scan integration.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  This file is the parity checker for
 * the HIP library and the timed single-core CPU baseline ("port").  Only the
 * test suite, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it;
 * nothing under bonxai_amd/ links, loads or calls it, and the GPU library has
 * no path that falls back to it.
 *
 * It is written from the reference's documented behaviour, one sequential
 * thread, in the same order of operations as the reference (reference paths
 * are relative to the reference source tree):
 *
 *   posToCoord              bonxai_core/include/bonxai/bonxai.hpp:404-410
 *   coordToPos              bonxai.hpp:412-417
 *   root/inner/leaf keys    bonxai.hpp:419-447
 *   Accessor::setValue      bonxai.hpp:449-466
 *   Accessor::value         bonxai.hpp:469-494
 *   ConstAccessor::value    bonxai.hpp:496-516, isCellOn 518-534
 *   setCellOn / setCellOff  bonxai.hpp:537-569
 *   getLeafGrid             bonxai.hpp:588-647
 *   activeCellsCount        bonxai.hpp:689-701
 *   forEachCell             bonxai.hpp:704-743
 *   clear                   bonxai.hpp:678-687
 *   releaseUnusedMemory     bonxai.hpp:367-387
 *   Mask bit n -> word n>>6, bit n&63   bonxai_core/include/bonxai/mask.hpp:329-365
 *   CellT {update_id:4, probability_log:28}
 *                           bonxai_map/include/bonxai_map/probabilistic_map.hpp:44-53
 *   Options defaults        probabilistic_map.hpp:56-64 (logods :34-36)
 *   insertPointCloud        probabilistic_map.hpp:141-160
 *   RayIterator             probabilistic_map.hpp:162-203
 *   addHitPoint/addMissPoint bonxai_map/src/probabilistic_map.cpp:30-54
 *   isOccupied/isUnknown/isFree  probabilistic_map.cpp:56-75
 *   updateFreeCells         probabilistic_map.cpp:77-106
 *   getOccupied/FreeVoxels  probabilistic_map.cpp:108-126
 *   ingest pre-step         bonxai_ros/src/bonxai_server.cpp:147-179 (NaN filter,
 *                           float transform, origin) -- parity unpinned, see
 *                           bxo_prob_insert_tf
 *
 * Eigen3 (absent from this image) supplies the arithmetic of
 * probabilistic_map.hpp:144-157.  Its squaredNorm() of a Vector3d is taken as
 * (x*x + y*y) + z*z, the SSE2 packet reduction of Eigen 3.4 with default x86-64
 * flags; this association is parity unpinned (no Eigen source to check), and
 * the test suite counts the points whose endpoint would change under the other
 * association x*x + (y*y + z*z).  Build with -ffp-contract=off: the reference
 * is built without FMA (default x86-64 target, RelWithDebInfo).
 *
 * Two departures that cannot change any observable cell state:
 *   - the root directory is an open-addressing table, not std::unordered_map
 *     (iteration order is never compared; exports are sorted);
 *   - Accessor::value(coord, true) re-resolves a cached null leaf instead of
 *     dereferencing it (the reference's behaviour there is undefined).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define BXO_EXPORT __attribute__((visibility("default")))

typedef struct {
  int32_t x, y, z;
} bxo_coord;

typedef struct bxo_leaf {
  uint64_t* mask;
  uint8_t* cells;
} bxo_leaf;

typedef struct bxo_inner {
  uint64_t* mask;
  bxo_leaf** child;
} bxo_inner;

typedef struct bxo_grid {
  uint32_t inner_bits, leaf_bits, log2n;
  uint32_t inner_size, leaf_size, inner_words, leaf_words;
  uint32_t value_bytes;
  double resolution, inv_resolution;
  /* root directory: open addressing on the root key */
  uint64_t cap, count;
  bxo_coord* keys;
  bxo_inner** inner;
  uint8_t* used;
  uint64_t leaf_count;
} bxo_grid;

typedef struct bxo_accessor {
  bxo_grid* g;
  bxo_coord prev_root;
  bxo_inner* prev_inner;
  bxo_coord prev_leaf_key;
  bxo_leaf* prev_leaf;
  int prev_valid;
} bxo_accessor;

/* ------------------------------------------------------------------ grid */

static uint64_t bxo_hash_coord(bxo_coord c) {
  uint64_t h = (uint64_t)(uint32_t)c.x * 0x9E3779B97F4A7C15ull;
  h ^= (uint64_t)(uint32_t)c.y * 0xC2B2AE3D27D4EB4Full;
  h ^= (uint64_t)(uint32_t)c.z * 0x165667B19E3779F9ull;
  h ^= h >> 29;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 32;
  return h;
}

static int bxo_coord_eq(bxo_coord a, bxo_coord b) {
  return a.x == b.x && a.y == b.y && a.z == b.z;
}

BXO_EXPORT bxo_grid* bxo_grid_create(double voxel_size, uint32_t inner_bits, uint32_t leaf_bits,
                                      uint32_t value_bytes) {
  if (inner_bits < 1 || leaf_bits < 1 || inner_bits > 4 || leaf_bits > 4) return NULL;
  bxo_grid* g = (bxo_grid*)calloc(1, sizeof(bxo_grid));
  g->inner_bits = inner_bits;
  g->leaf_bits = leaf_bits;
  g->log2n = inner_bits + leaf_bits;
  g->inner_size = 1u << (3 * inner_bits);
  g->leaf_size = 1u << (3 * leaf_bits);
  g->inner_words = g->inner_size >= 64 ? g->inner_size / 64 : 1;
  g->leaf_words = g->leaf_size >= 64 ? g->leaf_size / 64 : 1;
  g->value_bytes = value_bytes;
  g->resolution = voxel_size;
  g->inv_resolution = 1.0 / voxel_size; /* bonxai.hpp:395 */
  g->cap = 64;
  g->keys = (bxo_coord*)calloc(g->cap, sizeof(bxo_coord));
  g->inner = (bxo_inner**)calloc(g->cap, sizeof(bxo_inner*));
  g->used = (uint8_t*)calloc(g->cap, 1);
  return g;
}

static void bxo_free_leaf(bxo_leaf* l) {
  if (!l) return;
  free(l); /* header, mask and cells are one block */
}

static void bxo_free_inner(bxo_grid* g, bxo_inner* in) {
  for (uint32_t i = 0; i < g->inner_size; ++i) bxo_free_leaf(in->child[i]);
  free(in->child);
  free(in->mask);
  free(in);
}

static void bxo_grid_free_all(bxo_grid* g) {
  for (uint64_t s = 0; s < g->cap; ++s)
    if (g->used[s]) bxo_free_inner(g, g->inner[s]);
  free(g->keys);
  free(g->inner);
  free(g->used);
}

BXO_EXPORT void bxo_grid_destroy(bxo_grid* g) {
  if (!g) return;
  bxo_grid_free_all(g);
  free(g);
}

/* clear(CLEAR_MEMORY), bonxai.hpp:679-683 */
BXO_EXPORT void bxo_grid_clear_memory(bxo_grid* g) {
  bxo_grid_free_all(g);
  g->cap = 64;
  g->count = 0;
  g->leaf_count = 0;
  g->keys = (bxo_coord*)calloc(g->cap, sizeof(bxo_coord));
  g->inner = (bxo_inner**)calloc(g->cap, sizeof(bxo_inner*));
  g->used = (uint8_t*)calloc(g->cap, 1);
}

static bxo_inner** bxo_root_slot(bxo_grid* g, bxo_coord key, int create) {
  uint64_t m = g->cap - 1;
  uint64_t s = bxo_hash_coord(key) & m;
  while (g->used[s]) {
    if (bxo_coord_eq(g->keys[s], key)) return &g->inner[s];
    s = (s + 1) & m;
  }
  if (!create) return NULL;
  if ((g->count + 1) * 2 > g->cap) {
    /* grow and re-insert */
    uint64_t ocap = g->cap;
    bxo_coord* ok = g->keys;
    bxo_inner** oi = g->inner;
    uint8_t* ou = g->used;
    g->cap = ocap * 2;
    g->keys = (bxo_coord*)calloc(g->cap, sizeof(bxo_coord));
    g->inner = (bxo_inner**)calloc(g->cap, sizeof(bxo_inner*));
    g->used = (uint8_t*)calloc(g->cap, 1);
    m = g->cap - 1;
    for (uint64_t t = 0; t < ocap; ++t) {
      if (!ou[t]) continue;
      uint64_t u = bxo_hash_coord(ok[t]) & m;
      while (g->used[u]) u = (u + 1) & m;
      g->used[u] = 1;
      g->keys[u] = ok[t];
      g->inner[u] = oi[t];
    }
    free(ok);
    free(oi);
    free(ou);
    s = bxo_hash_coord(key) & m;
    while (g->used[s]) s = (s + 1) & m;
  }
  g->used[s] = 1;
  g->keys[s] = key;
  bxo_inner* in = (bxo_inner*)calloc(1, sizeof(bxo_inner));
  in->mask = (uint64_t*)calloc(g->inner_words, sizeof(uint64_t));
  in->child = (bxo_leaf**)calloc(g->inner_size, sizeof(bxo_leaf*));
  g->inner[s] = in;
  g->count++;
  return &g->inner[s];
}

/* getRootKey / getInnerKey / getInnerIndex / getLeafIndex, bonxai.hpp:419-447 */
static bxo_coord bxo_root_key(const bxo_grid* g, bxo_coord c) {
  const int32_t m = ~((1 << g->log2n) - 1);
  bxo_coord r = {c.x & m, c.y & m, c.z & m};
  return r;
}
static bxo_coord bxo_inner_key(const bxo_grid* g, bxo_coord c) {
  const int32_t m = ~((1 << g->leaf_bits) - 1);
  bxo_coord r = {c.x & m, c.y & m, c.z & m};
  return r;
}
static uint32_t bxo_inner_index(const bxo_grid* g, bxo_coord c) {
  const uint32_t lb = g->leaf_bits, ib = g->inner_bits, im = (1u << ib) - 1;
  return ((uint32_t)(c.x >> lb) & im) | (((uint32_t)(c.y >> lb) & im) << ib) |
         (((uint32_t)(c.z >> lb) & im) << (2 * ib));
}
static uint32_t bxo_leaf_index(const bxo_grid* g, bxo_coord c) {
  const uint32_t lb = g->leaf_bits, lm = (1u << lb) - 1;
  return ((uint32_t)c.x & lm) | (((uint32_t)c.y & lm) << lb) | (((uint32_t)c.z & lm) << (2 * lb));
}

static int bxo_bit(const uint64_t* w, uint32_t n) {
  return (w[n >> 6] >> (n & 63)) & 1u;
}
static int bxo_bit_set(uint64_t* w, uint32_t n) {
  const uint64_t b = 1ull << (n & 63);
  const int was = (w[n >> 6] & b) != 0;
  w[n >> 6] |= b;
  return was;
}
static int bxo_bit_clear(uint64_t* w, uint32_t n) {
  const uint64_t b = 1ull << (n & 63);
  const int was = (w[n >> 6] & b) != 0;
  w[n >> 6] &= ~b;
  return was;
}

/* getLeafGrid(coord, create), bonxai.hpp:588-621 */
static bxo_leaf* bxo_get_leaf(bxo_accessor* a, bxo_coord c, int create) {
  bxo_grid* g = a->g;
  bxo_coord rk = bxo_root_key(g, c);
  bxo_inner* in = a->prev_inner;
  if (!in || !bxo_coord_eq(rk, a->prev_root)) {
    bxo_inner** slot = bxo_root_slot(g, rk, create);
    if (!slot) return NULL;
    in = *slot;
    a->prev_root = rk;
    a->prev_inner = in;
  }
  const uint32_t ii = bxo_inner_index(g, c);
  if (create) {
    if (!bxo_bit_set(in->mask, ii)) {
      /* one block: the header, the mask words (inline in the leaf, as the
       * reference's Mask is, mask.hpp:23,258) and the cells */
      bxo_leaf* l = (bxo_leaf*)calloc(1, sizeof(bxo_leaf) + (size_t)g->leaf_words * sizeof(uint64_t) +
                                             (size_t)g->leaf_size * g->value_bytes);
      l->mask = (uint64_t*)(l + 1);
      l->cells = g->value_bytes ? (uint8_t*)(l->mask + g->leaf_words) : NULL;
      in->child[ii] = l;
      g->leaf_count++;
    }
  } else if (!bxo_bit(in->mask, ii)) {
    return NULL;
  }
  return in->child[ii];
}

static bxo_leaf* bxo_cached_leaf(bxo_accessor* a, bxo_coord c, int create) {
  bxo_coord ik = bxo_inner_key(a->g, c);
  if (!a->prev_valid || !bxo_coord_eq(ik, a->prev_leaf_key) || (create && !a->prev_leaf)) {
    a->prev_leaf = bxo_get_leaf(a, c, create);
    a->prev_leaf_key = ik;
    a->prev_valid = 1;
  }
  return a->prev_leaf;
}

static void bxo_accessor_init(bxo_accessor* a, bxo_grid* g) {
  memset(a, 0, sizeof(*a));
  a->g = g;
}

/* Accessor::value(coord, create), bonxai.hpp:469-494.  Returns a pointer to
 * the cell bytes or NULL. */
static uint8_t* bxo_value(bxo_accessor* a, bxo_coord c, int create) {
  bxo_leaf* l = bxo_cached_leaf(a, c, create);
  if (!l) return NULL;
  const uint32_t li = bxo_leaf_index(a->g, c);
  const uint32_t vb = a->g->value_bytes;
  if (bxo_bit(l->mask, li)) return l->cells + (size_t)li * vb;
  if (create) {
    bxo_bit_set(l->mask, li);
    memset(l->cells + (size_t)li * vb, 0, vb); /* value-initialised, bonxai.hpp:489 */
    return l->cells + (size_t)li * vb;
  }
  return NULL;
}

/* ---------------------------------------------------- bulk VoxelGrid ops */

/* Accessor::setValue in input order, bonxai.hpp:449-466 */
BXO_EXPORT void bxo_set_values(bxo_grid* g, const int32_t* coords, const void* vals, uint64_t n,
                               uint8_t* was_on) {
  bxo_accessor a;
  bxo_accessor_init(&a, g);
  const uint8_t* v = (const uint8_t*)vals;
  const uint32_t vb = g->value_bytes;
  for (uint64_t i = 0; i < n; ++i) {
    bxo_coord c = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
    bxo_leaf* l = bxo_cached_leaf(&a, c, 1);
    const uint32_t li = bxo_leaf_index(g, c);
    const int was = bxo_bit_set(l->mask, li);
    if (vb) memcpy(l->cells + (size_t)li * vb, v + i * vb, vb);
    if (was_on) was_on[i] = (uint8_t)was;
  }
}

/* Accessor::setCellOn(coord, default), bonxai.hpp:537-553 */
BXO_EXPORT void bxo_set_cells_on(bxo_grid* g, const int32_t* coords, const void* defaults,
                                 uint64_t n, uint8_t* was_on) {
  bxo_accessor a;
  bxo_accessor_init(&a, g);
  const uint8_t* v = (const uint8_t*)defaults;
  const uint32_t vb = g->value_bytes;
  for (uint64_t i = 0; i < n; ++i) {
    bxo_coord c = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
    bxo_leaf* l = bxo_cached_leaf(&a, c, 1);
    const uint32_t li = bxo_leaf_index(g, c);
    const int was = bxo_bit_set(l->mask, li);
    if (!was && vb) memcpy(l->cells + (size_t)li * vb, v + i * vb, vb);
    if (was_on) was_on[i] = (uint8_t)was;
  }
}

/* Accessor::setCellOff(coord), bonxai.hpp:556-569 */
BXO_EXPORT void bxo_set_cells_off(bxo_grid* g, const int32_t* coords, uint64_t n,
                                  uint8_t* was_on) {
  bxo_accessor a;
  bxo_accessor_init(&a, g);
  for (uint64_t i = 0; i < n; ++i) {
    bxo_coord c = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
    bxo_leaf* l = bxo_cached_leaf(&a, c, 0);
    int was = 0;
    if (l) was = bxo_bit_clear(l->mask, bxo_leaf_index(g, c));
    if (was_on) was_on[i] = (uint8_t)was;
  }
}

/* ConstAccessor::value / isCellOn, bonxai.hpp:496-534.  vals_out receives the
 * cell bytes for found cells and zero bytes otherwise. */
BXO_EXPORT void bxo_lookup(bxo_grid* g, const int32_t* coords, uint64_t n, void* vals_out,
                           uint8_t* found) {
  bxo_accessor a;
  bxo_accessor_init(&a, g);
  uint8_t* out = (uint8_t*)vals_out;
  const uint32_t vb = g->value_bytes;
  for (uint64_t i = 0; i < n; ++i) {
    bxo_coord c = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
    bxo_leaf* l = bxo_cached_leaf(&a, c, 0);
    int on = 0;
    if (l) {
      const uint32_t li = bxo_leaf_index(g, c);
      if (bxo_bit(l->mask, li)) {
        on = 1;
        if (vb && out) memcpy(out + i * vb, l->cells + (size_t)li * vb, vb);
      }
    }
    if (!on && vb && out) memset(out + i * vb, 0, vb);
    if (found) found[i] = (uint8_t)on;
  }
}

/* activeCellsCount, bonxai.hpp:689-701 */
BXO_EXPORT uint64_t bxo_active_cells(const bxo_grid* g) {
  uint64_t total = 0;
  for (uint64_t s = 0; s < g->cap; ++s) {
    if (!g->used[s]) continue;
    bxo_inner* in = g->inner[s];
    for (uint32_t i = 0; i < g->inner_size; ++i) {
      if (!bxo_bit(in->mask, i)) continue;
      bxo_leaf* l = in->child[i];
      for (uint32_t w = 0; w < g->leaf_words; ++w) total += (uint64_t)__builtin_popcountll(l->mask[w]);
    }
  }
  return total;
}

BXO_EXPORT uint64_t bxo_root_count(const bxo_grid* g) { return g->count; }

BXO_EXPORT uint64_t bxo_leaf_count(const bxo_grid* g) {
  uint64_t total = 0;
  for (uint64_t s = 0; s < g->cap; ++s) {
    if (!g->used[s]) continue;
    for (uint32_t i = 0; i < g->inner_size; ++i) total += (uint64_t)bxo_bit(g->inner[s]->mask, i);
  }
  return total;
}

/* forEachCell, bonxai.hpp:704-743: coord = root | inner<<LEAF_BITS | leaf.
 * Writes up to cap cells (coords xyz, raw value bytes) in directory order and
 * returns the number of on cells.  filter: 0 all, 1 occupied (p > thr),
 * 2 free (p < thr), the last two for 4-byte CellT grids only. */
BXO_EXPORT uint64_t bxo_export(const bxo_grid* g, int filter, int32_t threshold, int32_t* coords,
                               void* vals, uint64_t cap) {
  const int32_t lm = (1 << g->leaf_bits) - 1, im = (1 << g->inner_bits) - 1;
  const uint32_t lb = g->leaf_bits, ib = g->inner_bits, vb = g->value_bytes;
  uint8_t* out = (uint8_t*)vals;
  uint64_t k = 0;
  for (uint64_t s = 0; s < g->cap; ++s) {
    if (!g->used[s]) continue;
    const bxo_coord r = g->keys[s];
    const bxo_inner* in = g->inner[s];
    for (uint32_t ii = 0; ii < g->inner_size; ++ii) {
      if (!bxo_bit(in->mask, ii)) continue;
      const int32_t xb = r.x | ((int32_t)(ii & im) << lb);
      const int32_t yb = r.y | ((int32_t)((ii >> ib) & im) << lb);
      const int32_t zb = r.z | ((int32_t)((ii >> (2 * ib)) & im) << lb);
      const bxo_leaf* l = in->child[ii];
      for (uint32_t li = 0; li < g->leaf_size; ++li) {
        if (!bxo_bit(l->mask, li)) continue;
        if (filter) {
          int32_t w;
          memcpy(&w, l->cells + (size_t)li * 4, 4);
          const int32_t p = w >> 4;
          if (filter == 1 && !(p > threshold)) continue;
          if (filter == 2 && !(p < threshold)) continue;
        }
        if (k < cap) {
          if (coords) {
            coords[3 * k] = xb | (int32_t)(li & lm);
            coords[3 * k + 1] = yb | (int32_t)((li >> lb) & lm);
            coords[3 * k + 2] = zb | (int32_t)((li >> (2 * lb)) & lm);
          }
          if (out && vb) memcpy(out + k * vb, l->cells + (size_t)li * vb, vb);
        }
        ++k;
      }
    }
  }
  return k;
}

/* clear(SET_ALL_CELLS_OFF), bonxai.hpp:685-686: every on cell goes off, the
 * leaves stay allocated. */
BXO_EXPORT void bxo_grid_clear_cells(bxo_grid* g) {
  for (uint64_t s = 0; s < g->cap; ++s) {
    if (!g->used[s]) continue;
    bxo_inner* in = g->inner[s];
    for (uint32_t i = 0; i < g->inner_size; ++i)
      if (bxo_bit(in->mask, i)) memset(in->child[i]->mask, 0, g->leaf_words * sizeof(uint64_t));
  }
}

/* releaseUnusedMemory, bonxai.hpp:367-387: drop leaves with no on cell, then
 * roots with no leaf. */
BXO_EXPORT void bxo_grid_release_unused(bxo_grid* g) {
  for (uint64_t s = 0; s < g->cap; ++s) {
    if (!g->used[s]) continue;
    bxo_inner* in = g->inner[s];
    for (uint32_t i = 0; i < g->inner_size; ++i) {
      if (!bxo_bit(in->mask, i)) continue;
      bxo_leaf* l = in->child[i];
      int any = 0;
      for (uint32_t w = 0; w < g->leaf_words; ++w) any |= l->mask[w] != 0;
      if (!any) {
        bxo_bit_clear(in->mask, i);
        bxo_free_leaf(l);
        in->child[i] = NULL;
        g->leaf_count--;
      }
    }
  }
  /* rebuild the directory without empty roots */
  uint64_t ocap = g->cap;
  bxo_coord* ok = g->keys;
  bxo_inner** oi = g->inner;
  uint8_t* ou = g->used;
  g->keys = (bxo_coord*)calloc(ocap, sizeof(bxo_coord));
  g->inner = (bxo_inner**)calloc(ocap, sizeof(bxo_inner*));
  g->used = (uint8_t*)calloc(ocap, 1);
  g->count = 0;
  const uint64_t m = ocap - 1;
  for (uint64_t t = 0; t < ocap; ++t) {
    if (!ou[t]) continue;
    int any = 0;
    for (uint32_t w = 0; w < g->inner_words; ++w) any |= oi[t]->mask[w] != 0;
    if (!any) {
      bxo_free_inner(g, oi[t]);
      continue;
    }
    uint64_t u = bxo_hash_coord(ok[t]) & m;
    while (g->used[u]) u = (u + 1) & m;
    g->used[u] = 1;
    g->keys[u] = ok[t];
    g->inner[u] = oi[t];
    g->count++;
  }
  free(ok);
  free(oi);
  free(ou);
}

/* posToCoord, bonxai.hpp:404-410 */
BXO_EXPORT void bxo_pos_to_coord(const double* xyz, uint64_t n, double inv_resolution,
                                 int32_t* out) {
  for (uint64_t i = 0; i < 3 * n; ++i) out[i] = (int32_t)floor(xyz[i] * inv_resolution);
}

/* ----------------------------------------------------- order-free digest */

/* A multiset digest of the on cells: count, wrapping sum, xor and wrapping
 * sum of a second mix of a 64-bit hash of (x, y, z, value word).  It does not
 * depend on iteration order, so the GPU export can be compared without a sort.
 * The same function is implemented by the HIP library. */
static uint64_t bxo_mix64(uint64_t h) {
  h ^= h >> 33;
  h *= 0xFF51AFD7ED558CCDull;
  h ^= h >> 33;
  h *= 0xC4CEB9FE1A85EC53ull;
  h ^= h >> 33;
  return h;
}
static uint64_t bxo_cell_hash(int32_t x, int32_t y, int32_t z, uint64_t word) {
  uint64_t a = (uint64_t)(uint32_t)x | ((uint64_t)(uint32_t)y << 32);
  uint64_t b = (uint64_t)(uint32_t)z ^ (word * 0x9E3779B97F4A7C15ull);
  return bxo_mix64(a ^ bxo_mix64(b + 0x632BE59BD9B4E019ull));
}
static uint64_t bxo_value_word(const uint8_t* p, uint32_t vb) {
  uint64_t w = 0;
  if (vb <= 8) {
    memcpy(&w, p, vb);
    return w;
  }
  for (uint32_t o = 0; o < vb; o += 8) {
    uint64_t part = 0;
    memcpy(&part, p + o, (vb - o) < 8 ? (vb - o) : 8);
    w = bxo_mix64(w ^ part);
  }
  return w;
}

BXO_EXPORT void bxo_digest(const bxo_grid* g, uint64_t out[4]) {
  const int32_t lm = (1 << g->leaf_bits) - 1, im = (1 << g->inner_bits) - 1;
  const uint32_t lb = g->leaf_bits, ib = g->inner_bits, vb = g->value_bytes;
  uint64_t cnt = 0, sum = 0, x = 0, sum2 = 0;
  for (uint64_t s = 0; s < g->cap; ++s) {
    if (!g->used[s]) continue;
    const bxo_coord r = g->keys[s];
    const bxo_inner* in = g->inner[s];
    for (uint32_t ii = 0; ii < g->inner_size; ++ii) {
      if (!bxo_bit(in->mask, ii)) continue;
      const int32_t xb = r.x | ((int32_t)(ii & im) << lb);
      const int32_t yb = r.y | ((int32_t)((ii >> ib) & im) << lb);
      const int32_t zb = r.z | ((int32_t)((ii >> (2 * ib)) & im) << lb);
      const bxo_leaf* l = in->child[ii];
      for (uint32_t li = 0; li < g->leaf_size; ++li) {
        if (!bxo_bit(l->mask, li)) continue;
        const uint64_t w = vb ? bxo_value_word(l->cells + (size_t)li * vb, vb) : 0;
        const uint64_t h = bxo_cell_hash(xb | (int32_t)(li & lm), yb | (int32_t)((li >> lb) & lm),
                                         zb | (int32_t)((li >> (2 * lb)) & lm), w);
        cnt++;
        sum += h;
        x ^= h;
        sum2 += bxo_mix64(h ^ 0xA0761D6478BD642Full);
      }
    }
  }
  out[0] = cnt;
  out[1] = sum;
  out[2] = x;
  out[3] = sum2;
}

/* --------------------------------------------------- ProbabilisticMap */

typedef struct bxo_prob {
  bxo_grid* g;
  int32_t miss, hit, cmin, cmax, thr;
  uint8_t cur; /* _update_count, probabilistic_map.hpp:129 */
  bxo_accessor acc; /* the map's own accessor, probabilistic_map.hpp:134 */
  bxo_coord* hits;
  uint64_t nhits, caphits;
  bxo_coord* misses;
  uint64_t nmiss, capmiss;
  /* statistics of the last scan */
  uint64_t st_endpoint_updates, st_rays, st_visits, st_ray_updates;
} bxo_prob;

/* CellT raw word layout (x86-64 GCC bit-field allocation, LSB first):
 * bits 0..3 update_id (signed), bits 4..31 probability_log (signed). */
static int32_t bxo_cell_id(uint32_t w) { return ((int32_t)(w << 28)) >> 28; }
static int32_t bxo_cell_p(uint32_t w) { return ((int32_t)w) >> 4; }
static uint32_t bxo_cell_make(int32_t p, int32_t id) {
  return ((uint32_t)p << 4) | ((uint32_t)id & 0xFu);
}

BXO_EXPORT bxo_prob* bxo_prob_create(double resolution) {
  bxo_prob* m = (bxo_prob*)calloc(1, sizeof(bxo_prob));
  m->g = bxo_grid_create(resolution, 2, 3, 4); /* probabilistic_map.cpp:15 */
  /* Options defaults, logods() of 0.4f, 0.7f, 0.12f, 0.97f, 0.5 */
  m->miss = -405465;
  m->hit = 847297;
  m->cmin = -1992430;
  m->cmax = 3476099;
  m->thr = 0;
  m->cur = 1;
  bxo_accessor_init(&m->acc, m->g);
  return m;
}

BXO_EXPORT void bxo_prob_destroy(bxo_prob* m) {
  if (!m) return;
  bxo_grid_destroy(m->g);
  free(m->hits);
  free(m->misses);
  free(m);
}

BXO_EXPORT bxo_grid* bxo_prob_grid(bxo_prob* m) { return m->g; }

BXO_EXPORT void bxo_prob_set_options(bxo_prob* m, const int32_t opts[5]) {
  m->miss = opts[0];
  m->hit = opts[1];
  m->cmin = opts[2];
  m->cmax = opts[3];
  m->thr = opts[4];
}

BXO_EXPORT uint32_t bxo_prob_update_count(const bxo_prob* m) { return m->cur; }
BXO_EXPORT void bxo_prob_set_update_count(bxo_prob* m, uint32_t c) { m->cur = (uint8_t)c; }

BXO_EXPORT void bxo_prob_stats(const bxo_prob* m, uint64_t out[4]) {
  out[0] = m->st_endpoint_updates;
  out[1] = m->st_rays;
  out[2] = m->st_visits;
  out[3] = m->st_ray_updates;
}

static void bxo_push(bxo_coord** v, uint64_t* n, uint64_t* cap, bxo_coord c) {
  if (*n == *cap) {
    *cap = *cap ? *cap * 2 : 1024;
    *v = (bxo_coord*)realloc(*v, *cap * sizeof(bxo_coord));
  }
  (*v)[(*n)++] = c;
}

static bxo_coord bxo_pos2coord(const bxo_grid* g, double x, double y, double z) {
  bxo_coord c = {(int32_t)floor(x * g->inv_resolution), (int32_t)floor(y * g->inv_resolution),
                 (int32_t)floor(z * g->inv_resolution)};
  return c;
}

/* addHitPoint, probabilistic_map.cpp:30-41 */
static void bxo_add_hit(bxo_prob* m, double x, double y, double z) {
  const bxo_coord c = bxo_pos2coord(m->g, x, y, z);
  uint8_t* cell = bxo_value(&m->acc, c, 1);
  uint32_t w;
  memcpy(&w, cell, 4);
  if (bxo_cell_id(w) != (int32_t)m->cur) {
    int32_t p = (int32_t)((uint32_t)bxo_cell_p(w) + (uint32_t)m->hit);
    if (p > m->cmax) p = m->cmax;
    w = bxo_cell_make(p, m->cur);
    memcpy(cell, &w, 4);
    m->st_endpoint_updates++;
    bxo_push(&m->hits, &m->nhits, &m->caphits, c);
  }
}

/* addMissPoint, probabilistic_map.cpp:43-54 */
static void bxo_add_miss(bxo_prob* m, double x, double y, double z) {
  const bxo_coord c = bxo_pos2coord(m->g, x, y, z);
  uint8_t* cell = bxo_value(&m->acc, c, 1);
  uint32_t w;
  memcpy(&w, cell, 4);
  if (bxo_cell_id(w) != (int32_t)m->cur) {
    int32_t p = (int32_t)((uint32_t)bxo_cell_p(w) + (uint32_t)m->miss);
    if (p < m->cmin) p = m->cmin;
    w = bxo_cell_make(p, m->cur);
    memcpy(cell, &w, 4);
    m->st_endpoint_updates++;
    bxo_push(&m->misses, &m->nmiss, &m->capmiss, c);
  }
}

BXO_EXPORT void bxo_prob_add_hit(bxo_prob* m, const double p[3]) { bxo_add_hit(m, p[0], p[1], p[2]); }
BXO_EXPORT void bxo_prob_add_miss(bxo_prob* m, const double p[3]) { bxo_add_miss(m, p[0], p[1], p[2]); }

/* The clearPoint lambda of updateFreeCells, probabilistic_map.cpp:81-89, over
 * the cells of one ray.  The walk keeps the current leaf's mask and cell words
 * and the scan's constants in locals: cell stores then cannot alias them, so
 * the loop does not reload its state after every store (same results as going
 * through bxo_value for every cell). */
typedef struct {
  bxo_coord key; /* inner key (leaf origin) of the cached leaf */
  uint64_t* mask;
  uint32_t* cells;
  int valid;
} bxo_ray_leaf;

static inline uint32_t bxo_clear_cell(bxo_accessor* a, bxo_ray_leaf* rl, bxo_coord c,
                                      int32_t cur, int32_t miss, int32_t cmin) {
  const bxo_grid* g = a->g;
  const bxo_coord ik = bxo_inner_key(g, c);
  if (!rl->valid || !bxo_coord_eq(ik, rl->key)) {
    bxo_leaf* l = bxo_cached_leaf(a, c, 1); /* Accessor::value(c, true), bonxai.hpp:469-494 */
    rl->key = ik;
    rl->mask = l->mask;
    rl->cells = (uint32_t*)l->cells;
    rl->valid = 1;
  }
  const uint32_t li = bxo_leaf_index(g, c);
  const uint64_t bit = 1ull << (li & 63);
  uint32_t w = 0;
  if (rl->mask[li >> 6] & bit) {
    w = rl->cells[li];
  } else {
    rl->mask[li >> 6] |= bit; /* created value-initialised, bonxai.hpp:487-490 */
    rl->cells[li] = 0u;
  }
  if (bxo_cell_id(w) != cur) {
    int32_t p = (int32_t)((uint32_t)bxo_cell_p(w) + (uint32_t)miss);
    if (p < cmin) p = cmin;
    rl->cells[li] = bxo_cell_make(p, cur);
    return 1;
  }
  return 0;
}

/* RayIterator, probabilistic_map.hpp:162-203: visits the origin cell and then
 * max|delta|-1 further cells of an integer 3-D Bresenham walk; the end cell is
 * not visited; nothing is visited when origin == end. */
static void bxo_ray(bxo_prob* m, bxo_accessor* a, bxo_coord o, bxo_coord e) {
  if (bxo_coord_eq(o, e)) return;
  m->st_rays++;
  const int32_t cur = (int32_t)m->cur, miss = m->miss, cmin = m->cmin;
  bxo_ray_leaf rl = {{0, 0, 0}, NULL, NULL, 0};
  uint64_t upd = bxo_clear_cell(a, &rl, o, cur, miss, cmin);
  int32_t dx = e.x - o.x, dy = e.y - o.y, dz = e.z - o.z;
  const int32_t sx = dx < 0 ? -1 : 1, sy = dy < 0 ? -1 : 1, sz = dz < 0 ? -1 : 1;
  dx = dx < 0 ? -dx : dx;
  dy = dy < 0 ? -dy : dy;
  dz = dz < 0 ? -dz : dz;
  int32_t mx = dx > dy ? dx : dy;
  mx = mx > dz ? mx : dz;
  int32_t ex = 0, ey = 0, ez = 0;
  bxo_coord c = o;
  for (int32_t i = 0; i < mx - 1; ++i) {
    ex += dx;
    ey += dy;
    ez += dz;
    if (2 * ex >= mx) {
      c.x += sx;
      ex -= mx;
    }
    if (2 * ey >= mx) {
      c.y += sy;
      ey -= mx;
    }
    if (2 * ez >= mx) {
      c.z += sz;
      ez -= mx;
    }
    upd += bxo_clear_cell(a, &rl, c, cur, miss, cmin);
  }
  m->st_visits += (uint64_t)mx;
  m->st_ray_updates += upd;
}

/* updateFreeCells, probabilistic_map.cpp:77-106 */
static void bxo_update_free_cells(bxo_prob* m, double ox, double oy, double oz) {
  bxo_accessor a;
  bxo_accessor_init(&a, m->g);
  const bxo_coord o = bxo_pos2coord(m->g, ox, oy, oz);
  for (uint64_t i = 0; i < m->nhits; ++i) bxo_ray(m, &a, o, m->hits[i]);
  m->nhits = 0;
  for (uint64_t i = 0; i < m->nmiss; ++i) bxo_ray(m, &a, o, m->misses[i]);
  m->nmiss = 0;
  if (++m->cur == 4) m->cur = 1;
}

BXO_EXPORT void bxo_prob_update_free_cells(bxo_prob* m, const double origin[3]) {
  bxo_update_free_cells(m, origin[0], origin[1], origin[2]);
}

/* insertPointCloud, probabilistic_map.hpp:141-160.  Points are float32 or
 * float64 xyz triples (is_f64 selects); the origin is given in double, i.e.
 * already widened from the caller's point type. */
static void bxo_insert(bxo_prob* m, const void* pts, int is_f64, uint64_t n, const double origin[3],
                       double max_range) {
  m->st_endpoint_updates = m->st_rays = m->st_visits = m->st_ray_updates = 0;
  const double fx = origin[0], fy = origin[1], fz = origin[2];
  const double r2 = max_range * max_range;
  const float* pf = (const float*)pts;
  const double* pd = (const double*)pts;
  for (uint64_t i = 0; i < n; ++i) {
    double tx, ty, tz;
    if (is_f64) {
      tx = pd[3 * i];
      ty = pd[3 * i + 1];
      tz = pd[3 * i + 2];
    } else {
      tx = (double)pf[3 * i];
      ty = (double)pf[3 * i + 1];
      tz = (double)pf[3 * i + 2];
    }
    const double vx = tx - fx, vy = ty - fy, vz = tz - fz;
    const double sq = (vx * vx + vy * vy) + vz * vz;
    if (sq >= r2) {
      const double s = sqrt(sq);
      const double qx = fx + ((vx / s) * max_range);
      const double qy = fy + ((vy / s) * max_range);
      const double qz = fz + ((vz / s) * max_range);
      bxo_add_miss(m, qx, qy, qz);
    } else {
      bxo_add_hit(m, tx, ty, tz);
    }
  }
  bxo_update_free_cells(m, fx, fy, fz);
}

BXO_EXPORT void bxo_prob_insert_f32(bxo_prob* m, const float* pts, uint64_t n,
                                    const double origin[3], double max_range) {
  bxo_insert(m, pts, 0, n, origin, max_range);
}

BXO_EXPORT void bxo_prob_insert_f64(bxo_prob* m, const double* pts, uint64_t n,
                                    const double origin[3], double max_range) {
  bxo_insert(m, pts, 1, n, origin, max_range);
}

/* Endpoint coordinate and kind of each point (1 hit, 0 miss) as computed by
 * insertPointCloud, plus the endpoint under the alternative squaredNorm
 * association x*x + (y*y + z*z) for the ambiguity count. */
static void bxo_endpoint(double fx, double fy, double fz, double tx, double ty, double tz,
                         double max_range, double r2, double inv, int alt, int32_t* c,
                         uint8_t* kind) {
  const double vx = tx - fx, vy = ty - fy, vz = tz - fz;
  const double sq = alt ? vx * vx + (vy * vy + vz * vz) : (vx * vx + vy * vy) + vz * vz;
  double qx = tx, qy = ty, qz = tz;
  *kind = 1;
  if (sq >= r2) {
    const double s = sqrt(sq);
    qx = fx + ((vx / s) * max_range);
    qy = fy + ((vy / s) * max_range);
    qz = fz + ((vz / s) * max_range);
    *kind = 0;
  }
  c[0] = (int32_t)floor(qx * inv);
  c[1] = (int32_t)floor(qy * inv);
  c[2] = (int32_t)floor(qz * inv);
}

BXO_EXPORT uint64_t bxo_endpoints_f32(const float* pts, uint64_t n, const double origin[3],
                                      double max_range, double inv_resolution, int32_t* coords,
                                      uint8_t* kind) {
  uint64_t ambiguous = 0;
  const double fx = origin[0], fy = origin[1], fz = origin[2];
  const double r2 = max_range * max_range;
  for (uint64_t i = 0; i < n; ++i) {
    const double tx = (double)pts[3 * i], ty = (double)pts[3 * i + 1], tz = (double)pts[3 * i + 2];
    int32_t c[3], c2[3];
    uint8_t k, k2;
    bxo_endpoint(fx, fy, fz, tx, ty, tz, max_range, r2, inv_resolution, 0, c, &k);
    bxo_endpoint(fx, fy, fz, tx, ty, tz, max_range, r2, inv_resolution, 1, c2, &k2);
    if (coords) memcpy(coords + 3 * i, c, sizeof(c));
    if (kind) kind[i] = k;
    ambiguous += (k != k2 || c[0] != c2[0] || c[1] != c2[1] || c[2] != c2[2]);
  }
  return ambiguous;
}

/* The ingest pre-step of bonxai_server.cpp:147-179 in front of
 * insertPointCloud: points with a non-finite coordinate are dropped
 * (:147-153, order kept), the others moved by the float sensor-to-world
 * matrix (pcl::transformPointCloud, :166-170), and the sensor origin is the
 * matrix's translation widened to double (:172-175).  PCL (absent from this
 * image, version unpinned) computes x' = m00*x + (m01*y + (m02*z + m03)) in
 * its SSE build and ((m00*x + m01*y) + m02*z) + m03 in its scalar build; the
 * SSE association is taken (parity unpinned: bxo_tf_ambiguity counts the
 * points whose endpoint differs between the two).  stride: floats from one
 * point to the next. */
static void bxo_tf_point(const float* t, const float* p, int alt, float* q) {
  for (int r = 0; r < 3; ++r) {
    const float* row = t + 4 * r;
    q[r] = alt ? ((row[0] * p[0] + row[1] * p[1]) + row[2] * p[2]) + row[3]
               : row[0] * p[0] + (row[1] * p[1] + (row[2] * p[2] + row[3]));
  }
}

static uint64_t bxo_ingest(const float* pts, uint64_t n, uint64_t stride, const float m16[16], int alt,
                           float* out) {
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const float* p = pts + i * stride;
    if (!(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]))) continue;
    bxo_tf_point(m16, p, alt, out + 3 * k);
    ++k;
  }
  return k;
}

BXO_EXPORT void bxo_prob_insert_tf(bxo_prob* m, const float* pts, uint64_t n, uint64_t stride,
                                   const float m16[16], double max_range) {
  float* q = (float*)malloc((n ? n : 1) * 3 * sizeof(float));
  const uint64_t k = bxo_ingest(pts, n, stride, m16, 0, q);
  const double origin[3] = {(double)m16[3], (double)m16[7], (double)m16[11]};
  bxo_insert(m, q, 0, k, origin, max_range);
  free(q);
}

/* Points whose endpoint voxel or hit/miss kind differs between the two
 * associations of the float transform. */
BXO_EXPORT uint64_t bxo_tf_ambiguity(const float* pts, uint64_t n, uint64_t stride, const float m16[16],
                                     double max_range, double inv_resolution) {
  float* a = (float*)malloc((n ? n : 1) * 3 * sizeof(float));
  float* b = (float*)malloc((n ? n : 1) * 3 * sizeof(float));
  const uint64_t k = bxo_ingest(pts, n, stride, m16, 0, a);
  bxo_ingest(pts, n, stride, m16, 1, b);
  const double fx = m16[3], fy = m16[7], fz = m16[11], r2 = max_range * max_range;
  uint64_t diff = 0;
  for (uint64_t i = 0; i < k; ++i) {
    int32_t c[3], c2[3];
    uint8_t ka, kb;
    bxo_endpoint(fx, fy, fz, a[3 * i], a[3 * i + 1], a[3 * i + 2], max_range, r2, inv_resolution, 0, c, &ka);
    bxo_endpoint(fx, fy, fz, b[3 * i], b[3 * i + 1], b[3 * i + 2], max_range, r2, inv_resolution, 0, c2, &kb);
    diff += (ka != kb || c[0] != c2[0] || c[1] != c2[1] || c[2] != c2[2]);
  }
  free(a);
  free(b);
  return diff;
}

/* isOccupied / isUnknown / isFree, probabilistic_map.cpp:56-75:
 * state 1 occupied, -1 free, 0 unknown (missing cell or p == threshold). */
BXO_EXPORT void bxo_prob_classify(bxo_prob* m, const int32_t* coords, uint64_t n, int8_t* state) {
  bxo_accessor a;
  bxo_accessor_init(&a, m->g);
  for (uint64_t i = 0; i < n; ++i) {
    bxo_coord c = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
    uint8_t* cell = bxo_value(&a, c, 0);
    int8_t s = 0;
    if (cell) {
      uint32_t w;
      memcpy(&w, cell, 4);
      const int32_t p = bxo_cell_p(w);
      s = p > m->thr ? 1 : (p < m->thr ? -1 : 0);
    }
    state[i] = s;
  }
}

/* ComputeRay, probabilistic_map.hpp:13-19: writes up to cap cells of the ray
 * and returns its length. */
BXO_EXPORT uint64_t bxo_compute_ray(const int32_t o[3], const int32_t e[3], int32_t* out,
                                    uint64_t cap) {
  bxo_coord oc = {o[0], o[1], o[2]}, ec = {e[0], e[1], e[2]};
  if (bxo_coord_eq(oc, ec)) return 0;
  uint64_t k = 0;
  int32_t dx = ec.x - oc.x, dy = ec.y - oc.y, dz = ec.z - oc.z;
  const int32_t sx = dx < 0 ? -1 : 1, sy = dy < 0 ? -1 : 1, sz = dz < 0 ? -1 : 1;
  dx = dx < 0 ? -dx : dx;
  dy = dy < 0 ? -dy : dy;
  dz = dz < 0 ? -dz : dz;
  int32_t mx = dx > dy ? dx : dy;
  mx = mx > dz ? mx : dz;
  int32_t ex = 0, ey = 0, ez = 0;
  bxo_coord c = oc;
  if (k < cap) {
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
  }
  k++;
  for (int32_t i = 0; i < mx - 1; ++i) {
    ex += dx;
    ey += dy;
    ez += dz;
    if (2 * ex >= mx) {
      c.x += sx;
      ex -= mx;
    }
    if (2 * ey >= mx) {
      c.y += sy;
      ey -= mx;
    }
    if (2 * ez >= mx) {
      c.z += sz;
      ez -= mx;
    }
    if (k < cap) {
      out[3 * k] = c.x;
      out[3 * k + 1] = c.y;
      out[3 * k + 2] = c.z;
    }
    k++;
  }
  return k;
}
