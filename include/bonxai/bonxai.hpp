// bonxai.hpp -- Bonxai::VoxelGrid<T> backed by a sparse voxel map in GPU memory.
//
// A drop-in for the reference's header API (bonxai_core/include/bonxai/bonxai.hpp:114-333):
// same class, method and enum names, same argument meaning, same results.  The
// map itself -- root directory, inner grids, leaf pool -- lives on the GPU and is
// driven through the C ABI of libbonxai_gpu (bonxai_gpu.h).
//
// Accessors keep the reference's pointer semantics with host pages: the leaf a
// coordinate falls in is copied to host memory on first access (one leaf = its
// mask words + its cells), Accessor::value() returns a pointer into that page,
// and the grid keeps its pages coherent with the device copy:
//   - before any device-side operation (insertPointCloud, bulk calls, counts,
//     exports, forEachCell) every page created or changed on the host is
//     written back (pages are compared with the bytes last exchanged, so writes
//     made through returned pointers are found too);
//   - after a device-side operation that may change cells, every cached page is
//     read back in place, so pointers obtained earlier stay valid and see the
//     new values.
// Pointers stay valid until clear() or releaseUnusedMemory() (bonxai.hpp:166-171),
// as in the reference.  Leaves never seen by the host are not paged: while the
// device holds only leaves the host has seen (a grid filled through accessors),
// misses are answered without a device round trip.
//
// Not provided: rootMap() (the directory is a device hash table; use
// forEachCell or the bulk calls) and allocateLeafGrid() (an allocator hook).
// Cell types must be trivially copyable (they are stored as raw bytes on the GPU).
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "bonxai/gpu_runtime.hpp"
#include "bonxai/grid_coord.hpp"
#include "bonxai/mask.hpp"

namespace Bonxai {

/// Cell type of BinaryVoxelGrid: cells carry only their on/off bit (bonxai.hpp:29-30).
struct EmptyVoxel {};

/// bonxai.hpp:106-112
enum ClearOption {
  CLEAR_MEMORY,       // free every leaf and root
  SET_ALL_CELLS_OFF   // keep the leaves, turn every cell off
};

template <typename DataT>
class VoxelGrid;

/// A cube of 2^(3 log2dim) cells with its mask (bonxai.hpp:39-103).  In the
/// drop-in it is the host page of one leaf.
template <typename DataT>
class Grid {
 public:
  explicit Grid(size_t log2dim) : dim_(static_cast<uint8_t>(1u << log2dim)), mask_(log2dim) {
    if constexpr (!std::is_same_v<DataT, EmptyVoxel>) {
      size_ = uint32_t(dim_) * dim_ * dim_;
      data_.reset(new DataT[size_]());
    }
  }
  Grid(const Grid&) = delete;
  Grid& operator=(const Grid&) = delete;

  [[nodiscard]] size_t size() const { return size_; }
  [[nodiscard]] Mask& mask() { return mask_; }
  [[nodiscard]] const Mask& mask() const { return mask_; }
  [[nodiscard]] DataT& cell(size_t index) {
    static_assert(!std::is_same_v<DataT, EmptyVoxel>, "Cells have no value, when DataT == EmptyVoxel");
    return data_[index];
  }
  [[nodiscard]] const DataT& cell(size_t index) const {
    static_assert(!std::is_same_v<DataT, EmptyVoxel>, "Cells have no value, when DataT == EmptyVoxel");
    return data_[index];
  }
  [[nodiscard]] size_t memUsage() const {
    return mask_.memUsage() + sizeof(*this) + (std::is_same_v<DataT, EmptyVoxel> ? 0 : sizeof(DataT) * size_);
  }

 private:
  friend class VoxelGrid<DataT>;
  uint8_t dim_;
  uint32_t size_ = 0;
  Mask mask_;
  std::unique_ptr<DataT[]> data_;
  // paging state: present on the device, and the bytes last exchanged with it
  bool on_device_ = false;
  std::vector<uint8_t> shadow_;
};

template <typename DataT>
class VoxelGrid {
  static_assert(std::is_trivially_copyable_v<DataT>,
                "GPU-resident cells are raw bytes: DataT must be trivially copyable");
  static constexpr uint32_t kValueBytes = std::is_same_v<DataT, EmptyVoxel> ? 0u : uint32_t(sizeof(DataT));

 public:
  using LeafGrid = Grid<DataT>;

  /// bonxai.hpp:138 / 389-402, on the default GPU context.
  explicit VoxelGrid(double voxel_size, uint8_t inner_bits = 2, uint8_t leaf_bits = 3)
      : VoxelGrid(voxel_size, inner_bits, leaf_bits, gpu::Context::Default(), 0u) {}

  /// The same on a given context; map_flags are bxg_map_create's flags.
  VoxelGrid(double voxel_size, uint8_t inner_bits, uint8_t leaf_bits, gpu::Context& ctx,
            uint32_t map_flags = 0u)
      : st_(new State()) {
    if (leaf_bits < 1 || inner_bits < 1)
      throw std::runtime_error("The minimum value of the inner_bits and leaf_bits should be 1");
    State& s = *st_;
    s.ctx = ctx.get();
    s.inner_bits = inner_bits;
    s.leaf_bits = leaf_bits;
    s.log2n = uint32_t(inner_bits) + leaf_bits;
    s.res = voxel_size;
    s.inv_res = 1.0 / voxel_size;
    s.leaf_size = 1u << (3 * leaf_bits);
    s.leaf_words = std::max<uint32_t>(s.leaf_size >> 6, 1u);
    bxg_map* m = nullptr;
    gpu::check(bxg_map_create(s.ctx, voxel_size, inner_bits, leaf_bits, kValueBytes, map_flags, &m), s.ctx);
    s.map = m;
  }

  VoxelGrid(const VoxelGrid&) = delete;
  VoxelGrid& operator=(const VoxelGrid&) = delete;
  VoxelGrid(VoxelGrid&&) noexcept = default;
  VoxelGrid& operator=(VoxelGrid&&) noexcept = default;

  /// (sic) the reference's name for innerBits (bonxai.hpp:147)
  uint32_t innetBits() const { return st_->inner_bits; }
  uint32_t innerBits() const { return st_->inner_bits; }
  uint32_t leafBits() const { return st_->leaf_bits; }
  double voxelSize() const { return st_->res; }

  /// Bytes held for this grid: the device pools plus the host pages.
  [[nodiscard]] size_t memUsage() const {
    uint64_t dev = 0;
    gpu::check(bxg_map_mem_usage(st_->map, &dev), st_->ctx);
    std::lock_guard<std::recursive_mutex> lk(st_->mu);
    size_t host = 0;
    for (const auto& kv : st_->pages)
      if (kv.second) host += kv.second->memUsage() + kv.second->shadow_.size();
    return size_t(dev) + host;
  }

  /// Free every leaf whose cells are all off and every root left without
  /// leaves (bonxai.hpp:367-387); the device pools shrink to what is left.
  void releaseUnusedMemory() {
    flush();
    gpu::check(bxg_map_release_unused(st_->map), st_->ctx);
    refresh(false);
  }

  /// bonxai.hpp:689-701
  [[nodiscard]] size_t activeCellsCount() const {
    flush();
    uint64_t n = 0;
    gpu::check(bxg_map_active_cells(st_->map, &n), st_->ctx);
    return size_t(n);
  }

  /// bonxai.hpp:404-410
  [[nodiscard]] CoordT posToCoord(double x, double y, double z) const {
    return {static_cast<int32_t>(std::floor(x * st_->inv_res)),
            static_cast<int32_t>(std::floor(y * st_->inv_res)),
            static_cast<int32_t>(std::floor(z * st_->inv_res))};
  }
  [[nodiscard]] CoordT posToCoord(const Point3D& pos) const { return posToCoord(pos.x, pos.y, pos.z); }
  /// bonxai.hpp:412-417
  [[nodiscard]] Point3D coordToPos(const CoordT& c) const {
    return {double(c.x) * st_->res, double(c.y) * st_->res, double(c.z) * st_->res};
  }

  /// Visit every on cell with func(const DataT&, const CoordT&) (bonxai.hpp:704-743).
  /// The visiting order is the device pool's, not the reference's hash order.
  template <class VisitorFunction>
  void forEachCell(VisitorFunction func) const {
    visit(func, false);
  }
  /// Visit every on cell with func(DataT&, const CoordT&); changed values are
  /// written back to the device.
  template <class VisitorFunction>
  void forEachCell(VisitorFunction func) {
    visit(func, true);
  }

  /// bonxai.hpp:678-687
  void clear(ClearOption opt) {
    if (opt == CLEAR_MEMORY) {
      gpu::check(bxg_map_clear(st_->map, BXG_CLEAR_MEMORY), st_->ctx);
      std::lock_guard<std::recursive_mutex> lk(st_->mu);
      st_->pages.clear();
      st_->complete = true;
      ++st_->gen;
      return;
    }
    flush();
    gpu::check(bxg_map_clear(st_->map, BXG_SET_ALL_CELLS_OFF), st_->ctx);
    refresh(false);
  }

  class ConstAccessor {
   public:
    ConstAccessor(const VoxelGrid& grid) : grid_(grid) {}

    /// bonxai.hpp:496-516: pointer to the value, or nullptr if the cell is off.
    [[nodiscard]] const DataT* value(const CoordT& coord) const {
      static_assert(!std::is_same_v<DataT, EmptyVoxel>,
                    "You can not access a value when using type EmptyVoxel. Use isCellOn");
      const LeafGrid* leaf = leafFor(coord);
      if (!leaf) return nullptr;
      const uint32_t i = grid_.getLeafIndex(coord);
      return leaf->mask().isOn(i) ? &leaf->cell(i) : nullptr;
    }
    /// bonxai.hpp:518-534
    [[nodiscard]] bool isCellOn(const CoordT& coord) const {
      const LeafGrid* leaf = leafFor(coord);
      return leaf && leaf->mask().isOn(grid_.getLeafIndex(coord));
    }
    [[nodiscard]] const LeafGrid* lastLeafGrid() const { return prev_leaf_; }
    /// bonxai.hpp:623-647
    [[nodiscard]] const LeafGrid* getLeafGrid(const CoordT& coord) const { return leafFor(coord); }

   protected:
    const LeafGrid* leafFor(const CoordT& coord) const {
      const CoordT key = grid_.getInnerKey(coord);
      const uint64_t gen = grid_.st_->gen.load(std::memory_order_relaxed);
      if (key != prev_key_ || gen_ != gen) {
        prev_leaf_ = grid_.page(key, false);
        prev_key_ = key;
        gen_ = grid_.st_->gen.load(std::memory_order_relaxed);
      }
      return prev_leaf_;
    }
    const VoxelGrid& grid_;
    mutable CoordT prev_key_ = {std::numeric_limits<int32_t>::max(), 0, 0};
    mutable const LeafGrid* prev_leaf_ = nullptr;
    mutable uint64_t gen_ = ~uint64_t(0);
  };

  class Accessor : public ConstAccessor {
   public:
    Accessor(VoxelGrid& grid) : ConstAccessor(grid), mgrid_(grid) {}

    /// bonxai.hpp:449-466: create the cell if needed, store value; returns
    /// the previous on-state.
    bool setValue(const CoordT& coord, const DataT& value) {
      static_assert(!std::is_same_v<DataT, EmptyVoxel>,
                    "You can not access a value when using type EmptyVoxel. Use setCellOn / setCellOff");
      LeafGrid* leaf = getLeafGrid(coord, true);
      const uint32_t i = mgrid_.getLeafIndex(coord);
      const bool was_on = leaf->mask().setOn(i);
      leaf->cell(i) = value;
      return was_on;
    }

    /// bonxai.hpp:469-494: pointer to the value; with create_if_missing an
    /// off cell is turned on with a value-initialised DataT.
    [[nodiscard]] DataT* value(const CoordT& coord, bool create_if_missing = false) {
      static_assert(!std::is_same_v<DataT, EmptyVoxel>,
                    "You can not access a value when using type EmptyVoxel. Use isCellOn / setCellOn / setCellOff");
      LeafGrid* leaf = getLeafGrid(coord, create_if_missing);
      if (!leaf) return nullptr;
      const uint32_t i = mgrid_.getLeafIndex(coord);
      if (leaf->mask().isOn(i)) return &leaf->cell(i);
      if (!create_if_missing) return nullptr;
      leaf->mask().setOn(i);
      leaf->cell(i) = DataT{};
      return &leaf->cell(i);
    }

    /// bonxai.hpp:537-553: turn the cell on; default_value only if it was off.
    bool setCellOn(const CoordT& coord, const DataT& default_value = DataT()) {
      LeafGrid* leaf = getLeafGrid(coord, true);
      const uint32_t i = mgrid_.getLeafIndex(coord);
      const bool was_on = leaf->mask().setOn(i);
      if constexpr (!std::is_same_v<DataT, EmptyVoxel>) {
        if (!was_on) leaf->cell(i) = default_value;
      }
      return was_on;
    }

    /// bonxai.hpp:556-569: turn the cell off; returns the previous on-state.
    bool setCellOff(const CoordT& coord) {
      LeafGrid* leaf = getLeafGrid(coord, false);
      if (!leaf) return false;
      return leaf->mask().setOff(mgrid_.getLeafIndex(coord));
    }

    /// bonxai.hpp:588-621: the leaf holding coord, created if asked.
    [[nodiscard]] LeafGrid* getLeafGrid(const CoordT& coord, bool create_if_missing = false) {
      const CoordT key = mgrid_.getInnerKey(coord);
      if (key != this->prev_key_ || this->gen_ != mgrid_.st_->gen.load(std::memory_order_relaxed) ||
          (create_if_missing && !this->prev_leaf_)) {
        this->prev_leaf_ = mgrid_.page(key, create_if_missing);
        this->prev_key_ = key;
        this->gen_ = mgrid_.st_->gen.load(std::memory_order_relaxed);
      }
      return const_cast<LeafGrid*>(this->prev_leaf_);
    }

   private:
    VoxelGrid& mgrid_;
  };

  Accessor createAccessor() { return Accessor(*this); }
  ConstAccessor createConstAccessor() const { return ConstAccessor(*this); }

  /// bonxai.hpp:419-447
  [[nodiscard]] CoordT getRootKey(const CoordT& c) const {
    const int32_t m = ~((1 << st_->log2n) - 1);
    return {c.x & m, c.y & m, c.z & m};
  }
  [[nodiscard]] CoordT getInnerKey(const CoordT& c) const {
    const int32_t m = ~((1 << st_->leaf_bits) - 1);
    return {c.x & m, c.y & m, c.z & m};
  }
  [[nodiscard]] uint32_t getInnerIndex(const CoordT& c) const {
    const uint32_t lb = st_->leaf_bits, ib = st_->inner_bits, im = (1u << ib) - 1;
    return (uint32_t(c.x >> lb) & im) | ((uint32_t(c.y >> lb) & im) << ib) |
           ((uint32_t(c.z >> lb) & im) << (2 * ib));
  }
  [[nodiscard]] uint32_t getLeafIndex(const CoordT& c) const {
    const uint32_t lb = st_->leaf_bits, lm = (1u << lb) - 1;
    return (uint32_t(c.x) & lm) | ((uint32_t(c.y) & lm) << lb) | ((uint32_t(c.z) & lm) << (2 * lb));
  }

  // ------------------------------------------------------------ GPU extras

  bxg_map* deviceMap() const { return st_->map; }
  bxg_ctx* deviceContext() const { return st_->ctx; }

  /// Write every page created or changed on the host to the device.
  void flush() const {
    std::lock_guard<std::recursive_mutex> lk(st_->mu);
    flushLocked();
  }

  /// Re-read every cached page after device-side changes made directly
  /// through deviceMap(); `may_create` says whether leaves may have been added.
  void refresh(bool may_create) const {
    std::lock_guard<std::recursive_mutex> lk(st_->mu);
    refreshLocked(may_create, nullptr, 0.0);
  }

  /// refresh(true) for a device-side change that touched no cell farther than
  /// `radius` (world units) from `center`, such as a scan from `center` whose
  /// rays end within that reach: only the cached pages whose leaf box comes
  /// within `radius` of `center` are re-read.
  void refreshNear(const Point3D& center, double radius) const {
    std::lock_guard<std::recursive_mutex> lk(st_->mu);
    refreshLocked(true, &center, radius);
  }

  /// Accessor::setValue over a batch, in input order, on the GPU.
  void setValues(const std::vector<CoordT>& coords, const std::vector<DataT>& values,
                 std::vector<uint8_t>* was_on = nullptr) {
    static_assert(!std::is_same_v<DataT, EmptyVoxel>, "EmptyVoxel cells have no value");
    if (coords.size() != values.size()) throw std::invalid_argument("coords and values differ in size");
    if (was_on) was_on->assign(coords.size(), 0);
    flush();
    gpu::check(bxg_set_values_host(st_->map, coordPtr(coords), values.data(), coords.size(),
                                   was_on ? was_on->data() : nullptr),
               st_->ctx);
    refresh(true);
  }

  /// ConstAccessor::value over a batch: values (zero where off) and found flags.
  void values(const std::vector<CoordT>& coords, std::vector<DataT>& out,
              std::vector<uint8_t>& found) const {
    out.resize(coords.size());
    found.assign(coords.size(), 0);
    flush();
    gpu::check(bxg_lookup_host(st_->map, coordPtr(coords), coords.size(),
                               kValueBytes ? static_cast<void*>(out.data()) : nullptr, found.data()),
               st_->ctx);
  }

  /// Accessor::setCellOn / setCellOff / isCellOn over a batch.
  void setCellsOn(const std::vector<CoordT>& coords, std::vector<uint8_t>* was_on = nullptr) {
    if (was_on) was_on->assign(coords.size(), 0);
    flush();
    gpu::check(bxg_set_cells_on_host(st_->map, coordPtr(coords), nullptr, coords.size(),
                                     was_on ? was_on->data() : nullptr),
               st_->ctx);
    refresh(true);
  }
  void setCellsOff(const std::vector<CoordT>& coords, std::vector<uint8_t>* was_on = nullptr) {
    if (was_on) was_on->assign(coords.size(), 0);
    flush();
    gpu::check(bxg_set_cells_off_host(st_->map, coordPtr(coords), coords.size(),
                                      was_on ? was_on->data() : nullptr),
               st_->ctx);
    refresh(false);
  }
  void cellsOn(const std::vector<CoordT>& coords, std::vector<uint8_t>& on) const {
    on.assign(coords.size(), 0);
    flush();
    gpu::check(bxg_lookup_host(st_->map, coordPtr(coords), coords.size(), nullptr, on.data()), st_->ctx);
  }

  /// Every on cell (coordinates and values), in the device pool's order.
  void exportCells(std::vector<CoordT>& coords, std::vector<DataT>* vals = nullptr,
                   int filter = BXG_EXPORT_ALL) const {
    flush();
    uint64_t n = 0;
    int rc = bxg_export_host(st_->map, filter, nullptr, nullptr, 0, &n);
    if (rc != BXG_OK && rc != BXG_ERR_CAPACITY) gpu::check(rc, st_->ctx);
    coords.resize(n);
    if (vals) vals->resize(n);
    if (!n) return;
    uint64_t got = 0;
    gpu::check(bxg_export_host(st_->map, filter, &coords[0].x,
                               vals && kValueBytes ? static_cast<void*>(vals->data()) : nullptr, n, &got),
               st_->ctx);
    coords.resize(got);
    if (vals) vals->resize(got);
  }

  /// Wait for the device work of this grid; throws the map's device error, if any.
  void sync() const {
    flush();
    gpu::check(bxg_map_sync(st_->map), st_->ctx);
  }

 private:
  struct CoordHash {
    size_t operator()(const CoordT& c) const {
      uint64_t h = uint64_t(uint32_t(c.x)) * 0x9E3779B97F4A7C15ull;
      h ^= uint64_t(uint32_t(c.y)) * 0xC2B2AE3D27D4EB4Full;
      h ^= uint64_t(uint32_t(c.z)) * 0x165667B19E3779F9ull;
      return size_t(h ^ (h >> 29));
    }
  };
  struct State {
    bxg_ctx* ctx = nullptr;
    bxg_map* map = nullptr;
    uint32_t inner_bits = 0, leaf_bits = 0, log2n = 0, leaf_size = 0, leaf_words = 0;
    double res = 0, inv_res = 0;
    // host pages by leaf origin; nullptr = known to be absent on the device
    std::unordered_map<CoordT, std::unique_ptr<LeafGrid>, CoordHash> pages;
    // the device holds no leaf that is missing from `pages`
    bool complete = true;
    // bumped whenever a page appears or disappears (accessor caches check it)
    std::atomic<uint64_t> gen{0};
    std::recursive_mutex mu;
    ~State() {
      if (map) bxg_map_destroy(map);
    }
  };

  static const int32_t* coordPtr(const std::vector<CoordT>& v) {
    static_assert(sizeof(CoordT) == 12, "CoordT must be three packed int32");
    return v.empty() ? nullptr : &v[0].x;
  }
  size_t cellBytes() const { return size_t(st_->leaf_size) * kValueBytes; }
  size_t maskBytes() const { return size_t(st_->leaf_words) * 8; }
  uint8_t* cellPtr(LeafGrid* g) const {
    if constexpr (std::is_same_v<DataT, EmptyVoxel>) {
      return nullptr;
    } else {
      return reinterpret_cast<uint8_t*>(g->data_.get());
    }
  }
  void snapshot(LeafGrid* g) const {
    g->shadow_.resize(maskBytes() + cellBytes());
    std::memcpy(g->shadow_.data(), g->mask_.data(), maskBytes());
    if (cellBytes()) std::memcpy(g->shadow_.data() + maskBytes(), cellPtr(g), cellBytes());
  }
  bool changed(LeafGrid* g) const {
    if (g->shadow_.size() != maskBytes() + cellBytes()) return true;
    if (std::memcmp(g->shadow_.data(), g->mask_.data(), maskBytes())) return true;
    return cellBytes() && std::memcmp(g->shadow_.data() + maskBytes(), cellPtr(g), cellBytes());
  }

  /// The host page of the leaf whose origin is `key`, paged in from the device
  /// (or created) as needed; nullptr if absent and not created.
  LeafGrid* page(const CoordT& key, bool create) const {
    State& s = *st_;
    std::lock_guard<std::recursive_mutex> lk(s.mu);
    auto it = s.pages.find(key);
    if (it != s.pages.end()) {
      if (it->second || !create) return it->second.get();
      it->second.reset(new LeafGrid(s.leaf_bits));
      ++s.gen;
      return it->second.get();
    }
    if (s.complete) {
      if (!create) return nullptr;
      LeafGrid* g = new LeafGrid(s.leaf_bits);
      s.pages.emplace(key, std::unique_ptr<LeafGrid>(g));
      ++s.gen;
      return g;
    }
    std::unique_ptr<LeafGrid> g(new LeafGrid(s.leaf_bits));
    uint8_t exists = 0;
    gpu::check(bxg_read_leaves_host(s.map, &key.x, 1, g->mask_.data(), cellPtr(g.get()), &exists), s.ctx);
    if (exists) {
      g->on_device_ = true;
      snapshot(g.get());
    } else if (!create) {
      s.pages.emplace(key, nullptr);
      return nullptr;
    }
    LeafGrid* r = g.get();
    s.pages.emplace(key, std::move(g));
    ++s.gen;
    return r;
  }

  void flushLocked() const {
    State& s = *st_;
    std::vector<int32_t> origins;
    std::vector<uint64_t> masks;
    std::vector<uint8_t> cells;
    std::vector<LeafGrid*> out;
    for (auto& kv : s.pages) {
      LeafGrid* g = kv.second.get();
      if (!g || (g->on_device_ && !changed(g))) continue;
      origins.insert(origins.end(), {kv.first.x, kv.first.y, kv.first.z});
      masks.insert(masks.end(), g->mask_.data(), g->mask_.data() + s.leaf_words);
      if (cellBytes()) cells.insert(cells.end(), cellPtr(g), cellPtr(g) + cellBytes());
      out.push_back(g);
    }
    if (out.empty()) return;
    gpu::check(bxg_write_leaves_host(s.map, origins.data(), out.size(), masks.data(),
                                     cellBytes() ? cells.data() : nullptr),
               s.ctx);
    for (LeafGrid* g : out) {
      g->on_device_ = true;
      snapshot(g);
    }
  }

  // whether the box of the leaf at `key` comes within `radius` of `c`
  bool leafNear(const CoordT& key, const Point3D& c, double radius) const {
    const double res = st_->res, side = double(1u << st_->leaf_bits) * res;
    const double lo[3] = {double(key.x) * res, double(key.y) * res, double(key.z) * res};
    const double p[3] = {c.x, c.y, c.z};
    double d2 = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double below = lo[a] - p[a], above = p[a] - (lo[a] + side);
      const double d = below > 0 ? below : (above > 0 ? above : 0.0);
      d2 += d * d;
    }
    return !(d2 > radius * radius);  // NaN or infinite radius: near
  }

  void refreshLocked(bool may_create, const Point3D* center, double radius) const {
    State& s = *st_;
    if (may_create) s.complete = false;
    std::vector<CoordT> keys;
    for (auto it = s.pages.begin(); it != s.pages.end();) {
      if (!it->second) {
        it = s.pages.erase(it);
      } else {
        if (!center || leafNear(it->first, *center, radius)) keys.push_back(it->first);
        ++it;
      }
    }
    ++s.gen;
    if (keys.empty()) return;
    std::vector<uint64_t> masks(keys.size() * s.leaf_words);
    std::vector<uint8_t> cells(std::max<size_t>(keys.size() * cellBytes(), 1));
    std::vector<uint8_t> exists(keys.size());
    gpu::check(bxg_read_leaves_host(s.map, &keys[0].x, keys.size(), masks.data(),
                                    cellBytes() ? cells.data() : nullptr, exists.data()),
               s.ctx);
    for (size_t k = 0; k < keys.size(); ++k) {
      auto it = s.pages.find(keys[k]);
      if (!exists[k]) {  // freed on the device (releaseUnusedMemory)
        s.pages.erase(it);
        continue;
      }
      LeafGrid* g = it->second.get();
      std::memcpy(g->mask_.data(), &masks[k * s.leaf_words], maskBytes());
      if (cellBytes()) std::memcpy(cellPtr(g), &cells[k * cellBytes()], cellBytes());
      g->on_device_ = true;
      snapshot(g);
    }
  }

  template <class F>
  void visit(F& func, bool write_back) const {
    State& s = *st_;
    std::lock_guard<std::recursive_mutex> lk(s.mu);
    flushLocked();
    uint64_t nl = 0;
    int rc = bxg_leaf_origins_host(s.map, nullptr, 0, &nl);
    if (rc != BXG_OK && rc != BXG_ERR_CAPACITY) gpu::check(rc, s.ctx);
    if (!nl) return;
    std::vector<CoordT> origins(nl);
    gpu::check(bxg_leaf_origins_host(s.map, &origins[0].x, nl, &nl), s.ctx);
    const uint32_t lb = s.leaf_bits, lm = (1u << lb) - 1;
    const size_t chunk = 1 << 14;
    std::vector<uint64_t> masks;
    std::vector<uint8_t> cells, before, exists;
    std::vector<CoordT> dirty_keys;
    std::vector<uint64_t> dirty_masks;
    std::vector<uint8_t> dirty_cells;
    for (size_t b = 0; b < origins.size(); b += chunk) {
      const size_t n = std::min(chunk, origins.size() - b);
      masks.assign(n * s.leaf_words, 0);
      cells.assign(std::max<size_t>(n * cellBytes(), 1), 0);
      exists.assign(n, 0);
      gpu::check(bxg_read_leaves_host(s.map, &origins[b].x, n, masks.data(),
                                      cellBytes() ? cells.data() : nullptr, exists.data()),
                 s.ctx);
      if (write_back) before = cells;
      dirty_keys.clear();
      dirty_masks.clear();
      dirty_cells.clear();
      for (size_t k = 0; k < n; ++k) {
        const CoordT o = origins[b + k];
        // a cached page is the host's copy of record for its leaf
        auto it = s.pages.find(o);
        LeafGrid* g = it != s.pages.end() ? it->second.get() : nullptr;
        const uint64_t* mw = g ? g->mask_.data() : &masks[k * s.leaf_words];
        DataT* vals = nullptr;
        if constexpr (!std::is_same_v<DataT, EmptyVoxel>) {
          vals = g ? g->data_.get() : reinterpret_cast<DataT*>(&cells[k * cellBytes()]);
        }
        for (uint32_t w = 0; w < s.leaf_words; ++w) {
          uint64_t bits = mw[w];
          while (bits) {
            const uint32_t li = (w << 6) + uint32_t(__builtin_ctzll(bits));
            bits &= bits - 1;
            if (li >= s.leaf_size) break;
            const CoordT c = {o.x | int32_t(li & lm), o.y | int32_t((li >> lb) & lm),
                              o.z | int32_t((li >> (2 * lb)) & lm)};
            if constexpr (std::is_same_v<DataT, EmptyVoxel>) {
              EmptyVoxel dummy{};
              func(dummy, c);
            } else {
              func(vals[li], c);
            }
          }
        }
        if (write_back && !g && cellBytes() &&
            std::memcmp(&cells[k * cellBytes()], &before[k * cellBytes()], cellBytes())) {
          dirty_keys.push_back(o);
          dirty_masks.insert(dirty_masks.end(), &masks[k * s.leaf_words], &masks[(k + 1) * s.leaf_words]);
          dirty_cells.insert(dirty_cells.end(), &cells[k * cellBytes()], &cells[(k + 1) * cellBytes()]);
        }
      }
      if (!dirty_keys.empty())
        gpu::check(bxg_write_leaves_host(s.map, &dirty_keys[0].x, dirty_keys.size(), dirty_masks.data(),
                                         dirty_cells.data()),
                   s.ctx);
    }
  }

  std::unique_ptr<State> st_;
};

/// bonxai.hpp:335
using BinaryVoxelGrid = VoxelGrid<EmptyVoxel>;

}  // namespace Bonxai
