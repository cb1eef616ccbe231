// ref_grid_driver.cpp -- a C entry layer over the reference's own VoxelGrid,
// compiled from the reference headers where they lie (see oracle/Makefile).
//
// TEST INFRASTRUCTURE.  It is built only in the development container (the
// reference tree is not present elsewhere) into oracle/_ref/, and used to
// check the CPU restatement in bx_oracle.c and to write golden fixtures.
//
// Reference code used, unmodified, through its public API:
//   Bonxai::VoxelGrid<T>, Accessor, ConstAccessor, forEachCell
//       bonxai_core/include/bonxai/bonxai.hpp:114-333
//   Bonxai::Serialize / GetHeaderInfo / Deserialize
//       bonxai_core/include/bonxai/serialization.hpp:77-199
//
// bonxai_map/include/bonxai_map/probabilistic_map.hpp cannot be compiled in
// this image: it includes <eigen3/Eigen/Geometry> and Eigen3 is not
// installed.  ref_prob_* below therefore runs the scan-integration steps of
// that file as restated here (the same statements as bx_oracle.c) on top of
// the reference's VoxelGrid<Cell> and Accessor, which pins the storage and
// accessor semantics the restatement relies on, not the integration rules.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "bonxai/bonxai.hpp"
#include "bonxai/serialization.hpp"

namespace {

// Same layout as the reference's ProbabilisticMap::CellT
// (probabilistic_map.hpp:44-53): a signed 4-bit update id in the low bits and
// a signed 28-bit log-odds above it.
struct Cell {
  int32_t update_id : 4;
  int32_t probability_log : 28;
  Cell() : update_id(0), probability_log(0) {}
};

struct GridBase {
  virtual ~GridBase() = default;
  virtual void set_values(const int32_t* c, const void* v, uint64_t n, uint8_t* was_on) = 0;
  virtual void set_cells_on(const int32_t* c, const void* v, uint64_t n, uint8_t* was_on) = 0;
  virtual void set_cells_off(const int32_t* c, uint64_t n, uint8_t* was_on) = 0;
  virtual void lookup(const int32_t* c, uint64_t n, void* out, uint8_t* found) = 0;
  virtual uint64_t export_cells(int32_t* c, void* v, uint64_t cap) = 0;
  virtual uint64_t active_cells() = 0;
  virtual uint64_t serialize(char* out, uint64_t cap) = 0;
  virtual int deserialize(const char* in, uint64_t n) = 0;
  virtual void clear(int opt) = 0;
  virtual void release_unused() = 0;
  virtual uint32_t value_bytes() = 0;
};

template <typename T>
struct Grid final : GridBase {
  Bonxai::VoxelGrid<T> g;
  Grid(double res, int ib, int lb) : g(res, ib, lb) {}

  static constexpr bool kEmpty = std::is_same_v<T, Bonxai::EmptyVoxel>;

  uint32_t value_bytes() override { return kEmpty ? 0 : sizeof(T); }

  void set_values(const int32_t* c, const void* v, uint64_t n, uint8_t* was_on) override {
    auto acc = g.createAccessor();
    for (uint64_t i = 0; i < n; ++i) {
      Bonxai::CoordT k{c[3 * i], c[3 * i + 1], c[3 * i + 2]};
      bool w;
      if constexpr (kEmpty) {
        w = acc.setCellOn(k);
      } else {
        T val;
        std::memcpy(&val, static_cast<const char*>(v) + i * sizeof(T), sizeof(T));
        w = acc.setValue(k, val);
      }
      if (was_on) was_on[i] = w;
    }
  }
  void set_cells_on(const int32_t* c, const void* v, uint64_t n, uint8_t* was_on) override {
    auto acc = g.createAccessor();
    for (uint64_t i = 0; i < n; ++i) {
      Bonxai::CoordT k{c[3 * i], c[3 * i + 1], c[3 * i + 2]};
      bool w;
      if constexpr (kEmpty) {
        w = acc.setCellOn(k);
      } else {
        T val;
        std::memcpy(&val, static_cast<const char*>(v) + i * sizeof(T), sizeof(T));
        w = acc.setCellOn(k, val);
      }
      if (was_on) was_on[i] = w;
    }
  }
  void set_cells_off(const int32_t* c, uint64_t n, uint8_t* was_on) override {
    auto acc = g.createAccessor();
    for (uint64_t i = 0; i < n; ++i) {
      Bonxai::CoordT k{c[3 * i], c[3 * i + 1], c[3 * i + 2]};
      bool w = acc.setCellOff(k);
      if (was_on) was_on[i] = w;
    }
  }
  void lookup(const int32_t* c, uint64_t n, void* out, uint8_t* found) override {
    auto acc = g.createConstAccessor();
    for (uint64_t i = 0; i < n; ++i) {
      Bonxai::CoordT k{c[3 * i], c[3 * i + 1], c[3 * i + 2]};
      if constexpr (kEmpty) {
        found[i] = acc.isCellOn(k);
      } else {
        const T* p = acc.value(k);
        found[i] = p != nullptr;
        if (out) {
          char* o = static_cast<char*>(out) + i * sizeof(T);
          if (p)
            std::memcpy(o, p, sizeof(T));
          else
            std::memset(o, 0, sizeof(T));
        }
      }
    }
  }
  uint64_t export_cells(int32_t* c, void* v, uint64_t cap) override {
    uint64_t k = 0;
    g.forEachCell([&](const auto& val, const Bonxai::CoordT& coord) {
      if (k < cap) {
        if (c) {
          c[3 * k] = coord.x;
          c[3 * k + 1] = coord.y;
          c[3 * k + 2] = coord.z;
        }
        if constexpr (!kEmpty) {
          if (v) std::memcpy(static_cast<char*>(v) + k * sizeof(T), &val, sizeof(T));
        }
      }
      ++k;
    });
    return k;
  }
  uint64_t active_cells() override { return g.activeCellsCount(); }
  uint64_t serialize(char* out, uint64_t cap) override {
    if constexpr (kEmpty) {
      return 0;
    } else {
      std::ostringstream os(std::ios::binary);
      Bonxai::Serialize(os, g);
      const std::string s = os.str();
      if (out) std::memcpy(out, s.data(), s.size() < cap ? s.size() : cap);
      return s.size();
    }
  }
  int deserialize(const char* in, uint64_t n) override {
    if constexpr (kEmpty) {
      return -1;
    } else {
      std::istringstream is(std::string(in, n), std::ios::binary);
      char header[256];
      is.getline(header, 256);
      auto info = Bonxai::GetHeaderInfo(header);
      g = Bonxai::Deserialize<T>(is, info);
      return 0;
    }
  }
  void clear(int opt) override {
    g.clear(opt == 0 ? Bonxai::CLEAR_MEMORY : Bonxai::SET_ALL_CELLS_OFF);
  }
  void release_unused() override { g.releaseUnusedMemory(); }
};

}  // namespace

extern "C" {

// kind: 0 float, 1 int32, 2 uint32, 3 int64, 4 EmptyVoxel (binary), 5 Cell
__attribute__((visibility("default"))) void* ref_grid_create(int kind, double res, int ib, int lb) {
  switch (kind) {
    case 0: return new Grid<float>(res, ib, lb);
    case 1: return new Grid<int32_t>(res, ib, lb);
    case 2: return new Grid<uint32_t>(res, ib, lb);
    case 3: return new Grid<int64_t>(res, ib, lb);
    case 4: return new Grid<Bonxai::EmptyVoxel>(res, ib, lb);
    case 5: return new Grid<Cell>(res, ib, lb);
    default: return nullptr;
  }
}
__attribute__((visibility("default"))) void ref_grid_destroy(void* g) {
  delete static_cast<GridBase*>(g);
}
__attribute__((visibility("default"))) void ref_set_values(void* g, const int32_t* c, const void* v,
                                                           uint64_t n, uint8_t* was_on) {
  static_cast<GridBase*>(g)->set_values(c, v, n, was_on);
}
__attribute__((visibility("default"))) void ref_set_cells_on(void* g, const int32_t* c,
                                                             const void* v, uint64_t n,
                                                             uint8_t* was_on) {
  static_cast<GridBase*>(g)->set_cells_on(c, v, n, was_on);
}
__attribute__((visibility("default"))) void ref_set_cells_off(void* g, const int32_t* c,
                                                              uint64_t n, uint8_t* was_on) {
  static_cast<GridBase*>(g)->set_cells_off(c, n, was_on);
}
__attribute__((visibility("default"))) void ref_lookup(void* g, const int32_t* c, uint64_t n,
                                                       void* out, uint8_t* found) {
  static_cast<GridBase*>(g)->lookup(c, n, out, found);
}
__attribute__((visibility("default"))) uint64_t ref_export(void* g, int32_t* c, void* v,
                                                           uint64_t cap) {
  return static_cast<GridBase*>(g)->export_cells(c, v, cap);
}
__attribute__((visibility("default"))) uint64_t ref_active_cells(void* g) {
  return static_cast<GridBase*>(g)->active_cells();
}
__attribute__((visibility("default"))) uint64_t ref_serialize(void* g, char* out, uint64_t cap) {
  return static_cast<GridBase*>(g)->serialize(out, cap);
}
__attribute__((visibility("default"))) int ref_deserialize(void* g, const char* in, uint64_t n) {
  return static_cast<GridBase*>(g)->deserialize(in, n);
}
__attribute__((visibility("default"))) void ref_clear(void* g, int opt) {
  static_cast<GridBase*>(g)->clear(opt);
}
__attribute__((visibility("default"))) void ref_release_unused(void* g) {
  static_cast<GridBase*>(g)->release_unused();
}
__attribute__((visibility("default"))) void ref_pos_to_coord(double res, const double* xyz,
                                                             uint64_t n, int32_t* out) {
  Bonxai::VoxelGrid<float> g(res);
  for (uint64_t i = 0; i < n; ++i) {
    const auto c = g.posToCoord(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    out[3 * i] = c.x;
    out[3 * i + 1] = c.y;
    out[3 * i + 2] = c.z;
  }
}
__attribute__((visibility("default"))) void ref_coord_hash(const int32_t* c, uint64_t n,
                                                           uint64_t* out) {
  std::hash<Bonxai::CoordT> h;
  for (uint64_t i = 0; i < n; ++i) out[i] = h(Bonxai::CoordT{c[3 * i], c[3 * i + 1], c[3 * i + 2]});
}

// ---- scan integration over the reference's VoxelGrid<Cell> ----------------

struct RefProb {
  Bonxai::VoxelGrid<Cell> grid;
  Bonxai::VoxelGrid<Cell>::Accessor acc;
  int32_t miss = -405465, hit = 847297, cmin = -1992430, cmax = 3476099;
  uint8_t cur = 1;
  std::vector<Bonxai::CoordT> hits, misses;
  explicit RefProb(double res) : grid(res), acc(grid.createAccessor()) {}
};

__attribute__((visibility("default"))) void* ref_prob_create(double res) { return new RefProb(res); }
__attribute__((visibility("default"))) void ref_prob_destroy(void* m) {
  delete static_cast<RefProb*>(m);
}

static void ref_endpoint(RefProb* m, const Bonxai::CoordT& c, bool is_hit) {
  Cell* cell = m->acc.value(c, true);
  if (cell->update_id != m->cur) {
    if (is_hit)
      cell->probability_log = std::min<int32_t>(cell->probability_log + m->hit, m->cmax);
    else
      cell->probability_log = std::max<int32_t>(cell->probability_log + m->miss, m->cmin);
    cell->update_id = m->cur;
    (is_hit ? m->hits : m->misses).push_back(c);
  }
}

__attribute__((visibility("default"))) void ref_prob_insert_f32(void* mp, const float* pts,
                                                                uint64_t n, const double origin[3],
                                                                double max_range) {
  RefProb* m = static_cast<RefProb*>(mp);
  const double fx = origin[0], fy = origin[1], fz = origin[2];
  const double r2 = max_range * max_range;
  for (uint64_t i = 0; i < n; ++i) {
    const double tx = pts[3 * i], ty = pts[3 * i + 1], tz = pts[3 * i + 2];
    const double vx = tx - fx, vy = ty - fy, vz = tz - fz;
    const double sq = (vx * vx + vy * vy) + vz * vz;
    if (sq >= r2) {
      const double s = std::sqrt(sq);
      ref_endpoint(m,
                   m->grid.posToCoord(fx + ((vx / s) * max_range), fy + ((vy / s) * max_range),
                                      fz + ((vz / s) * max_range)),
                   false);
    } else {
      ref_endpoint(m, m->grid.posToCoord(tx, ty, tz), true);
    }
  }
  auto acc = m->grid.createAccessor();
  const auto o = m->grid.posToCoord(fx, fy, fz);
  auto clear_point = [&](const Bonxai::CoordT& c) {
    Cell* cell = acc.value(c, true);
    if (cell->update_id != m->cur) {
      cell->probability_log = std::max<int32_t>(cell->probability_log + m->miss, m->cmin);
      cell->update_id = m->cur;
    }
  };
  auto ray = [&](const Bonxai::CoordT& e) {
    if (o == e) return;
    clear_point(o);
    Bonxai::CoordT d = e - o;
    const Bonxai::CoordT st{d.x < 0 ? -1 : 1, d.y < 0 ? -1 : 1, d.z < 0 ? -1 : 1};
    d = {d.x < 0 ? -d.x : d.x, d.y < 0 ? -d.y : d.y, d.z < 0 ? -d.z : d.z};
    const int mx = std::max(std::max(d.x, d.y), d.z);
    Bonxai::CoordT err{0, 0, 0}, c = o;
    for (int i = 0; i < mx - 1; ++i) {
      err = err + d;
      if (2 * err.x >= mx) { c.x += st.x; err.x -= mx; }
      if (2 * err.y >= mx) { c.y += st.y; err.y -= mx; }
      if (2 * err.z >= mx) { c.z += st.z; err.z -= mx; }
      clear_point(c);
    }
  };
  for (const auto& e : m->hits) ray(e);
  m->hits.clear();
  for (const auto& e : m->misses) ray(e);
  m->misses.clear();
  if (++m->cur == 4) m->cur = 1;
}

// export of the Cell grid as (x, y, z, raw 32-bit word); returns the count
__attribute__((visibility("default"))) uint64_t ref_prob_export(void* mp, int32_t* c,
                                                                uint32_t* words, uint64_t cap) {
  RefProb* m = static_cast<RefProb*>(mp);
  uint64_t k = 0;
  m->grid.forEachCell([&](const Cell& cell, const Bonxai::CoordT& coord) {
    if (k < cap) {
      c[3 * k] = coord.x;
      c[3 * k + 1] = coord.y;
      c[3 * k + 2] = coord.z;
      std::memcpy(&words[k], &cell, 4);
    }
    ++k;
  });
  return k;
}
__attribute__((visibility("default"))) uint32_t ref_prob_update_count(void* mp) {
  return static_cast<RefProb*>(mp)->cur;
}

}  // extern "C"
