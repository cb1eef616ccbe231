// Scan integration through the drop-in C++ API: Bonxai::ProbabilisticMap driven
// the way a reference caller drives it (insertPointCloud with float or double
// point types, addHitPoint / addMissPoint, isOccupied / isFree / isUnknown,
// getOccupiedVoxels / getFreeVoxels, accessors on grid()), replaying a script
// and dumping the results for comparison with an independent model.
//
// usage: dropin_probmap script.bin result.bin
//
// script (little-endian):  double resolution, uint32 op count, then per op a
// uint32 kind and its payload:
//   0 insert, float points   uint32 n, double origin[3], double max_range, float xyz[3n]
//   1 insert, double points  uint32 n, double origin[3], double max_range, double xyz[3n]
//   2 addHitPoint            double xyz[3]
//   3 addMissPoint           double xyz[3]
//   4 query                  uint32 n, int32 xyz[3n]
//   5 accessor write         uint32 n, int32 xyz[3n], int32 word[n]  (value(coord, true) = word)
//   6 insert, sensor frame   uint32 n, float sensor_to_world[16] (row-major), double max_range,
//                            float xyzw[4n] (a pcl::PointXYZ-like point: x y z pad)
//   7 checkpoint and resume  (no payload) saveCheckpoint to a file beside result.bin, check
//                            that it opens with the bytes of Serialize(grid()), then replace
//                            the map by a new one restored with loadCheckpoint from that file
// result: per query op, per coordinate: int8 isOccupied, isFree, isUnknown,
// classify state, int64 cell word through a const accessor (-1 if off); then
// uint64 n + n x (int32 xyz, int32 word) from forEachCell, uint64 n + n x int32
// xyz from getOccupiedVoxels, the same from getFreeVoxels, uint64 n + n x double
// xyz from getOccupiedVoxels<Point3D>, uint64 activeCellsCount, uint32 update count.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

#include <memory>
#include <sstream>
#include <string>

#include "bonxai_map/probabilistic_map.hpp"

namespace {

struct FloatPoint {  // a float point type with x, y, z members
  float x, y, z;
};
struct PaddedPoint {  // pcl::PointXYZ's layout: x y z and a pad word
  float x, y, z, pad;
};

struct Reader {
  std::ifstream in;
  explicit Reader(const char* path) : in(path, std::ios::binary) {
    if (!in) throw std::runtime_error("cannot open script");
  }
  template <typename T>
  T get() {
    T v{};
    in.read(reinterpret_cast<char*>(&v), sizeof(T));
    if (!in) throw std::runtime_error("truncated script");
    return v;
  }
  template <typename T>
  void get(T* p, size_t n) {
    in.read(reinterpret_cast<char*>(p), sizeof(T) * n);
    if (!in) throw std::runtime_error("truncated script");
  }
};

struct Writer {
  std::ofstream out;
  explicit Writer(const char* path) : out(path, std::ios::binary) {}
  template <typename T>
  void put(const T& v) {
    out.write(reinterpret_cast<const char*>(&v), sizeof(T));
  }
};

int32_t word_of(const Bonxai::ProbabilisticMap::CellT& c) {
  int32_t w;
  std::memcpy(&w, &c, 4);
  return w;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::cerr << "usage: dropin_probmap script.bin result.bin" << std::endl;
    return 2;
  }
  try {
    Reader r(argv[1]);
    Writer w(argv[2]);
    const double res = r.get<double>();
    // a pointer, so that a checkpoint op can replace the map by its restored copy
    auto owner = std::make_unique<Bonxai::ProbabilisticMap>(res);
    Bonxai::ProbabilisticMap* map = owner.get();
    using CellT = Bonxai::ProbabilisticMap::CellT;
    const uint32_t nops = r.get<uint32_t>();
    for (uint32_t op = 0; op < nops; ++op) {
      const uint32_t kind = r.get<uint32_t>();
      if (kind == 0 || kind == 1) {
        const uint32_t n = r.get<uint32_t>();
        double o[3];
        r.get(o, 3);
        const double range = r.get<double>();
        if (kind == 0) {
          std::vector<FloatPoint> pts(n);
          r.get(reinterpret_cast<float*>(pts.data()), 3 * size_t(n));
          const FloatPoint origin{float(o[0]), float(o[1]), float(o[2])};
          if (double(origin.x) != o[0] || double(origin.y) != o[1] || double(origin.z) != o[2])
            throw std::runtime_error("float-point scans need a float origin");
          map->insertPointCloud(pts, origin, range);
        } else {
          std::vector<Bonxai::Point3D> pts(n);
          for (auto& p : pts) {
            p.x = r.get<double>();
            p.y = r.get<double>();
            p.z = r.get<double>();
          }
          map->insertPointCloud(pts, Bonxai::Point3D(o[0], o[1], o[2]), range);
        }
      } else if (kind == 2 || kind == 3) {
        double p[3];
        r.get(p, 3);
        const Bonxai::ProbabilisticMap::Vector3D v(p[0], p[1], p[2]);
        kind == 2 ? map->addHitPoint(v) : map->addMissPoint(v);
      } else if (kind == 4) {
        const uint32_t n = r.get<uint32_t>();
        std::vector<Bonxai::CoordT> q(n);
        r.get(reinterpret_cast<int32_t*>(q.data()), 3 * size_t(n));
        std::vector<int8_t> state;
        map->classify(q, state);
        const auto& grid = map->grid();
        auto acc = grid.createConstAccessor();
        for (uint32_t i = 0; i < n; ++i) {
          w.put(int8_t(map->isOccupied(q[i])));
          w.put(int8_t(map->isFree(q[i])));
          w.put(int8_t(map->isUnknown(q[i])));
          w.put(state[i]);
          const CellT* c = acc.value(q[i]);
          w.put(int64_t(c ? word_of(*c) : -1));
        }
      } else if (kind == 5) {
        const uint32_t n = r.get<uint32_t>();
        std::vector<Bonxai::CoordT> q(n);
        r.get(reinterpret_cast<int32_t*>(q.data()), 3 * size_t(n));
        std::vector<int32_t> words(n);
        r.get(words.data(), n);
        auto acc = map->grid().createAccessor();
        for (uint32_t i = 0; i < n; ++i) {
          CellT* c = acc.value(q[i], true);
          c->update_id = int32_t(uint32_t(words[i]) << 28) >> 28;
          c->probability_log = words[i] >> 4;
        }
      } else if (kind == 6) {
        const uint32_t n = r.get<uint32_t>();
        float tf[16];
        r.get(tf, 16);
        const double range = r.get<double>();
        std::vector<PaddedPoint> pts(n);
        r.get(reinterpret_cast<float*>(pts.data()), 4 * size_t(n));
        map->insertPointCloudTransformed(pts, tf, range);
      } else if (kind == 7) {
        const std::string path = std::string(argv[2]) + ".checkpoint";
        {
          std::ofstream f(path, std::ios::binary);
          map->saveCheckpoint(f);
        }
        std::ifstream f(path, std::ios::binary);
        const std::string bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        std::ostringstream grid_only;
        Bonxai::Serialize(grid_only, map->grid());
        const std::string g = grid_only.str();
        if (bytes.size() <= g.size() || bytes.compare(0, g.size(), g) != 0)
          throw std::runtime_error("the checkpoint does not open with the grid stream");
        const std::string trailer = bytes.substr(g.size());
        if (trailer.rfind("Bonxai::ProbabilisticMap(update_count=", 0) != 0 || trailer.back() != '\n')
          throw std::runtime_error("unexpected checkpoint trailer");
        auto restored = std::make_unique<Bonxai::ProbabilisticMap>(res);
        std::istringstream in(bytes);
        restored->loadCheckpoint(in);
        owner = std::move(restored);  // the saved map is gone
        map = owner.get();
        std::cout << "checkpoint " << bytes.size() << " bytes, update_count " << map->updateCount() << std::endl;
      } else {
        throw std::runtime_error("unknown op");
      }
    }
    std::vector<std::pair<Bonxai::CoordT, int32_t>> cells;
    map->grid().forEachCell([&](const CellT& c, const Bonxai::CoordT& xyz) { cells.push_back({xyz, word_of(c)}); });
    w.put(uint64_t(cells.size()));
    for (const auto& c : cells) {
      w.put(c.first);
      w.put(c.second);
    }
    std::vector<Bonxai::CoordT> occ, fre;
    map->getOccupiedVoxels(occ);
    map->getFreeVoxels(fre);
    for (auto* v : {&occ, &fre}) {
      w.put(uint64_t(v->size()));
      for (const auto& c : *v) w.put(c);
    }
    std::vector<Bonxai::Point3D> pts;
    map->getOccupiedVoxels(pts);
    w.put(uint64_t(pts.size()));
    for (const auto& p : pts) {
      w.put(p.x);
      w.put(p.y);
      w.put(p.z);
    }
    w.put(uint64_t(map->grid().activeCellsCount()));
    w.put(map->updateCount());
    std::cout << "ops " << nops << " cells " << cells.size() << " occupied " << occ.size() << " free "
              << fre.size() << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
