// VoxelGrid<float> through the drop-in C++ API, the way reference callers use
// it: Accessor::setValue / value / value(create) with writes through the
// returned pointer / setCellOn / setCellOff, ConstAccessor::isCellOn, the
// non-const forEachCell, clear(), releaseUnusedMemory(), a Serialize /
// Deserialize round trip, plus the GPU batch calls.  Replays a script and
// dumps what every call returned.
//
// usage: dropin_grid script.bin result.bin
//
// script: double voxel_size, uint32 inner_bits, uint32 leaf_bits, uint32 op
// count, then per op a uint32 kind and its payload:
//   0 setValue        uint32 n, int32 xyz[3n], float v[n]   -> int8 was_on[n]
//   1 value           uint32 n, int32 xyz[3n]               -> int8 found, float value (0 if none)
//   2 value(create)   uint32 n, int32 xyz[3n], float v[n]   -> float value before the write; then *p = v
//   3 setCellOn       uint32 n, int32 xyz[3n], float d[n]   -> int8 was_on[n]
//   4 setCellOff      uint32 n, int32 xyz[3n]               -> int8 was_on[n]
//   5 setValues (GPU batch)  uint32 n, int32 xyz[3n], float v[n] -> int8 was_on[n]
//   6 forEachCell: v = 2 v + 1 on every on cell
//   7 clear(SET_ALL_CELLS_OFF)
//   8 releaseUnusedMemory
//   9 clear(CLEAR_MEMORY)
//  10 Serialize then Deserialize into the same variable
//  11 isCellOn (const accessor)  uint32 n, int32 xyz[3n]   -> int8 on[n]
// result: the per-op outputs above in order, then uint64 n + n x (int32 xyz,
// float v) from the const forEachCell and uint64 activeCellsCount.
#include <cstdint>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "bonxai/bonxai.hpp"
#include "bonxai/serialization.hpp"

namespace {

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
  std::vector<T> vec(size_t n) {
    std::vector<T> v(n);
    in.read(reinterpret_cast<char*>(v.data()), sizeof(T) * n);
    if (!in) throw std::runtime_error("truncated script");
    return v;
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

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    std::cerr << "usage: dropin_grid script.bin result.bin" << std::endl;
    return 2;
  }
  try {
    Reader r(argv[1]);
    Writer w(argv[2]);
    const double vs = r.get<double>();
    const uint32_t ib = r.get<uint32_t>(), lb = r.get<uint32_t>();
    Bonxai::VoxelGrid<float> grid(vs, uint8_t(ib), uint8_t(lb));
    const uint32_t nops = r.get<uint32_t>();
    for (uint32_t op = 0; op < nops; ++op) {
      const uint32_t kind = r.get<uint32_t>();
      std::vector<Bonxai::CoordT> c;
      std::vector<float> v;
      auto coords = [&]() {
        const uint32_t n = r.get<uint32_t>();
        c = r.vec<Bonxai::CoordT>(n);
        return n;
      };
      // a fresh accessor per op: accessors do not survive clear() or a move
      auto acc = grid.createAccessor();
      switch (kind) {
        case 0: {
          const uint32_t n = coords();
          v = r.vec<float>(n);
          for (uint32_t i = 0; i < n; ++i) w.put(int8_t(acc.setValue(c[i], v[i])));
          break;
        }
        case 1: {
          const uint32_t n = coords();
          for (uint32_t i = 0; i < n; ++i) {
            const float* p = acc.value(c[i]);
            w.put(int8_t(p != nullptr));
            w.put(p ? *p : 0.0f);
          }
          break;
        }
        case 2: {
          const uint32_t n = coords();
          v = r.vec<float>(n);
          for (uint32_t i = 0; i < n; ++i) {
            float* p = acc.value(c[i], true);
            w.put(*p);
            *p = v[i];
          }
          break;
        }
        case 3: {
          const uint32_t n = coords();
          v = r.vec<float>(n);
          for (uint32_t i = 0; i < n; ++i) w.put(int8_t(acc.setCellOn(c[i], v[i])));
          break;
        }
        case 4: {
          const uint32_t n = coords();
          for (uint32_t i = 0; i < n; ++i) w.put(int8_t(acc.setCellOff(c[i])));
          break;
        }
        case 5: {
          const uint32_t n = coords();
          v = r.vec<float>(n);
          std::vector<uint8_t> was;
          grid.setValues(c, v, &was);
          for (uint32_t i = 0; i < n; ++i) w.put(int8_t(was[i]));
          break;
        }
        case 6:
          grid.forEachCell([](float& x, const Bonxai::CoordT&) { x = 2.0f * x + 1.0f; });
          break;
        case 7:
          grid.clear(Bonxai::SET_ALL_CELLS_OFF);
          break;
        case 8:
          grid.releaseUnusedMemory();
          break;
        case 9:
          grid.clear(Bonxai::CLEAR_MEMORY);
          break;
        case 10: {
          std::stringstream ss(std::ios::in | std::ios::out | std::ios::binary);
          Bonxai::Serialize(ss, grid);
          char header[256];
          ss.getline(header, sizeof(header));
          grid = Bonxai::Deserialize<float>(ss, Bonxai::GetHeaderInfo(header));
          break;
        }
        case 11: {
          const uint32_t n = coords();
          const auto& cg = grid;
          auto cacc = cg.createConstAccessor();
          for (uint32_t i = 0; i < n; ++i) w.put(int8_t(cacc.isCellOn(c[i])));
          break;
        }
        default:
          throw std::runtime_error("unknown op");
      }
    }
    std::vector<std::pair<Bonxai::CoordT, float>> cells;
    const auto& cg = grid;
    cg.forEachCell([&](const float& x, const Bonxai::CoordT& xyz) { cells.push_back({xyz, x}); });
    w.put(uint64_t(cells.size()));
    for (const auto& e : cells) {
      w.put(e.first);
      w.put(e.second);
    }
    w.put(uint64_t(grid.activeCellsCount()));
    std::cout << "ops " << nops << " cells " << cells.size() << " mem " << grid.memUsage() << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
