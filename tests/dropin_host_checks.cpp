// Host-only checks of the drop-in headers (no GPU call is made): prints the
// values the Python side compares with the oracle and with closed forms.
//
// usage: dropin_host_checks rays.bin
//   rays.bin: uint32 n, then n x (int32 origin[3], int32 end[3])
// prints one line per check; for each ray "ray k: x y z x y z ..." (cells of
// Bonxai::ComputeRay).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "bonxai/bonxai.hpp"
#include "bonxai/serialization.hpp"
#include "bonxai_map/probabilistic_map.hpp"

struct MemberPoint {
  float x, y, z;
};
struct MethodPoint {
  double v[3];
  double x() const { return v[0]; }
  double y() const { return v[1]; }
  double z() const { return v[2]; }
};

int main(int argc, char** argv) {
  using PM = Bonxai::ProbabilisticMap;
  std::cout << "logods " << PM::logods(0.4f) << " " << PM::logods(0.7f) << " " << PM::logods(0.12f) << " "
            << PM::logods(0.97f) << " " << PM::logods(0.5f) << std::endl;
  const PM::Options o;
  std::cout << "options " << o.prob_miss_log << " " << o.prob_hit_log << " " << o.clamp_min_log << " "
            << o.clamp_max_log << " " << o.occupancy_threshold_log << std::endl;
  PM::CellT c;
  c.update_id = -3;
  c.probability_log = -123456;
  uint32_t w;
  std::memcpy(&w, &c, 4);
  std::cout << "cell " << w << std::endl;

  const auto h = Bonxai::GetHeaderInfo("Bonxai::VoxelGrid<Bonxai::ProbabilisticMap::CellT,2,3>(0.050000)\n");
  std::cout << "header " << h.type_name << " " << h.inner_bits << " " << h.leaf_bits << " " << h.resolution
            << std::endl;
  const auto h2 = Bonxai::GetHeaderInfo("Bonxai::VoxelGrid<int,1,4>(0.100000)");
  std::cout << "header " << h2.type_name << " " << h2.inner_bits << " " << h2.leaf_bits << " " << h2.resolution
            << std::endl;

  Bonxai::Mask m(3);
  m.setOn(0);
  m.setOn(63);
  m.setOn(64);
  m.setOn(511);
  const bool was = m.setOn(64);
  std::cout << "mask " << m.wordCount() << " " << m.countOn() << " " << was << " " << m.setOff(63) << " "
            << m.setOff(63) << " on:";
  for (auto it = m.beginOn(); it; ++it) std::cout << " " << *it;
  std::cout << std::endl;
  Bonxai::Mask m1(1);
  std::cout << "mask1 " << m1.wordCount() << " " << m1.bitCount() << std::endl;

  const Bonxai::Point3D a = Bonxai::ConvertPoint<Bonxai::Point3D>(MemberPoint{1.5f, -2.25f, 3.0f});
  const Bonxai::Point3D b = Bonxai::ConvertPoint<Bonxai::Point3D>(MethodPoint{{0.1, 0.2, 0.3}});
  const Bonxai::Point3D d = Bonxai::ConvertPoint<Bonxai::Point3D>(std::vector<double>{7, 8, 9});
  std::cout << "convert " << a.x << " " << a.y << " " << a.z << " " << b.x << " " << b.y << " " << b.z << " "
            << d.x << " " << d.y << " " << d.z << std::endl;
  const Bonxai::CoordT pc = Bonxai::PosToCoord({-0.3, 0.05, 1.0}, 20.0);
  std::cout << "pos2coord " << pc.x << " " << pc.y << " " << pc.z << std::endl;
  std::cout << "hash " << std::hash<Bonxai::CoordT>()({-7, 12345, 99}) << " "
            << std::hash<Bonxai::CoordT>()({1 << 20, -(1 << 20), 3}) << std::endl;

  if (argc > 1) {
    std::ifstream in(argv[1], std::ios::binary);
    uint32_t n = 0;
    in.read(reinterpret_cast<char*>(&n), 4);
    std::vector<Bonxai::CoordT> ray;
    for (uint32_t k = 0; k < n; ++k) {
      Bonxai::CoordT e[2];
      in.read(reinterpret_cast<char*>(e), sizeof(e));
      Bonxai::ComputeRay(e[0], e[1], ray);
      std::cout << "ray " << k << ":";
      for (const auto& r : ray) std::cout << " " << r.x << " " << r.y << " " << r.z;
      std::cout << std::endl;
    }
  }
  return 0;
}
