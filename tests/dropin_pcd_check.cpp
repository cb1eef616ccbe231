// Host-only check of include/bonxai_map/pcl_utils.hpp (no GPU call):
// ReadPointsFromPCD of a file, its points as doubles to a binary file, then
// WritePointsFromPCD and a second read of what was written.
//
// usage: dropin_pcd_check in.pcd points.bin copy.pcd
// prints "points N" and "reread equal 1" when the copy reads back the same.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "bonxai_map/pcl_utils.hpp"

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  std::vector<Bonxai::Point3D> pts;
  Bonxai::ReadPointsFromPCD(argv[1], pts);
  std::ofstream out(argv[2], std::ios::binary);
  for (const auto& p : pts) {
    out.write(reinterpret_cast<const char*>(&p.x), 8);
    out.write(reinterpret_cast<const char*>(&p.y), 8);
    out.write(reinterpret_cast<const char*>(&p.z), 8);
  }
  out.close();
  Bonxai::WritePointsFromPCD(argv[3], pts);
  std::vector<Bonxai::Point3D> again;
  Bonxai::ReadPointsFromPCD(argv[3], again);
  bool same = again.size() == pts.size();
  for (size_t i = 0; same && i < pts.size(); ++i)
    same = again[i].x == pts[i].x && again[i].y == pts[i].y && again[i].z == pts[i].z;
  std::cout << "points " << pts.size() << std::endl;
  std::cout << "reread equal " << same << std::endl;
  try {
    Bonxai::ReadPointsFromPCD("/nonexistent/file.pcd", again);
  } catch (const std::runtime_error& e) {
    std::cout << "missing file: " << e.what() << std::endl;
  }
  return 0;
}
