// Round trip of a VoxelGrid<int> through Serialize / GetHeaderInfo /
// Deserialize, the scenario of the reference's examples/test_serialization.cpp:6-64
// (a 10x10x10 walk over [-0.5, 0.5) m in steps of 0.1 m, value = visiting order).
// The walk accumulates x += 0.1 in double, and posToCoord maps two of the ten
// steps of every axis to one voxel (-0.30000000000000004 floors to -4), so the
// grid holds 729 cells, each with the value of its last visit.
//
// usage: dropin_serialization [stream_out]
//   stream_out  also write the serialized bytes there
// Exit status 0 and "Round trip looks good!" when every value came back.
//
// usage: dropin_serialization --load stream_in cells_out stream_out
//   Deserialize<float> a stream written elsewhere, dump its cells (uint64 n,
//   then n x (int32 xyz, float value) from forEachCell) and Serialize it again.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "bonxai/bonxai.hpp"
#include "bonxai/serialization.hpp"

namespace {

int load(const char* in_path, const char* cells_path, const char* out_path) {
  std::ifstream in(in_path, std::ios::binary);
  if (!in) return 2;
  std::string header;
  std::getline(in, header);
  auto grid = Bonxai::Deserialize<float>(in, Bonxai::GetHeaderInfo(header));
  std::vector<std::pair<Bonxai::CoordT, float>> cells;
  grid.forEachCell([&](const float& v, const Bonxai::CoordT& c) { cells.push_back({c, v}); });
  std::ofstream co(cells_path, std::ios::binary);
  const uint64_t n = cells.size();
  co.write(reinterpret_cast<const char*>(&n), 8);
  for (const auto& e : cells) {
    co.write(reinterpret_cast<const char*>(&e.first), 12);
    co.write(reinterpret_cast<const char*>(&e.second), 4);
  }
  std::ofstream so(out_path, std::ios::binary);
  Bonxai::Serialize(so, grid);
  std::cout << "loaded " << n << " cells, " << grid.activeCellsCount() << " active" << std::endl;
  return n == grid.activeCellsCount() ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 5 && std::string(argv[1]) == "--load") return load(argv[2], argv[3], argv[4]);
  const double res = 0.1;
  Bonxai::VoxelGrid<int> grid(res);
  auto acc = grid.createAccessor();
  if (acc.value({0, 0, 0}) != nullptr) {
    std::cout << "a new grid is not empty" << std::endl;
    return 1;
  }
  std::cout << "Empty as expected" << std::endl;

  int n = 0;
  std::unordered_map<Bonxai::CoordT, int> expected;  // last write wins
  for (double x = -0.5; x < 0.5; x += res)
    for (double y = -0.5; y < 0.5; y += res)
      for (double z = -0.5; z < 0.5; z += res) {
        const Bonxai::CoordT c = grid.posToCoord(x, y, z);
        acc.setValue(c, n);
        expected[c] = n++;
      }

  std::ostringstream os(std::ios::binary);
  Bonxai::Serialize(os, grid);
  const std::string bytes = os.str();
  if (argc > 1) std::ofstream(argv[1], std::ios::binary).write(bytes.data(), bytes.size());

  std::istringstream is(bytes, std::ios::binary);
  char header[256];
  is.getline(header, sizeof(header));
  const Bonxai::HeaderInfo info = Bonxai::GetHeaderInfo(header);
  auto back = Bonxai::Deserialize<int>(is, info);
  std::cout << "header: " << header << std::endl;
  std::cout << "Original grid memory: " << grid.memUsage() << std::endl;
  std::cout << "New grid memory: " << back.memUsage() << std::endl;

  auto bacc = back.createAccessor();
  bool ok = back.activeCellsCount() == expected.size() && grid.activeCellsCount() == expected.size();
  std::cout << "cells: " << grid.activeCellsCount() << " -> " << back.activeCellsCount() << " (expected "
            << expected.size() << ")" << std::endl;
  for (const auto& e : expected) {
    const int* v = bacc.value(e.first);
    if (!v || *v != e.second) {
      std::cout << " Problem at cell " << e.first.x << " " << e.first.y << " " << e.first.z << std::endl;
      ok = false;
    }
  }
  if (!ok) return 1;
  std::cout << "Round trip looks good!" << std::endl;
  return 0;
}
