// Timing of the drop-in C++ API the way the reference's ROS node drives it
// (bonxai_server.cpp:140-181, 221): per scan, host points in a std::vector
// go through Bonxai::ProbabilisticMap::insertPointCloud (copy to the GPU,
// scan, wait), then getOccupiedVoxels reads the occupied cells back, as the
// node's publisher does after every cloud.
//
// usage: dropin_bench scans.bin
//
// scans.bin (little-endian): double resolution, double max_range, uint32
// scan count, then per scan uint32 n, double origin[3], float xyz[3n].
// Prints one JSON object: scans, points, per-scan insert and insert+export
// milliseconds (p50, p99, mean) and rays per second over the inserts.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "bonxai_map/probabilistic_map.hpp"

namespace {

struct FloatPoint {  // a float point type with x, y, z members, as pcl::PointXYZ
  float x, y, z;
};

struct Scan {
  std::vector<FloatPoint> points;
  FloatPoint origin;
};

template <class T>
T get(std::ifstream& in) {
  T v;
  if (!in.read(reinterpret_cast<char*>(&v), sizeof(T))) throw std::runtime_error("truncated scan file");
  return v;
}

double percentile(std::vector<double> v, double q) {
  std::sort(v.begin(), v.end());
  const double pos = q * double(v.size() - 1);
  const size_t lo = size_t(pos);
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - double(lo));
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::cerr << "usage: dropin_bench scans.bin\n";
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    if (!in) throw std::runtime_error("cannot open the scan file");
    const double res = get<double>(in);
    const double max_range = get<double>(in);
    const uint32_t nscans = get<uint32_t>(in);
    std::vector<Scan> scans(nscans);
    uint64_t npoints = 0;
    for (auto& s : scans) {
      const uint32_t n = get<uint32_t>(in);
      double o[3];
      for (double& v : o) v = get<double>(in);
      s.origin = FloatPoint{float(o[0]), float(o[1]), float(o[2])};
      s.points.resize(n);
      if (!in.read(reinterpret_cast<char*>(s.points.data()), std::streamsize(sizeof(FloatPoint) * n)))
        throw std::runtime_error("truncated scan file");
      npoints += n;
    }
    // the origins are written as doubles and read back as floats here, as a
    // ROS node holds its sensor position in a pcl::PointXYZ: the float origin
    // is what the map sees
    Bonxai::ProbabilisticMap warm(res);
    warm.insertPointCloud(scans[0].points, scans[0].origin, max_range);

    Bonxai::ProbabilisticMap map(res);
    std::vector<double> t_insert, t_total;
    std::vector<Bonxai::CoordT> occupied;
    size_t last_occupied = 0;
    using clock = std::chrono::steady_clock;
    for (const auto& s : scans) {
      const auto t0 = clock::now();
      map.insertPointCloud(s.points, s.origin, max_range);
      const auto t1 = clock::now();
      occupied.clear();
      map.getOccupiedVoxels(occupied);
      const auto t2 = clock::now();
      last_occupied = occupied.size();
      t_insert.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
      t_total.push_back(std::chrono::duration<double, std::milli>(t2 - t0).count());
    }
    double sum_insert = 0, sum_total = 0;
    for (double v : t_insert) sum_insert += v;
    for (double v : t_total) sum_total += v;
    std::printf(
        "{\"scans\": %u, \"points\": %llu, \"occupied_after_last\": %zu, "
        "\"insert_ms\": {\"p50\": %.4f, \"p99\": %.4f, \"mean\": %.4f}, "
        "\"insert_and_export_ms\": {\"p50\": %.4f, \"p99\": %.4f, \"mean\": %.4f}, "
        "\"insert_rays_per_s\": %.1f, \"active_cells\": %zu, \"update_count\": %u}\n",
        nscans, (unsigned long long)npoints, last_occupied, percentile(t_insert, 0.5),
        percentile(t_insert, 0.99), sum_insert / nscans, percentile(t_total, 0.5), percentile(t_total, 0.99),
        sum_total / nscans, double(npoints) / (sum_insert / 1e3), map.grid().activeCellsCount(),
        unsigned(map.updateCount()));
  } catch (const std::exception& e) {
    std::cerr << "dropin_bench: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
