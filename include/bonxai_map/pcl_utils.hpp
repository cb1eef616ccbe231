// pcl_utils.hpp -- point-cloud file I/O of the GPU-backed Bonxai drop-in.
//
// Same functions as the reference's bonxai_map/include/bonxai_map/pcl_utils.hpp
// (implemented there with PCL, bonxai_map/src/pcl_utils.cpp:10-73), written here
// without PCL: ReadPointsFromPCD reads the x, y, z fields of a PCD v0.5-0.7 file
// with DATA ascii or binary (F 4 or F 8 fields; binary_compressed is refused),
// WritePointsFromPCD writes an ascii PCD with FIELDS x y z as float, 9
// significant digits (every float reads back unchanged).  Values are read as
// PCL's PointXYZ does: as float, then widened to double.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "bonxai/grid_coord.hpp"

#if __has_include(<eigen3/Eigen/Core>)
#include <eigen3/Eigen/Core>
#define BONXAI_GPU_PCD_EIGEN 1
#endif

namespace Bonxai {

namespace pcd_detail {

struct Field {
  std::string name;
  int size = 4;
  char type = 'F';
  int count = 1;
  size_t offset = 0;
};

inline std::vector<float> read_xyz(const std::string& filepath) {
  std::ifstream in(filepath, std::ios::binary);
  if (!in) throw std::runtime_error("File doesn't exist");
  std::vector<Field> fields;
  size_t npoints = 0, width = 0, height = 1;
  std::string data, line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ls(line);
    std::string key;
    ls >> key;
    if (key == "FIELDS" || key == "COLUMNS") {
      std::string f;
      while (ls >> f) fields.push_back({f});
    } else if (key == "SIZE") {
      for (auto& f : fields) ls >> f.size;
    } else if (key == "TYPE") {
      for (auto& f : fields) ls >> f.type;
    } else if (key == "COUNT") {
      for (auto& f : fields) ls >> f.count;
    } else if (key == "WIDTH") {
      ls >> width;
    } else if (key == "HEIGHT") {
      ls >> height;
    } else if (key == "POINTS") {
      ls >> npoints;
    } else if (key == "DATA") {
      ls >> data;
      break;
    }
  }
  if (!npoints) npoints = width * height;
  int ix = -1, iy = -1, iz = -1;
  size_t rec = 0, col = 0;
  std::vector<size_t> first_col(fields.size());
  for (size_t k = 0; k < fields.size(); ++k) {
    fields[k].offset = rec;
    first_col[k] = col;
    rec += size_t(fields[k].size) * fields[k].count;
    col += fields[k].count;
    if (fields[k].name == "x") ix = int(k);
    if (fields[k].name == "y") iy = int(k);
    if (fields[k].name == "z") iz = int(k);
  }
  if (ix < 0 || iy < 0 || iz < 0) throw std::runtime_error("Problem loading file: no x, y, z fields");
  for (int k : {ix, iy, iz})
    if (fields[k].type != 'F' || (fields[k].size != 4 && fields[k].size != 8))
      throw std::runtime_error("Problem loading file: x, y, z must be float fields");
  std::vector<float> xyz(3 * npoints);
  if (data == "ascii") {
    std::vector<std::string> tok(col);
    for (size_t i = 0; i < npoints; ++i) {
      for (auto& t : tok)
        if (!(in >> t)) throw std::runtime_error("Problem loading file: truncated ascii data");
      const int axes[3] = {ix, iy, iz};
      for (int a = 0; a < 3; ++a) xyz[3 * i + a] = std::strtof(tok[first_col[axes[a]]].c_str(), nullptr);
    }
  } else if (data == "binary") {
    std::vector<char> buf(rec);
    for (size_t i = 0; i < npoints; ++i) {
      if (!in.read(buf.data(), rec)) throw std::runtime_error("Problem loading file: truncated binary data");
      const int axes[3] = {ix, iy, iz};
      for (int a = 0; a < 3; ++a) {
        const Field& f = fields[axes[a]];
        if (f.size == 4) {
          std::memcpy(&xyz[3 * i + a], buf.data() + f.offset, 4);
        } else {
          double d;
          std::memcpy(&d, buf.data() + f.offset, 8);
          xyz[3 * i + a] = float(d);
        }
      }
    }
  } else {
    throw std::runtime_error("Problem loading file: DATA " + data + " is not supported");
  }
  return xyz;
}

template <typename P>
inline void write_xyz(const std::string& filepath, const std::vector<P>& points) {
  std::FILE* f = std::fopen(filepath.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + filepath);
  std::fprintf(f,
               "# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\nFIELDS x y z\nSIZE 4 4 4\n"
               "TYPE F F F\nCOUNT 1 1 1\nWIDTH %zu\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS %zu\n"
               "DATA ascii\n",
               points.size(), points.size());
  for (const auto& p : points)
    std::fprintf(f, "%.9g %.9g %.9g\n", double(float(point_traits::get(p, 0))),
                 double(float(point_traits::get(p, 1))), double(float(point_traits::get(p, 2))));
  std::fclose(f);
}

}  // namespace pcd_detail

/// pcl_utils.cpp:10-26: the x, y, z of every point of a PCD file.
inline bool ReadPointsFromPCD(const std::string& filepath, std::vector<Point3D>& points) {
  const std::vector<float> xyz = pcd_detail::read_xyz(filepath);
  points.clear();
  points.reserve(xyz.size() / 3);
  for (size_t i = 0; i < xyz.size(); i += 3) points.push_back({xyz[i], xyz[i + 1], xyz[i + 2]});
  return true;
}

/// pcl_utils.cpp:54-73
inline void WritePointsFromPCD(const std::string& filepath, const std::vector<Point3D>& points) {
  pcd_detail::write_xyz(filepath, points);
}
inline void WritePointsFromPCD(const std::string& filepath, const std::vector<CoordT>& points) {
  pcd_detail::write_xyz(filepath, points);
}

#ifdef BONXAI_GPU_PCD_EIGEN
inline bool ReadPointsFromPCD(const std::string& filepath, std::vector<Eigen::Vector3d>& points) {
  const std::vector<float> xyz = pcd_detail::read_xyz(filepath);
  points.clear();
  points.reserve(xyz.size() / 3);
  for (size_t i = 0; i < xyz.size(); i += 3) points.emplace_back(xyz[i], xyz[i + 1], xyz[i + 2]);
  return true;
}
inline void WritePointsFromPCD(const std::string& filepath, const std::vector<Eigen::Vector3d>& points) {
  pcd_detail::write_xyz(filepath, points);
}
#endif

}  // namespace Bonxai
