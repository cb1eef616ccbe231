// serialization.hpp -- the reference's grid stream format for GPU-resident grids.
//
// Serialize / GetHeaderInfo / Deserialize of bonxai_core/include/bonxai/serialization.hpp:77-199,
// byte-compatible with it: a text header line
//     Bonxai::VoxelGrid<TYPE,INNER_BITS,LEAF_BITS>(RESOLUTION)\n
// (TYPE the demangled C++ type name, RESOLUTION printed with %lf), then
//     uint32 root count
//     per root:  int32 x, y, z of the root's min corner
//                the inner mask words (uint64, max(2^(3 inner_bits) / 64, 1) of them)
//                per inner bit, upward: the leaf mask words, then the bytes of
//                every on cell, upward.
// All values little-endian as stored.  Roots are written in ascending (z, y, x)
// order of their min corner (the reference's order is its hash table's); the
// bytes of a root do not depend on the order.  A whole grid moves in one
// device-to-host (or host-to-device) page transfer.
#pragma once

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <typeinfo>
#include <vector>

#include "bonxai/bonxai.hpp"

#ifdef __GNUG__
#include <cxxabi.h>
#include <cstdlib>
#endif

namespace Bonxai {

struct HeaderInfo {
  std::string type_name;
  int inner_bits = 0;
  int leaf_bits = 0;
  double resolution = 0;
};

namespace details {

inline std::string demangle(const char* name) {
#ifdef __GNUG__
  int status = -1;
  char* s = abi::__cxa_demangle(name, nullptr, nullptr, &status);
  std::string out = (status == 0 && s) ? std::string(s) : std::string(name);
  std::free(s);
  return out;
#else
  return name;
#endif
}

template <typename T>
inline void put(std::ostream& out, const T& v) {
  out.write(reinterpret_cast<const char*>(&v), sizeof(T));
}
template <typename T>
inline T get(std::istream& in) {
  T v{};
  in.read(reinterpret_cast<char*>(&v), sizeof(T));
  if (!in) throw std::runtime_error("truncated Bonxai grid stream");
  return v;
}

}  // namespace details

/// Write grid to out (serialization.hpp:77-118).
template <typename DataT>
inline void Serialize(std::ostream& out, const VoxelGrid<DataT>& grid) {
  constexpr size_t vb = std::is_same_v<DataT, EmptyVoxel> ? 0 : sizeof(DataT);
  const uint32_t ib = grid.innerBits(), lb = grid.leafBits(), log2n = ib + lb;
  const uint32_t inner_size = 1u << (3 * ib), leaf_size = 1u << (3 * lb);
  const uint32_t inner_words = std::max<uint32_t>(inner_size >> 6, 1u);
  const uint32_t leaf_words = std::max<uint32_t>(leaf_size >> 6, 1u);
  char header[256];
  std::snprintf(header, sizeof(header), "Bonxai::VoxelGrid<%s,%d,%d>(%lf)\n",
                details::demangle(typeid(DataT).name()).c_str(), int(ib), int(lb), grid.voxelSize());
  out.write(header, std::strlen(header));

  // every leaf of the device map, in one transfer
  grid.flush();
  bxg_map* m = grid.deviceMap();
  bxg_ctx* c = grid.deviceContext();
  uint64_t nl = 0;
  int rc = bxg_leaf_origins_host(m, nullptr, 0, &nl);
  if (rc != BXG_OK && rc != BXG_ERR_CAPACITY) gpu::check(rc, c);
  std::vector<CoordT> origins(nl);
  std::vector<uint64_t> masks(nl * leaf_words);
  std::vector<uint8_t> cells(std::max<size_t>(nl * leaf_size * vb, 1));
  std::vector<uint8_t> exists(nl);
  if (nl) {
    gpu::check(bxg_leaf_origins_host(m, &origins[0].x, nl, &nl), c);
    gpu::check(bxg_read_leaves_host(m, &origins[0].x, nl, masks.data(), vb ? cells.data() : nullptr,
                                    exists.data()),
               c);
  }
  // group the leaves by root, then by inner index
  const int32_t rmask = ~((1 << log2n) - 1);
  std::map<std::tuple<int32_t, int32_t, int32_t>, std::vector<std::pair<uint32_t, size_t>>> roots;
  for (size_t k = 0; k < nl; ++k) {
    const CoordT& o = origins[k];
    roots[{o.z & rmask, o.y & rmask, o.x & rmask}].push_back({grid.getInnerIndex(o), k});
  }
  details::put(out, uint32_t(roots.size()));
  std::vector<uint64_t> inner(inner_words);
  for (auto& r : roots) {
    details::put(out, std::get<2>(r.first));
    details::put(out, std::get<1>(r.first));
    details::put(out, std::get<0>(r.first));
    auto& leaves = r.second;
    std::sort(leaves.begin(), leaves.end());
    std::fill(inner.begin(), inner.end(), 0);
    for (const auto& l : leaves) inner[l.first >> 6] |= uint64_t(1) << (l.first & 63);
    for (uint64_t w : inner) details::put(out, w);
    for (const auto& l : leaves) {
      const uint64_t* mw = &masks[l.second * leaf_words];
      for (uint32_t w = 0; w < leaf_words; ++w) details::put(out, mw[w]);
      if (!vb) continue;
      const uint8_t* cb = &cells[l.second * leaf_size * vb];
      for (uint32_t w = 0; w < leaf_words; ++w) {
        uint64_t bits = mw[w];
        while (bits) {
          const uint32_t i = (w << 6) + uint32_t(__builtin_ctzll(bits));
          bits &= bits - 1;
          if (i >= leaf_size) break;
          out.write(reinterpret_cast<const char*>(cb + size_t(i) * vb), vb);
        }
      }
    }
  }
}

/// Parse the header line (serialization.hpp:127-150).
inline HeaderInfo GetHeaderInfo(std::string header) {
  const std::string prefix = "Bonxai::VoxelGrid<";
  if (header.rfind(prefix, 0) != 0) throw std::runtime_error("Header wasn't recognized");
  const size_t open = header.find('(', prefix.size());
  const size_t close = header.find(')', open == std::string::npos ? 0 : open);
  const size_t gt = header.rfind('>', open);
  if (open == std::string::npos || close == std::string::npos || gt == std::string::npos || gt < prefix.size())
    throw std::runtime_error("Header wasn't recognized");
  const std::string args = header.substr(prefix.size(), gt - prefix.size());
  const size_t c2 = args.rfind(',');
  const size_t c1 = c2 == std::string::npos || c2 == 0 ? std::string::npos : args.rfind(',', c2 - 1);
  if (c1 == std::string::npos) throw std::runtime_error("Header wasn't recognized");
  HeaderInfo info;
  info.type_name = args.substr(0, c1);
  info.inner_bits = std::stoi(args.substr(c1 + 1, c2 - c1 - 1));
  info.leaf_bits = std::stoi(args.substr(c2 + 1));
  info.resolution = std::stod(header.substr(open + 1, close - open - 1));
  return info;
}

/// Read the grid stream after its header line into `grid`, an empty grid of
/// the header's type and shape (the body of Deserialize).
template <typename DataT>
inline void DeserializeInto(std::istream& input, const HeaderInfo& info, VoxelGrid<DataT>& grid) {
  if (details::demangle(typeid(DataT).name()) != info.type_name)
    throw std::runtime_error("DataT does not match");
  if (int(grid.innerBits()) != info.inner_bits || int(grid.leafBits()) != info.leaf_bits)
    throw std::runtime_error("grid shape does not match the stream");
  constexpr size_t vb = std::is_same_v<DataT, EmptyVoxel> ? 0 : sizeof(DataT);
  const uint32_t ib = info.inner_bits, lb = info.leaf_bits;
  const uint32_t inner_size = 1u << (3 * ib), leaf_size = 1u << (3 * lb);
  const uint32_t inner_words = std::max<uint32_t>(inner_size >> 6, 1u);
  const uint32_t leaf_words = std::max<uint32_t>(leaf_size >> 6, 1u);
  std::vector<int32_t> origins;
  std::vector<uint64_t> masks;
  std::vector<uint8_t> cells;
  const uint32_t roots = details::get<uint32_t>(input);
  std::vector<uint64_t> inner(inner_words);
  for (uint32_t r = 0; r < roots; ++r) {
    const int32_t rx = details::get<int32_t>(input);
    const int32_t ry = details::get<int32_t>(input);
    const int32_t rz = details::get<int32_t>(input);
    for (auto& w : inner) w = details::get<uint64_t>(input);
    for (uint32_t ii = 0; ii < inner_size; ++ii) {
      if (!((inner[ii >> 6] >> (ii & 63)) & 1)) continue;
      const uint32_t im = (1u << ib) - 1;
      origins.push_back(rx | int32_t((ii & im) << lb));
      origins.push_back(ry | int32_t(((ii >> ib) & im) << lb));
      origins.push_back(rz | int32_t(((ii >> (2 * ib)) & im) << lb));
      const size_t m0 = masks.size();
      for (uint32_t w = 0; w < leaf_words; ++w) masks.push_back(details::get<uint64_t>(input));
      if (!vb) continue;
      const size_t c0 = cells.size();
      cells.resize(c0 + size_t(leaf_size) * vb, 0);
      for (uint32_t w = 0; w < leaf_words; ++w) {
        uint64_t bits = masks[m0 + w];
        while (bits) {
          const uint32_t i = (w << 6) + uint32_t(__builtin_ctzll(bits));
          bits &= bits - 1;
          if (i >= leaf_size) break;
          input.read(reinterpret_cast<char*>(&cells[c0 + size_t(i) * vb]), vb);
          if (!input) throw std::runtime_error("truncated Bonxai grid stream");
        }
      }
    }
  }
  const size_t n = origins.size() / 3;
  if (n) {
    gpu::check(bxg_write_leaves_host(grid.deviceMap(), origins.data(), n, masks.data(),
                                     vb ? cells.data() : nullptr),
               grid.deviceContext());
    grid.refresh(true);
  }
}

/// Build a grid from the stream after its header line (serialization.hpp:159-199).
template <typename DataT>
inline VoxelGrid<DataT> Deserialize(std::istream& input, HeaderInfo info) {
  if (details::demangle(typeid(DataT).name()) != info.type_name)
    throw std::runtime_error("DataT does not match");
  VoxelGrid<DataT> grid(info.resolution, uint8_t(info.inner_bits), uint8_t(info.leaf_bits));
  DeserializeInto(input, info, grid);
  return grid;
}

}  // namespace Bonxai
