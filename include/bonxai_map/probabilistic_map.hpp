// probabilistic_map.hpp -- Bonxai::ProbabilisticMap with scan integration on the GPU.
//
// A drop-in for the reference's bonxai_map/include/bonxai_map/probabilistic_map.hpp:27-137
// and bonxai_map/src/probabilistic_map.cpp: same class, types, options, method
// names and results.  insertPointCloud runs the whole scan -- max-range clip,
// posToCoord, first-point-per-voxel hit/miss update, the free-space ray of every
// endpoint, the update_id cycle -- as GPU kernels on the map's device copy; the
// cell values are bit-identical to the reference's (probabilistic_map.hpp:141-160,
// probabilistic_map.cpp:30-106).
//
// Vector3D is Eigen::Vector3d when Eigen is installed, otherwise a small xyz type
// with the same x()/y()/z() accessors.
//
// Differences, all on the safe side:
//   - the map cannot be copied or moved (the reference's implicit move leaves
//     its cached accessor pointing at the moved-from grid);
//   - isOccupied/isFree/isUnknown on a coordinate whose leaf does not exist,
//     followed by an insert into that leaf, is well defined here;
//   - points and origin are widened to double exactly as ConvertPoint<Vector3D>
//     does; float clouds travel to the GPU as float and are widened there.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#if __has_include(<eigen3/Eigen/Geometry>)
#include <eigen3/Eigen/Geometry>
#define BONXAI_GPU_HAS_EIGEN 1
#endif

#include "bonxai/bonxai.hpp"
#include "bonxai/serialization.hpp"

#if defined(__GNUC__) && !defined(__clang__)
#define BONXAI_MATH_CONSTEXPR constexpr  // GCC folds std::log / std::exp
#else
#define BONXAI_MATH_CONSTEXPR
#endif

namespace Bonxai {

/// Visit the voxels of the segment key_origin -> key_end (probabilistic_map.hpp:162-203):
/// key_origin first, then max|delta| - 1 more cells, key_end excluded; nothing if
/// the two are equal; func returns false to stop.  Cell k of the walk is
/// origin + sign(delta) * floor((2 k |delta| + m) / (2 m)) per axis, m = max|delta|,
/// sign(0) = +1, which is the same sequence as the reference's error-accumulating loop.
template <class Functor>
inline void RayIterator(const CoordT& key_origin, const CoordT& key_end, const Functor& func) {
  if (key_origin == key_end) return;
  const int64_t dx = int64_t(key_end.x) - key_origin.x, dy = int64_t(key_end.y) - key_origin.y,
                dz = int64_t(key_end.z) - key_origin.z;
  const int64_t ax = std::llabs(dx), ay = std::llabs(dy), az = std::llabs(dz);
  const int64_t m = std::max(ax, std::max(ay, az));
  const int64_t sx = dx < 0 ? -1 : 1, sy = dy < 0 ? -1 : 1, sz = dz < 0 ? -1 : 1;
  for (int64_t k = 0; k < m; ++k) {
    const CoordT c = {int32_t(key_origin.x + sx * ((2 * k * ax + m) / (2 * m))),
                      int32_t(key_origin.y + sy * ((2 * k * ay + m) / (2 * m))),
                      int32_t(key_origin.z + sz * ((2 * k * az + m) / (2 * m)))};
    if (!func(c)) return;
  }
}

/// The cells RayIterator visits, in order (probabilistic_map.hpp:13-19).
inline void ComputeRay(const CoordT& key_origin, const CoordT& key_end, std::vector<CoordT>& ray) {
  ray.clear();
  RayIterator(key_origin, key_end, [&ray](const CoordT& c) {
    ray.push_back(c);
    return true;
  });
}

class ProbabilisticMap {
 public:
#ifdef BONXAI_GPU_HAS_EIGEN
  using Vector3D = Eigen::Vector3d;
#else
  struct Vector3D {
    double v[3] = {0, 0, 0};
    Vector3D() = default;
    Vector3D(double x, double y, double z) : v{x, y, z} {}
    double x() const { return v[0]; }
    double y() const { return v[1]; }
    double z() const { return v[2]; }
    double& x() { return v[0]; }
    double& y() { return v[1]; }
    double& z() { return v[2]; }
    double operator[](int i) const { return v[i]; }
  };
#endif

  /// Log-odds of prob as a fixed-point integer, 6 decimals (probabilistic_map.hpp:34-36).
  [[nodiscard]] static BONXAI_MATH_CONSTEXPR int32_t logods(float prob) {
    return int32_t(1e6 * std::log(prob / (1.0 - prob)));
  }
  /// Inverse of logods (probabilistic_map.hpp:38-42).
  [[nodiscard]] static BONXAI_MATH_CONSTEXPR float prob(int32_t logods_fixed) {
    const float l = float(logods_fixed) * 1e-6;
    return (1.0 - 1.0 / (1.0 + std::exp(l)));
  }

  /// logods(0.5f) (probabilistic_map.cpp:8)
  static constexpr int32_t UnknownProbability = 0;

  /// The 32-bit cell (probabilistic_map.hpp:44-53): 4-bit signed update id and
  /// 28-bit fixed-point log-odds, laid out as GCC/Clang do on x86-64 and as the
  /// GPU map stores it (id in the low 4 bits).
  struct CellT {
    int32_t update_id : 4;
    int32_t probability_log : 28;
    CellT() : update_id(0), probability_log(UnknownProbability) {}
  };
  static_assert(sizeof(CellT) == 4, "CellT must be one 32-bit word");

  /// OctoMap's defaults (probabilistic_map.hpp:56-64): logods(0.4f), logods(0.7f),
  /// logods(0.12f), logods(0.97f), logods(0.5).
  struct Options {
    int32_t prob_miss_log = -405465;
    int32_t prob_hit_log = 847297;
    int32_t clamp_min_log = -1992430;
    int32_t clamp_max_log = 3476099;
    int32_t occupancy_threshold_log = 0;
  };

  /// probabilistic_map.cpp:14-16 (inner_bits 2, leaf_bits 3), on the default GPU context.
  explicit ProbabilisticMap(double resolution)
      : ProbabilisticMap(resolution, gpu::Context::Default()) {}

  ProbabilisticMap(double resolution, gpu::Context& ctx)
      : _grid(resolution, 2, 3, ctx, BXG_MAP_PROBABILISTIC), _accessor(_grid.createConstAccessor()) {}

  ProbabilisticMap(const ProbabilisticMap&) = delete;
  ProbabilisticMap& operator=(const ProbabilisticMap&) = delete;
  ProbabilisticMap(ProbabilisticMap&&) = delete;
  ProbabilisticMap& operator=(ProbabilisticMap&&) = delete;

  [[nodiscard]] VoxelGrid<CellT>& grid() { return _grid; }
  [[nodiscard]] const VoxelGrid<CellT>& grid() const { return _grid; }

  [[nodiscard]] const Options& options() const { return _options; }
  void setOptions(const Options& options) {
    const bxg_prob_options o{options.prob_miss_log, options.prob_hit_log, options.clamp_min_log,
                             options.clamp_max_log, options.occupancy_threshold_log};
    gpu::check(bxg_set_options(_grid.deviceMap(), &o), _grid.deviceContext());
    _options = options;
  }

  /// Integrate one scan (probabilistic_map.hpp:141-160): points further than
  /// max_range from origin become misses at max_range along their ray, the
  /// others hits; then every endpoint's ray from origin is cleared.
  template <typename PointT, typename Allocator>
  void insertPointCloud(const std::vector<PointT, Allocator>& points, const PointT& origin,
                        double max_range) {
    const auto from = ConvertPoint<Vector3D>(origin);
    const double o[3] = {from.x(), from.y(), from.z()};
    _grid.flush();
    using S = point_traits::scalar_t<PointT>;
    if constexpr (std::is_same_v<S, float>) {
      _fbuf.resize(points.size() * 3);
      for (size_t i = 0; i < points.size(); ++i) {
        _fbuf[3 * i] = point_traits::get(points[i], 0);
        _fbuf[3 * i + 1] = point_traits::get(points[i], 1);
        _fbuf[3 * i + 2] = point_traits::get(points[i], 2);
      }
      gpu::check(bxg_insert_point_cloud_host(_grid.deviceMap(), _fbuf.data(), points.size(), 0u, o, max_range),
                 _grid.deviceContext());
    } else {
      _dbuf.resize(points.size() * 3);
      for (size_t i = 0; i < points.size(); ++i) {
        _dbuf[3 * i] = double(point_traits::get(points[i], 0));
        _dbuf[3 * i + 1] = double(point_traits::get(points[i], 1));
        _dbuf[3 * i + 2] = double(point_traits::get(points[i], 2));
      }
      gpu::check(bxg_insert_point_cloud_host(_grid.deviceMap(), _dbuf.data(), points.size(),
                                             BXG_POINTS_F64, o, max_range),
                 _grid.deviceContext());
    }
    afterScan(o, max_range);
  }

  /// The reference server's ingest pre-step fused into the scan
  /// (bonxai_server.cpp:147-179): sensor-frame points with a non-finite
  /// coordinate are dropped, the others moved by the float sensor-to-world
  /// matrix (row-major 4x4) as pcl::transformPointCloud does, and the sensor
  /// origin is the matrix's translation.  Equivalent to that filter, the
  /// transform and insertPointCloud(points, PointXYZ(t), max_range), in one pass.
  template <typename PointT, typename Allocator>
  void insertPointCloudTransformed(const std::vector<PointT, Allocator>& sensor_points,
                                   const float sensor_to_world[16], double max_range) {
    _grid.flush();
    _fbuf.resize(sensor_points.size() * 3);
    for (size_t i = 0; i < sensor_points.size(); ++i)
      for (int a = 0; a < 3; ++a) _fbuf[3 * i + a] = float(point_traits::get(sensor_points[i], a));
    gpu::check(bxg_insert_point_cloud_tf_host(_grid.deviceMap(), _fbuf.data(), sensor_points.size(), 12,
                                              sensor_to_world, max_range),
               _grid.deviceContext());
    const double o[3] = {double(sensor_to_world[3]), double(sensor_to_world[7]), double(sensor_to_world[11])};
    afterScan(o, max_range);
  }
#ifdef BONXAI_GPU_HAS_EIGEN
  template <typename PointT, typename Allocator>
  void insertPointCloudTransformed(const std::vector<PointT, Allocator>& sensor_points,
                                   const Eigen::Matrix4f& sensor_to_world, double max_range) {
    float m[16];
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) m[4 * r + c] = sensor_to_world(r, c);
    insertPointCloudTransformed(sensor_points, m, max_range);
  }
#endif

  /// The same for xyz triples already in GPU memory (float, or double with
  /// f64); enqueued on the map's stream without waiting when no host page is
  /// cached.  The map keeps the call until its next confirmation (any other
  /// call on the map, or bxg_map_sync), which re-runs it after growing the
  /// pools should they have overflowed: keep points_dev valid until then.
  void insertPointCloudDevice(const void* points_dev, size_t n, bool f64, const Vector3D& origin,
                              double max_range) {
    const double o[3] = {origin.x(), origin.y(), origin.z()};
    _grid.flush();
    gpu::check(bxg_insert_point_cloud(_grid.deviceMap(), points_dev, n, f64 ? BXG_POINTS_F64 : 0u, o,
                                      max_range),
               _grid.deviceContext());
    afterScan(o, max_range);
  }

  /// probabilistic_map.cpp:30-41: the endpoint cell is updated now; its ray is
  /// cleared by the next insertPointCloud (the reference's updateFreeCells).
  void addHitPoint(const Vector3D& point) { addPoint(point, 1); }
  /// probabilistic_map.cpp:43-54
  void addMissPoint(const Vector3D& point) { addPoint(point, 0); }

  /// probabilistic_map.cpp:56-75 (a missing cell is unknown)
  [[nodiscard]] bool isOccupied(const CoordT& coord) const {
    const CellT* c = _accessor.value(coord);
    return c && c->probability_log > _options.occupancy_threshold_log;
  }
  [[nodiscard]] bool isUnknown(const CoordT& coord) const {
    const CellT* c = _accessor.value(coord);
    return !c || c->probability_log == _options.occupancy_threshold_log;
  }
  [[nodiscard]] bool isFree(const CoordT& coord) const {
    const CellT* c = _accessor.value(coord);
    return c && c->probability_log < _options.occupancy_threshold_log;
  }

  /// isOccupied / isFree / isUnknown of many cells on the GPU: 1, -1, 0.
  void classify(const std::vector<CoordT>& coords, std::vector<int8_t>& state) const {
    state.assign(coords.size(), 0);
    _grid.flush();
    gpu::check(bxg_classify_host(_grid.deviceMap(), coords.empty() ? nullptr : &coords[0].x, coords.size(),
                                 state.data()),
               _grid.deviceContext());
  }

  /// probabilistic_map.cpp:108-116: clears coords, then every cell with
  /// probability_log above the threshold (order unspecified).
  void getOccupiedVoxels(std::vector<CoordT>& coords) {
    coords.clear();
    _grid.exportCells(coords, nullptr, BXG_EXPORT_OCCUPIED);
  }
  /// probabilistic_map.cpp:118-126
  void getFreeVoxels(std::vector<CoordT>& coords) {
    coords.clear();
    _grid.exportCells(coords, nullptr, BXG_EXPORT_FREE);
  }
  /// probabilistic_map.hpp:115-124: appends the min corner of every occupied voxel.
  template <typename PointT>
  void getOccupiedVoxels(std::vector<PointT>& points) {
    std::vector<CoordT> coords;
    getOccupiedVoxels(coords);
    for (const auto& c : coords) {
      const auto p = _grid.coordToPos(c);
      points.emplace_back(p.x, p.y, p.z);
    }
  }

  /// The 1 -> 2 -> 3 scan counter (probabilistic_map.hpp:129), for checkpoints.
  [[nodiscard]] uint32_t updateCount() const {
    uint32_t v = 0;
    gpu::check(bxg_get_update_count(_grid.deviceMap(), &v), _grid.deviceContext());
    return v;
  }
  void setUpdateCount(uint32_t v) {
    gpu::check(bxg_set_update_count(_grid.deviceMap(), v), _grid.deviceContext());
  }

  /// Checkpoint of the map.  The reference has none: its stream
  /// (serialization.hpp:77-118) holds a grid only, and a map restored from it
  /// restarts the 1 -> 2 -> 3 scan counter (probabilistic_map.hpp:129), so its
  /// next scans diverge from the uninterrupted map's.  Written here: the grid
  /// stream Serialize(grid()) writes, byte for byte (the reference's
  /// GetHeaderInfo / Deserialize<CellT> read it and stop before what follows),
  /// then the text line "Bonxai::ProbabilisticMap(update_count=N)\n".  The
  /// options are not part of it (the caller sets them, as for a new map).
  /// Throws std::logic_error while addHitPoint / addMissPoint rays wait for the
  /// next insertPointCloud: they are not part of the checkpoint.
  void saveCheckpoint(std::ostream& out) const {
    if (!_pending_pts.empty())
      throw std::logic_error("addHitPoint/addMissPoint rays are pending: insert a scan before a checkpoint");
    const uint32_t n = updateCount();
    Serialize(out, _grid);
    char line[64];
    std::snprintf(line, sizeof(line), "Bonxai::ProbabilisticMap(update_count=%u)\n", unsigned(n));
    out << line;
    if (!out) throw std::runtime_error("checkpoint write failed");
  }

  /// Restore a checkpoint written by saveCheckpoint into this map (built with
  /// the same resolution): its cells and scan counter are replaced, and the
  /// next insertPointCloud continues as the saved map would have.
  void loadCheckpoint(std::istream& in) {
    std::string header;
    if (!std::getline(in, header)) throw std::runtime_error("empty checkpoint");
    const HeaderInfo info = GetHeaderInfo(header + "\n");
    char want[64], got[64];
    std::snprintf(want, sizeof(want), "%lf", _grid.voxelSize());
    std::snprintf(got, sizeof(got), "%lf", info.resolution);
    if (std::string(want) != got) throw std::runtime_error("checkpoint resolution differs from the map's");
    _grid.clear(CLEAR_MEMORY);
    _pending_pts.clear();
    DeserializeInto(in, info, _grid);
    std::string trailer;
    unsigned n = 0;
    char tail = 0;
    if (!std::getline(in, trailer) ||
        std::sscanf(trailer.c_str(), "Bonxai::ProbabilisticMap(update_count=%u%c", &n, &tail) != 2 ||
        tail != ')' || n < 1 || n > 3)
      throw std::runtime_error("checkpoint trailer missing or malformed");
    setUpdateCount(n);
  }

 private:
  void addPoint(const Vector3D& p, int hit) {
    const double xyz[3] = {p.x(), p.y(), p.z()};
    _grid.flush();
    gpu::check(bxg_add_point_host(_grid.deviceMap(), xyz, hit), _grid.deviceContext());
    // only the endpoint's cell changed; its ray is walked by the next scan
    _pending_pts.push_back({xyz[0], xyz[1], xyz[2]});
    _grid.refreshNear(Point3D(xyz[0], xyz[1], xyz[2]), kReachMargin * _grid.voxelSize());
  }

  // After a scan from origin o with range max_range: re-read the cached pages
  // it may have changed.  Every cell it touched lies on a voxel ray from o's
  // voxel to an endpoint voxel within max_range of o, or to a pending
  // addHit/addMissPoint endpoint, so nothing beyond that reach plus a few
  // voxels (kReachMargin > 2 sqrt(3): both ends' voxel offsets) changed.
  void afterScan(const double o[3], double max_range) {
    double reach = max_range;
    for (const auto& p : _pending_pts) {
      const double dx = p.x - o[0], dy = p.y - o[1], dz = p.z - o[2];
      reach = std::max(reach, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    _pending_pts.clear();
    _grid.refreshNear(Point3D(o[0], o[1], o[2]), reach + kReachMargin * _grid.voxelSize());
  }

  static constexpr double kReachMargin = 4.0;

  VoxelGrid<CellT> _grid;
  Options _options;
  mutable VoxelGrid<CellT>::ConstAccessor _accessor;
  std::vector<float> _fbuf;
  std::vector<double> _dbuf;
  std::vector<Point3D> _pending_pts;  // addHit/addMissPoint endpoints whose rays are due
};

}  // namespace Bonxai
