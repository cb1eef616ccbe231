/*
 * bonxai_gpu.h -- C ABI of the MI355X (gfx950) sparse voxel map.
 *
 * Plain pointers and sizes only.  Every function returns a BXG_* status code
 * and never throws.  "_dev" pointers are device memory (from bxg_malloc or
 * any HIP allocation on the context's device); "_host" functions take host
 * memory, stage it, and return when the result is complete.  Work is
 * ordered on the context's stream; the device-pointer functions return once
 * the work is enqueued, and errors raised on the device (pool exhausted,
 * coordinate out of range) are reported by bxg_map_sync / bxg_map_stats.
 *
 * The reference interface each entry point replaces is cited as
 * file:line under the reference source tree (facontidavide/Bonxai):
 *   bonxai_core/include/bonxai/bonxai.hpp                    (VoxelGrid<T>)
 *   bonxai_map/include/bonxai_map/probabilistic_map.hpp      (ProbabilisticMap)
 *   bonxai_map/src/probabilistic_map.cpp
 *
 * Coordinates are int32 (x, y, z) triples, Bonxai::CoordT (grid_coord.hpp:62-78).
 * Supported grid shapes: inner_bits and leaf_bits in 1..4; coordinates with
 * |coord >> (inner_bits + leaf_bits)| < 2^20 per axis (for the default 2/3
 * bits: |coord| < 2^25 voxels, i.e. +-1677 km at 5 cm).  A coordinate outside
 * that range makes the call fail with BXG_ERR_COORD_RANGE.
 */
#ifndef BONXAI_GPU_H
#define BONXAI_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status */
#define BXG_OK 0
#define BXG_ERR_INVALID_ARG 1
#define BXG_ERR_OUT_OF_POOL 2   /* node pool or root directory full: cells unchanged */
#define BXG_ERR_HIP 3           /* HIP runtime error */
#define BXG_ERR_COORD_RANGE 4   /* coordinate outside the supported range */
#define BXG_ERR_BUDGET 5        /* the context's memory budget would be exceeded */
#define BXG_ERR_NOT_SUPPORTED 6
#define BXG_ERR_CAPACITY 7      /* an output buffer is too small (count still returned) */
#define BXG_ERR_MAP_FAILED 8    /* the map holds an unacknowledged device error */

/* map flags */
#define BXG_MAP_PROBABILISTIC 1u /* 4-byte ProbabilisticMap::CellT cells */

/* point flags */
#define BXG_POINTS_F64 1u /* xyz triples are double; default float */

/* clear options, Bonxai::ClearOption (bonxai.hpp:107-112) */
#define BXG_CLEAR_MEMORY 0
#define BXG_SET_ALL_CELLS_OFF 1

/* export filters */
#define BXG_EXPORT_ALL 0
#define BXG_EXPORT_OCCUPIED 1 /* probability_log > threshold (probabilistic_map.cpp:108-116) */
#define BXG_EXPORT_FREE 2     /* probability_log < threshold (probabilistic_map.cpp:118-126) */

typedef struct bxg_ctx bxg_ctx;
typedef struct bxg_map bxg_map;

/* ProbabilisticMap::Options (probabilistic_map.hpp:56-64), fixed-point log-odds */
typedef struct bxg_prob_options {
  int32_t prob_miss_log;
  int32_t prob_hit_log;
  int32_t clamp_min_log;
  int32_t clamp_max_log;
  int32_t occupancy_threshold_log;
} bxg_prob_options;

typedef struct bxg_map_stats {
  uint64_t leaf_count;
  uint64_t leaf_capacity;
  uint64_t root_count;
  uint64_t root_capacity;
  uint64_t scans_applied;       /* insertPointCloud calls completed */
  uint64_t device_bytes;        /* device memory held by the map */
  uint64_t last_rays;           /* counted runs only (bxg_set_counting) */
  uint64_t last_ray_updates;
  uint32_t update_count;        /* ProbabilisticMap::_update_count (1..3) */
  uint32_t status;              /* device error bits, 0 if none */
  uint64_t last_ray_visits;     /* counted runs: cells the rays crossed (sum of ray lengths) */
  uint64_t replays;             /* asynchronous inserts run again after a pool overflow */
} bxg_map_stats;

/* one scan of a batched insert */
typedef struct bxg_scan_desc {
  bxg_map* map;
  const void* points_dev; /* xyz triples (x, y, z first in every point) */
  uint64_t n;
  uint32_t flags;         /* BXG_POINTS_F64 */
  uint32_t stride;        /* bytes from one point to the next; 0 = packed xyz */
  double origin[3];       /* ignored when sensor_to_world is set */
  double max_range;
  /* NULL, or the ingest pre-step of bonxai_server.cpp:147-179 for float
   * points: 16 floats, row-major sensor-to-world matrix; non-finite points are
   * dropped, the others transformed, and the origin is the matrix's
   * translation column */
  const float* sensor_to_world;
} bxg_scan_desc;

typedef struct bxg_graph bxg_graph;

/* ----------------------------------------------------------- context */

/* One context per GPU: a stream, scratch buffers and a memory budget for
 * the node pools of its maps (0 = no limit). */
int bxg_ctx_create(int device, uint64_t budget_bytes, bxg_ctx** out);
int bxg_ctx_destroy(bxg_ctx* ctx);
int bxg_ctx_sync(bxg_ctx* ctx);
const char* bxg_ctx_last_error(bxg_ctx* ctx);
const char* bxg_status_string(int status);
int bxg_ctx_stream(bxg_ctx* ctx, void** stream_out);
/* bytes held by the maps of the context and the budget */
int bxg_ctx_memory(bxg_ctx* ctx, uint64_t* used_bytes, uint64_t* budget_bytes);
int bxg_ctx_device_info(bxg_ctx* ctx, char* name, uint64_t name_cap, int* cu_count,
                        uint64_t* hbm_bytes);

/* device memory and copies on the context's stream (blocking for d2h) */
int bxg_malloc(bxg_ctx* ctx, uint64_t bytes, void** out);
int bxg_free(bxg_ctx* ctx, void* ptr);
int bxg_memcpy_h2d(bxg_ctx* ctx, void* dst_dev, const void* src_host, uint64_t bytes);
int bxg_memcpy_d2h(bxg_ctx* ctx, void* dst_host, const void* src_dev, uint64_t bytes);
int bxg_memcpy_d2d(bxg_ctx* ctx, void* dst_dev, const void* src_dev, uint64_t bytes);

/* events on the context's stream (64 slots) and per-kernel profiling */
int bxg_event_record(bxg_ctx* ctx, uint32_t slot);
int bxg_event_elapsed(bxg_ctx* ctx, uint32_t slot_a, uint32_t slot_b, float* ms);
int bxg_profile_enable(bxg_ctx* ctx, int on);
int bxg_profile_read(bxg_ctx* ctx, const char* kernel, uint64_t* launches, double* total_ms);
int bxg_profile_reset(bxg_ctx* ctx);
/* exact per-scan counting of ray updates (slower; for statistics runs) */
int bxg_set_counting(bxg_ctx* ctx, int on);
/* launch tuning of the context, set by the autotuner: "ray_blocks_per_cu"
 * and "leaf_blocks_per_cu" (1..256), the workgroups per CU of the
 * work-stealing update pass and leaf-allocation pass over the rays;
 * "small_ray_blocks_per_cu" (1..64, default 2), the workgroups per CU of the
 * update pass and the fresh pass of a call with fewer points than the tuned
 * grid has lanes;
 * "ray_block", "ray_chunk", "ep_table_factor", "root_slack"; "ray_window"
 * (default 8) and "leaf_window" (default 32), how many cells / leaves a ray
 * may run ahead of the slowest ray of its wavefront; "fresh_pass" (default
 * 1): the rays of a map that is empty when its scan starts create their
 * leaves and mark their cells in one walk (0: the leaf pass and the update
 * pass, as for other maps); "fresh_blocks_per_cu" (default 6) and
 * "fresh_window" (default 16), that walk's workgroups per CU and lockstep
 * window in cells; "ray_order" (0..5, default 3): the order in which the rays
 * of a scan are walked, by direction only (0) or by classes of length (equal
 * parts of the max range) and direction: 1 and 2 are 4 classes over 32 x 32
 * and 16 classes over 16 x 16 direction bins, 3 and 4 the latter under 8 and
 * 16 direction sectors, 5 the former under 8 sectors; a scan with fewer than
 * 1/8 of its points beyond the max range keeps the direction order. Every
 * setting gives the same maps. */
int bxg_ctx_set_tuning(bxg_ctx* ctx, const char* key, int64_t value);
int bxg_ctx_get_tuning(bxg_ctx* ctx, const char* key, int64_t* value);

/* -------------------------------------------------------------- maps */

/* VoxelGrid(voxel_size, inner_bits, leaf_bits), bonxai.hpp:138,389-402.
 * value_bytes = sizeof(T) (0 for BinaryVoxelGrid, bonxai.hpp:335);
 * BXG_MAP_PROBABILISTIC makes a ProbabilisticMap grid (probabilistic_map.cpp:14-16,
 * value_bytes must be 4). */
int bxg_map_create(bxg_ctx* ctx, double voxel_size, uint32_t inner_bits, uint32_t leaf_bits,
                   uint32_t value_bytes, uint32_t flags, bxg_map** out);
int bxg_map_destroy(bxg_map* map);
/* grow the node pool / root directory ahead of use */
int bxg_map_reserve(bxg_map* map, uint64_t leaves, uint64_t roots);
/* VoxelGrid::clear(ClearOption), bonxai.hpp:678-687 */
int bxg_map_clear(bxg_map* map, int option);
/* VoxelGrid::releaseUnusedMemory, bonxai.hpp:367-387 */
int bxg_map_release_unused(bxg_map* map);
/* wait for the map's work; returns the device error of the map, if any */
int bxg_map_sync(bxg_map* map);
int bxg_map_clear_error(bxg_map* map);
int bxg_map_get_stats(bxg_map* map, bxg_map_stats* out);
/* VoxelGrid::activeCellsCount, bonxai.hpp:689-701 */
int bxg_map_active_cells(bxg_map* map, uint64_t* out);
/* VoxelGrid::memUsage, bonxai.hpp:649-676 (device bytes of the map) */
int bxg_map_mem_usage(bxg_map* map, uint64_t* out);
/* order-free digest of the on cells: count, sum, xor and remixed sum of a
 * 64-bit hash of (x, y, z, value bytes) */
int bxg_map_digest(bxg_map* map, uint64_t out[4]);
double bxg_map_voxel_size(bxg_map* map);

/* ------------------------------------------------ VoxelGrid<T> bulk */

/* Accessor::setValue over a batch in input order, bonxai.hpp:449-466.
 * was_on_dev (optional) gets the previous on-state of every input. */
int bxg_set_values(bxg_map* map, const int32_t* coords_dev, const void* vals_dev, uint64_t n,
                   uint8_t* was_on_dev);
/* Accessor::setCellOn(coord, default), bonxai.hpp:537-553 (defaults may be NULL: zero) */
int bxg_set_cells_on(bxg_map* map, const int32_t* coords_dev, const void* defaults_dev,
                     uint64_t n, uint8_t* was_on_dev);
/* Accessor::setCellOff, bonxai.hpp:556-569 */
int bxg_set_cells_off(bxg_map* map, const int32_t* coords_dev, uint64_t n, uint8_t* was_on_dev);
/* ConstAccessor::value / isCellOn, bonxai.hpp:496-534: vals_dev gets the cell
 * bytes (zero where absent), found_dev 1/0; either may be NULL */
int bxg_lookup(bxg_map* map, const int32_t* coords_dev, uint64_t n, void* vals_dev,
               uint8_t* found_dev);
/* forEachCell as a compaction, bonxai.hpp:704-743 (order unspecified).
 * Writes min(count, cap) cells; *count receives the number of matching
 * cells (BXG_ERR_CAPACITY if it exceeds cap).  Blocks until done. */
int bxg_export(bxg_map* map, int filter, int32_t* coords_dev, void* vals_dev, uint64_t cap,
               uint64_t* count);

/* host-memory forms of the above (stage, run, copy back, block) */
int bxg_set_values_host(bxg_map* map, const int32_t* coords, const void* vals, uint64_t n,
                        uint8_t* was_on);
int bxg_set_cells_on_host(bxg_map* map, const int32_t* coords, const void* defaults, uint64_t n,
                          uint8_t* was_on);
int bxg_set_cells_off_host(bxg_map* map, const int32_t* coords, uint64_t n, uint8_t* was_on);
int bxg_lookup_host(bxg_map* map, const int32_t* coords, uint64_t n, void* vals, uint8_t* found);
int bxg_export_host(bxg_map* map, int filter, int32_t* coords, void* vals, uint64_t cap,
                    uint64_t* count);

/* whole-leaf paging for host-side accessors: leaves named by their origin
 * (min-corner) coordinate; masks are leaf_words uint64 per leaf, cells
 * leaf_size * value_bytes bytes per leaf */
int bxg_read_leaves_host(bxg_map* map, const int32_t* origins, uint64_t n, uint64_t* masks,
                         void* cells, uint8_t* exists);
int bxg_write_leaves_host(bxg_map* map, const int32_t* origins, uint64_t n,
                          const uint64_t* masks, const void* cells);
int bxg_leaf_origins_host(bxg_map* map, int32_t* origins, uint64_t cap, uint64_t* count);

/* ------------------------------------------------- ProbabilisticMap */

/* ProbabilisticMap::setOptions / options, probabilistic_map.hpp:74-76 */
int bxg_set_options(bxg_map* map, const bxg_prob_options* opts);
int bxg_get_options(bxg_map* map, bxg_prob_options* opts);
/* the 1 -> 2 -> 3 scan counter (probabilistic_map.hpp:129), e.g. for resume */
int bxg_get_update_count(bxg_map* map, uint32_t* out);
int bxg_set_update_count(bxg_map* map, uint32_t value);

/* ProbabilisticMap::insertPointCloud(points, origin, max_range),
 * probabilistic_map.hpp:91-93,141-160.  points_dev: n xyz triples (float,
 * or double with BXG_POINTS_F64); origin already widened to double.
 * Enqueues only.  The map keeps the call until it is next confirmed: by
 * bxg_map_sync, by any other call on the map, or after 256 such inserts.  If
 * the map ran out of nodes, the scan that did changed no cell, later scans
 * of the map were skipped the same way, and the confirmation grows the pools
 * and runs them again in order.  Keep points_dev valid until then.  Other
 * device errors (a coordinate out of range) are reported by bxg_map_sync. */
int bxg_insert_point_cloud(bxg_map* map, const void* points_dev, uint64_t n, uint32_t flags,
                           const double origin[3], double max_range);
/* same, host points; blocks, and grows the pools and re-runs the scan if it
 * ran out of nodes (the failed run leaves every cell unchanged) */
int bxg_insert_point_cloud_host(bxg_map* map, const void* points, uint64_t n, uint32_t flags,
                                const double origin[3], double max_range);
/* The ingest pre-step of the reference's server (bonxai_server.cpp:147-179)
 * fused in front of the scan: float points in the sensor frame (stride bytes
 * apart, x y z first; 0 = packed), points with a non-finite coordinate
 * dropped, the others moved by the row-major sensor-to-world matrix in float
 * (pcl::transformPointCloud: x' = m00 x + (m01 y + (m02 z + m03)), the
 * association of PCL's SSE transform), then insertPointCloud(points,
 * origin = (m03, m13, m23), max_range).  Enqueues only, like
 * bxg_insert_point_cloud. */
int bxg_insert_point_cloud_tf(bxg_map* map, const float* points_dev, uint64_t n, uint32_t stride,
                              const float sensor_to_world[16], double max_range);
/* same, host points; blocks and grows the pools as bxg_insert_point_cloud_host */
int bxg_insert_point_cloud_tf_host(bxg_map* map, const float* points, uint64_t n, uint32_t stride,
                                   const float sensor_to_world[16], double max_range);
/* Allocation-free form of bxg_insert_point_cloud.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code that
 * must not call hipMalloc, such as a tight serving loop or a harness that
 * forbids it).  Same scan and results, but the call's scratch is a static
 * device buffer of the library (32 MiB on each GPU it runs on, held
 * until the process ends) for at most BXG_NOALLOC_MAX_POINTS points, and the
 * map's pools are used as reserved (bxg_map_reserve): nothing is allocated
 * or grown.  Blocks until the scan is done.  Returns BXG_ERR_INVALID_ARG for
 * more points than the maximum; BXG_ERR_OUT_OF_POOL if the pools cannot hold
 * the scan (no cell changed and update_count did not advance; the error is
 * cleared, so reserve more and call again); BXG_ERR_NOT_SUPPORTED while
 * addHitPoint/addMissPoint rays are due (bxg_insert_point_cloud traces
 * them); BXG_ERR_MAP_FAILED if an earlier asynchronous insert of the map
 * overflowed (bxg_map_sync repairs that, growing the pools).  Every call on
 * one GPU uses that GPU's one buffer: the library runs such calls one at a
 * time per GPU, and other users of the buffer would race with them. */
#define BXG_NOALLOC_MAX_POINTS 524288u
int bxg_insert_point_cloud_noalloc(bxg_map* map, const void* points_dev, uint64_t n, uint32_t flags,
                                   const double origin[3], double max_range);
/* many scans in one pass over a flattened point queue; scans naming the
 * same map run in input order.  Blocks; a scan whose map ran out of nodes is
 * re-run after the pool grows (its failed run left the cells untouched). */
int bxg_insert_batch(bxg_ctx* ctx, const bxg_scan_desc* scans, uint32_t nscans);
/* addHitPoint / addMissPoint, probabilistic_map.hpp:98-103 (cpp:30-54):
 * endpoint update now, ray traced by the next insert (host point, blocks) */
int bxg_add_point_host(bxg_map* map, const double xyz[3], int is_hit);
/* isOccupied / isUnknown / isFree, probabilistic_map.cpp:56-75:
 * state 1 occupied, -1 free, 0 unknown */
int bxg_classify(bxg_map* map, const int32_t* coords_dev, uint64_t n, int8_t* state_dev);
int bxg_classify_host(bxg_map* map, const int32_t* coords, uint64_t n, int8_t* state);

/* ---------------------------------------------------- batched graphs */

/* Capture the kernels of a batch of up to max_scans scans and max_points
 * points into a HIP graph; bxg_graph_set_scans loads a frame's scans into
 * the graph's descriptor buffer and bxg_graph_launch replays it.  A launch
 * only enqueues; each map keeps its scan of the frame until it is confirmed
 * (as for bxg_insert_point_cloud), which re-runs it should the frame have
 * overflowed the map's pools.  bxg_graph_set_scans and bxg_graph_destroy
 * confirm the previous frame's maps first; keep a frame's points valid
 * until then. */
int bxg_graph_create(bxg_ctx* ctx, uint32_t max_scans, uint64_t max_points, bxg_graph** out);
int bxg_graph_set_scans(bxg_graph* g, const bxg_scan_desc* scans, uint32_t nscans);
int bxg_graph_launch(bxg_graph* g);
int bxg_graph_destroy(bxg_graph* g);

/* ------------------------------------------------------------ serving */

/* A request engine over maps resident on one context (no reference
 * counterpart: the reference's callers integrate one scan at a time,
 * bonxai_server.cpp:178-181).  Requests of one map run in submission order,
 * one at a time, so each sees exactly the requests submitted before it on
 * that map; requests of different maps overlap on the GPU (inserts of
 * several maps form one scan wave, lookups and exports run on the engine's
 * own streams beside it).  An insert is admitted when its map's node pools
 * can grow to the scan's head-room estimate within the context's memory
 * budget, else it waits for in-flight work (BXG_ERR_BUDGET if there is
 * none).  All buffers are device memory owned by the caller and must stay
 * valid until the request is waited for.  While a server runs, its maps
 * must be used through the server only. */
typedef struct bxg_server bxg_server;
typedef struct bxg_server_options {
  uint32_t max_inflight;    /* requests on the GPU at once (default 32) */
  uint32_t max_batch_scans; /* inserts merged into one scan wave (default 16, at most 64) */
  uint32_t streams;         /* streams for lookups and exports (default 4) */
  uint32_t reserved;
} bxg_server_options;
typedef struct bxg_request_info {
  int32_t status;     /* BXG_OK or an error; -1 while pending */
  uint32_t kind;      /* 0 insert, 1 lookup, 2 export */
  uint64_t count;     /* export: matching cells */
  uint64_t submit_ns; /* steady-clock times of submission, launch, completion */
  uint64_t admit_ns;
  uint64_t done_ns;
} bxg_request_info;
typedef struct bxg_server_stats {
  uint64_t submitted, completed, failed, retried, deferred;
  uint64_t waves, wave_scans; /* scan waves launched, scans in them */
  uint64_t queued, inflight;
} bxg_server_stats;

int bxg_server_create(bxg_ctx* ctx, const bxg_server_options* opts, bxg_server** out);
/* waits for every request, then stops the engine */
int bxg_server_destroy(bxg_server* s);
int bxg_server_drain(bxg_server* s);
/* ProbabilisticMap::insertPointCloud (probabilistic_map.hpp:141-160) */
int bxg_server_insert(bxg_server* s, bxg_map* map, const void* points_dev, uint64_t n, uint32_t flags,
                      const double origin[3], double max_range, uint64_t* id);
/* ConstAccessor::value / isCellOn over a batch (bonxai.hpp:496-534) */
int bxg_server_lookup(bxg_server* s, bxg_map* map, const int32_t* coords_dev, uint64_t n, void* vals_dev,
                      uint8_t* found_dev, uint64_t* id);
/* forEachCell / getOccupiedVoxels / getFreeVoxels (bonxai.hpp:704-743,
 * probabilistic_map.cpp:108-126): filter BXG_EXPORT_*; count in the info */
int bxg_server_export(bxg_server* s, bxg_map* map, int filter, int32_t* coords_dev, void* vals_dev,
                      uint64_t cap, uint64_t* id);
/* block until done; returns the request's status and forgets the id */
int bxg_server_wait(bxg_server* s, uint64_t id, bxg_request_info* info);
/* non-blocking: status -1 while pending */
int bxg_server_query(bxg_server* s, uint64_t id, bxg_request_info* info);
int bxg_server_get_stats(bxg_server* s, bxg_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* BONXAI_GPU_H */
