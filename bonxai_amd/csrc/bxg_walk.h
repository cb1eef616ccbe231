// bxg_walk.h -- the integer 3-D ray walk of RayIterator
// (probabilistic_map.hpp:162-203) as the scan kernels run it, one run of
// cells inside one leaf at a time.  Plain C++ with no HIP dependence, so the
// same code compiles for the GPU (bxg_kernels.hip) and for the host check
// of tests/test_ray_walk.py.
#pragma once

#include <stdint.h>

#if defined(__HIP__)
#define BXG_HD __host__ __device__ __forceinline__
#else
#define BXG_HD inline
#endif

namespace bxg {

constexpr int kRun = 8;  // cells of one leaf handled per step of a ray

// Cell 0 is the origin voxel, then max|d|-1 further cells; the end voxel is
// not part of the ray and a zero-length ray is empty.  (cx, cy, cz) is cell
// k.  The error terms are kept biased by h = ceil(mx / 2): the reference's
// "2 e >= mx" is "e - h >= 0".
struct RayWalk {
  int32_t cx, cy, cz;
  int32_t sx, sy, sz;
  int32_t dx, dy, dz;
  int32_t ex, ey, ez;  // error terms minus h
  int32_t mx, k, kend;
  BXG_HD void init(int32_t ox, int32_t oy, int32_t oz, int32_t tx, int32_t ty, int32_t tz) {
    cx = ox;
    cy = oy;
    cz = oz;
    dx = tx - ox;
    dy = ty - oy;
    dz = tz - oz;
    sx = dx < 0 ? -1 : 1;
    sy = dy < 0 ? -1 : 1;
    sz = dz < 0 ? -1 : 1;
    dx = dx < 0 ? -dx : dx;
    dy = dy < 0 ? -dy : dy;
    dz = dz < 0 ? -dz : dz;
    mx = dx > dy ? dx : dy;
    mx = mx > dz ? mx : dz;
    ex = ey = ez = -((mx + 1) >> 1);
    k = 0;
    kend = mx;
  }
  // walk only cells [k0, k1) of the ray: cell k0 and its error terms by the
  // closed form (n steps along an axis after k cells: n = floor((2k|d| +
  // mx) / (2mx)), error k|d| - n mx), the same state stepping from 0 reaches
  BXG_HD void seek(int32_t k0, int32_t k1) {
    const int64_t m2 = 2 * (int64_t)mx;
    const int64_t nx = (2 * (int64_t)k0 * dx + mx) / m2;
    const int64_t ny = (2 * (int64_t)k0 * dy + mx) / m2;
    const int64_t nz = (2 * (int64_t)k0 * dz + mx) / m2;
    cx += sx * (int32_t)nx;
    cy += sy * (int32_t)ny;
    cz += sz * (int32_t)nz;
    const int32_t h = (mx + 1) >> 1;
    ex = (int32_t)((int64_t)k0 * dx - nx * mx) - h;
    ey = (int32_t)((int64_t)k0 * dy - ny * mx) - h;
    ez = (int32_t)((int64_t)k0 * dz - nz * mx) - h;
    k = k0;
    kend = k1;
  }
  BXG_HD bool done() const { return k >= kend; }
};

// Advance a walk through the cells of its current leaf (2^lb cells a side),
// at most kRun of them, recording their leaf-local indices; returns how many.
// Slots past the run hold the run's first index, so that every slot names a
// cell of the leaf and a load from it needs no branch.
//
// Per axis the walk keeps T = (steps the axis may still take inside the leaf)
// + (cells advanced so far): a step along the axis leaves T unchanged, a cell
// without one adds 1, and cell k + j is in the leaf iff T >= j on every
// axis.  The local index moves by the axis's stride on a step.  With the
// error terms biased by h, the step test is the sign of e, so one cell costs
// a few adds, shifts and masks an axis and no compare-and-select chain.
BXG_HD uint32_t collect_cells(RayWalk& w, uint32_t lb, uint32_t li[kRun]) {
  const int32_t lm = (1 << lb) - 1;
  const int32_t ux = w.cx & lm, uy = w.cy & lm, uz = w.cz & lm;
  const int32_t tx0 = w.sx > 0 ? lm - ux : ux;
  const int32_t ty0 = w.sy > 0 ? lm - uy : uy;
  const int32_t tz0 = w.sz > 0 ? lm - uz : uz;
  // a step's stride in the leaf-local index, negated (added where the axis
  // does not step, after all three were added)
  const int32_t qx = w.sx, qy = w.sy * (1 << lb), qz = w.sz * (1 << (2 * lb));
  const int32_t qall = qx + qy + qz;
  int32_t idx = ux | (uy << lb) | (uz << (2 * lb));
  const uint32_t l0 = (uint32_t)idx;
  int32_t ex = w.ex, ey = w.ey, ez = w.ez;
  int32_t tx = tx0, ty = ty0, tz = tz0;
  const int32_t rem = w.kend - w.k;
  const int32_t mx = w.mx, dx = w.dx, dy = w.dy, dz = w.dz;
  uint32_t n = 0;
#pragma unroll
  for (int j = 0; j < kRun; ++j) {
    int32_t t = tx < ty ? tx : ty;
    t = t < tz ? t : tz;
    const bool in = (j < rem) & (t >= j);
    li[j] = in ? (uint32_t)idx : l0;
    if (in) {
      n = (uint32_t)j + 1u;
      ex += dx;
      ey += dy;
      ez += dz;
      const int32_t nx = ex >> 31, ny = ey >> 31, nz = ez >> 31;  // -1: no step
      ex -= ~nx & mx;
      ey -= ~ny & mx;
      ez -= ~nz & mx;
      tx -= nx;
      ty -= ny;
      tz -= nz;
      idx += qall - (nx & qx) - (ny & qy) - (nz & qz);
    }
  }
  // steps taken along an axis: n less the cells without one (T's growth)
  w.cx += w.sx * ((int32_t)n - (tx - tx0));
  w.cy += w.sy * ((int32_t)n - (ty - ty0));
  w.cz += w.sz * ((int32_t)n - (tz - tz0));
  w.ex = ex;
  w.ey = ey;
  w.ez = ez;
  w.k += (int32_t)n;
  return n;
}

}  // namespace bxg
