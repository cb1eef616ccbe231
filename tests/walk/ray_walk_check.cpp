// Host check of the scan kernels' cell walk (bonxai_amd/csrc/bxg_walk.h):
// every run collect_cells returns, from the ray's start or from a seek,
// against the closed form of RayIterator's cells (probabilistic_map.hpp:
// 162-203), cell_k = o + sgn(d) floor((2k|d| + m) / (2m)): the run's cells in
// order, all in one leaf, as long as possible (ended by a leaf border, the
// walk's end or kRun cells), their leaf-local indices, and the walk's state
// after the run equal to a seek to its next cell.
//
// usage: ray_walk_check SEED COUNT; prints "ok <runs>" or the first mismatch.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "bxg_walk.h"

using bxg::RayWalk;

static int64_t cell(int32_t o, int32_t s, int64_t d, int64_t m, int64_t k) {
  return o + s * ((2 * k * d + m) / (2 * m));
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  std::mt19937_64 rng(std::strtoull(argv[1], nullptr, 10));
  const long count = std::strtol(argv[2], nullptr, 10);
  long runs = 0;
  for (long t = 0; t < count; ++t) {
    const uint32_t lb = 1 + (uint32_t)(rng() % 4);
    const int32_t span = (t % 5 == 0) ? (1 << 12) : (t % 5 == 1 ? (1 << 22) : 40 + (int32_t)(rng() % 300));
    const int32_t base = (int32_t)(rng() % (1u << 24)) - (1 << 23);
    int32_t o[3], e[3];
    for (int a = 0; a < 3; ++a) {
      o[a] = base + (int32_t)(rng() % (2 * span + 1)) - span;
      e[a] = o[a] + (int32_t)(rng() % (2 * span + 1)) - span;
      if (rng() % 4 == 0) e[a] = o[a];  // axis-aligned rays
    }
    RayWalk w;
    w.init(o[0], o[1], o[2], e[0], e[1], e[2]);
    const int64_t m = w.mx;
    if (m == 0) continue;
    const int64_t d[3] = {w.dx, w.dy, w.dz};
    const int32_t s[3] = {w.sx, w.sy, w.sz};
    // the range walked: the ray [0, m), or with the endpoint cell [0, m], or a
    // segment from a seek (long rays: a window of at most 4096 cells)
    const int64_t last = m + (int64_t)(rng() % 2);
    int64_t k0 = (rng() % 3 == 0) ? 0 : (int64_t)(rng() % (uint64_t)last);
    int64_t k1 = k0 + 1 + (int64_t)(rng() % (uint64_t)(last - k0));
    if (k1 - k0 > 4096) k1 = k0 + 4096;
    if (k0 == 0)
      w.kend = (int32_t)k1;
    else
      w.seek((int32_t)k0, (int32_t)k1);
    const int32_t lm = (1 << lb) - 1;
    int64_t k = k0;
    while (!w.done()) {
      uint32_t li[bxg::kRun];
      int64_t c[3];
      for (int a = 0; a < 3; ++a) c[a] = cell(o[a], s[a], d[a], m, k);
      if (w.k != k || w.cx != c[0] || w.cy != c[1] || w.cz != c[2]) {
        std::printf("state mismatch ray %ld at cell %lld\n", t, (long long)k);
        return 1;
      }
      const uint32_t n = bxg::collect_cells(w, lb, li);
      ++runs;
      if (n < 1 || n > (uint32_t)bxg::kRun) {
        std::printf("run length %u ray %ld\n", n, t);
        return 1;
      }
      int64_t leaf[3];
      for (int a = 0; a < 3; ++a) leaf[a] = c[a] >> lb;
      for (uint32_t j = 0; j < (uint32_t)bxg::kRun; ++j) {
        if (j < n) {
          int64_t q[3];
          for (int a = 0; a < 3; ++a) q[a] = cell(o[a], s[a], d[a], m, k + j);
          for (int a = 0; a < 3; ++a)
            if ((q[a] >> lb) != leaf[a]) {
              std::printf("cell outside the run's leaf ray %ld cell %lld\n", t, (long long)(k + j));
              return 1;
            }
          const uint32_t want = (uint32_t)((q[0] & lm) | ((q[1] & lm) << lb) | ((q[2] & lm) << (2 * lb)));
          if (li[j] != want) {
            std::printf("index mismatch ray %ld cell %lld: %u vs %u\n", t, (long long)(k + j), li[j], want);
            return 1;
          }
        } else if (li[j] != li[0]) {
          std::printf("padding slot %u ray %ld\n", j, t);
          return 1;
        }
      }
      const int64_t kn = k + n;
      if (kn < k1 && n < (uint32_t)bxg::kRun) {
        bool same = true;
        for (int a = 0; a < 3; ++a) same = same && (cell(o[a], s[a], d[a], m, kn) >> lb) == leaf[a];
        if (same) {
          std::printf("run ended early ray %ld cell %lld\n", t, (long long)kn);
          return 1;
        }
      }
      // the state after the run is the seek to its next cell
      RayWalk r;
      r.init(o[0], o[1], o[2], e[0], e[1], e[2]);
      if (kn > 0) r.seek((int32_t)kn, (int32_t)k1);
      if (w.k != kn || w.cx != r.cx || w.cy != r.cy || w.cz != r.cz || w.ex != r.ex || w.ey != r.ey ||
          w.ez != r.ez) {
        std::printf("state after run ray %ld cell %lld differs from seek\n", t, (long long)kn);
        return 1;
      }
      k = kn;
    }
    if (k != k1) {
      std::printf("walk ended at %lld, not %lld, ray %ld\n", (long long)k, (long long)k1, t);
      return 1;
    }
  }
  std::printf("ok %ld\n", runs);
  return 0;
}
