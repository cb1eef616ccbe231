/*
 * scenes.c -- seeded synthetic workloads for the voxel-map benchmarks and
 * parity tests.
 *
 * Every generator is deterministic: a SplitMix64 stream seeded by the caller,
 * normals from an Irwin-Hall sum of 12 uniforms, and sine/cosine from a
 * fixed-order polynomial written here, so no output depends on the platform's
 * libm (only floor and sqrt, which IEEE 754 fixes exactly, are used).  Build
 * with -ffp-contract=off so no multiply-add is fused differently between
 * compilers.  Outputs are float32 xyz triples (AoS, 12 bytes a point) or int32
 * coordinate triples.
 *
 * Workloads:
 *   bxs_box_scan      100k-point box room, 90% wall/floor/ceiling points with
 *                     N(0, sigma) noise, 10% uniform inside, origin at centre.
 *   bxs_clusters      int32 voxel coords from Gaussian clusters (VoxelGrid
 *                     insert workload) and the matching lookup set.
 *   bxs_lidar_*       spinning-LiDAR scans ray-cast on a ground plane plus
 *                     axis-aligned boxes, sensor on a smooth seeded path.
 *   bxs_depth_frame   pinhole depth frame ray-cast in a closed box room.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define BXS_EXPORT __attribute__((visibility("default")))

/* ----------------------------------------------------------------- RNG */

typedef struct {
  uint64_t s;
} bxs_rng;

static uint64_t bxs_next(bxs_rng* r) {
  uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* uniform double in [0, 1) with 53 random bits */
static double bxs_u01(bxs_rng* r) { return (double)(bxs_next(r) >> 11) * (1.0 / 9007199254740992.0); }

static double bxs_uniform(bxs_rng* r, double a, double b) { return a + (b - a) * bxs_u01(r); }

/* approximately N(0, 1): Irwin-Hall sum of twelve uniforms minus six */
static double bxs_normal(bxs_rng* r) {
  double s = 0.0;
  for (int i = 0; i < 12; ++i) s += bxs_u01(r);
  return s - 6.0;
}

static bxs_rng bxs_seed(uint64_t seed, uint64_t stream) {
  bxs_rng r;
  r.s = seed * 0xD1342543DE82EF95ull + stream * 0x2545F4914F6CDD1Dull + 0x1234567ull;
  bxs_next(&r);
  return r;
}

/* ------------------------------------------------------- trigonometry */

static const double BXS_PI = 3.14159265358979323846;

/* sin and cos of x: reduce by multiples of pi/2 (Cody-Waite with a two-part
 * constant), then Taylor polynomials of fixed degree on [-pi/4, pi/4]
 * (truncation error below 2e-17). */
static void bxs_sincos(double x, double* s, double* c) {
  const double two_over_pi = 0.63661977236758134308;
  const double p1 = 1.5707963267341256e+00; /* pi/2 high part (33 bits) */
  const double p2 = 6.0770943833657612e-11; /* pi/2 - p1 */
  const double k = floor(x * two_over_pi + 0.5);
  const double y = (x - k * p1) - k * p2;
  const double y2 = y * y;
  double sn = 1.0 / 6227020800.0;             /* 1/13! */
  sn = sn * y2 - 1.0 / 39916800.0;            /* 1/11! */
  sn = sn * y2 + 1.0 / 362880.0;              /* 1/9!  */
  sn = sn * y2 - 1.0 / 5040.0;                /* 1/7!  */
  sn = sn * y2 + 1.0 / 120.0;
  sn = sn * y2 - 1.0 / 6.0;
  sn = y + y * (y2 * sn);
  double cs = 1.0 / 87178291200.0;            /* 1/14! */
  cs = cs * y2 - 1.0 / 479001600.0;           /* 1/12! */
  cs = cs * y2 + 1.0 / 3628800.0;             /* 1/10! */
  cs = cs * y2 - 1.0 / 40320.0;               /* 1/8!  */
  cs = cs * y2 + 1.0 / 720.0;
  cs = cs * y2 - 1.0 / 24.0;
  cs = cs * y2 + 0.5;
  cs = 1.0 - y2 * cs;
  const int64_t q = ((int64_t)k) & 3;
  switch (q) {
    case 0: *s = sn; *c = cs; break;
    case 1: *s = cs; *c = -sn; break;
    case 2: *s = -sn; *c = -cs; break;
    default: *s = -cs; *c = sn; break;
  }
}

BXS_EXPORT void bxs_sincos_array(const double* x, uint64_t n, double* s, double* c) {
  for (uint64_t i = 0; i < n; ++i) bxs_sincos(x[i], &s[i], &c[i]);
}

/* ---------------------------------------------------------- box room */

/* A closed box room [-hx, hx] x [-hy, hy] x [0, h], origin at the centre
 * (0, 0, h/2).  A fraction (1 - inside_frac) of the points lie on the six
 * faces, the face chosen with probability proportional to its area, with
 * N(0, sigma) noise on every axis; the rest are uniform inside the room. */
BXS_EXPORT void bxs_box_scan(uint64_t seed, uint64_t n, double sx, double sy, double sz,
                             double sigma, double inside_frac, float* out, double origin[3]) {
  bxs_rng r = bxs_seed(seed, 1);
  const double hx = 0.5 * sx, hy = 0.5 * sy;
  const double axy = sx * sy, axz = sx * sz, ayz = sy * sz;
  const double atot = 2.0 * (axy + axz + ayz);
  origin[0] = 0.0;
  origin[1] = 0.0;
  origin[2] = 0.5 * sz;
  for (uint64_t i = 0; i < n; ++i) {
    double x, y, z;
    if (bxs_u01(&r) < inside_frac) {
      x = bxs_uniform(&r, -hx, hx);
      y = bxs_uniform(&r, -hy, hy);
      z = bxs_uniform(&r, 0.0, sz);
    } else {
      const double a = bxs_u01(&r) * atot;
      const double u = bxs_u01(&r), v = bxs_u01(&r);
      if (a < axy) { /* floor */
        x = -hx + sx * u; y = -hy + sy * v; z = 0.0;
      } else if (a < 2.0 * axy) { /* ceiling */
        x = -hx + sx * u; y = -hy + sy * v; z = sz;
      } else if (a < 2.0 * axy + axz) {
        x = -hx + sx * u; y = -hy; z = sz * v;
      } else if (a < 2.0 * axy + 2.0 * axz) {
        x = -hx + sx * u; y = hy; z = sz * v;
      } else if (a < 2.0 * axy + 2.0 * axz + ayz) {
        x = -hx; y = -hy + sy * u; z = sz * v;
      } else {
        x = hx; y = -hy + sy * u; z = sz * v;
      }
      x += sigma * bxs_normal(&r);
      y += sigma * bxs_normal(&r);
      z += sigma * bxs_normal(&r);
    }
    out[3 * i] = (float)x;
    out[3 * i + 1] = (float)y;
    out[3 * i + 2] = (float)z;
  }
}

/* ------------------------------------------------------ voxel clusters */

/* n int32 voxel coordinates drawn from nclusters isotropic Gaussian clusters
 * (std sigma metres) whose centres are uniform in [0, extent)^3 metres; a
 * position becomes a coordinate as floor(p / voxel), all in double. */
BXS_EXPORT void bxs_clusters(uint64_t seed, uint64_t n, uint32_t nclusters, double extent,
                             double sigma, double voxel, int32_t* out) {
  bxs_rng r = bxs_seed(seed, 2);
  double* c = (double*)malloc(3 * sizeof(double) * nclusters);
  for (uint32_t k = 0; k < 3 * nclusters; ++k) c[k] = bxs_uniform(&r, 0.0, extent);
  const double inv = 1.0 / voxel;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t k = (uint32_t)(bxs_next(&r) % nclusters);
    for (int a = 0; a < 3; ++a) {
      const double p = c[3 * k + a] + sigma * bxs_normal(&r);
      out[3 * i + a] = (int32_t)floor(p * inv);
    }
  }
  free(c);
}

/* Lookup set: each query is, with probability hit_frac, a uniformly chosen
 * coordinate of `inserted` (n_ins triples), otherwise uniform in the cube of
 * coordinates [0, extent/voxel)^3. */
BXS_EXPORT void bxs_lookups(uint64_t seed, uint64_t n, const int32_t* inserted, uint64_t n_ins,
                            double hit_frac, double extent, double voxel, int32_t* out) {
  bxs_rng r = bxs_seed(seed, 3);
  const uint64_t side = (uint64_t)floor(extent / voxel);
  for (uint64_t i = 0; i < n; ++i) {
    if (n_ins && bxs_u01(&r) < hit_frac) {
      const uint64_t j = bxs_next(&r) % n_ins;
      memcpy(out + 3 * i, inserted + 3 * j, 3 * sizeof(int32_t));
    } else {
      for (int a = 0; a < 3; ++a) out[3 * i + a] = (int32_t)(bxs_next(&r) % side);
    }
  }
}

/* -------------------------------------------------------- LiDAR world */

typedef struct {
  double lo[3], hi[3];
} bxs_box;

typedef struct {
  uint64_t seed;
  double a1, l1, f1, a2, l2, f2; /* lateral path y(x) = a1 sin(x/l1+f1) + a2 sin(x/l2+f2) */
  uint32_t nboxes;
  bxs_box* boxes;
} bxs_world;

static double bxs_path_y(const bxs_world* w, double x, double* dy) {
  double s1, c1, s2, c2;
  bxs_sincos(x / w->l1 + w->f1, &s1, &c1);
  bxs_sincos(x / w->l2 + w->f2, &s2, &c2);
  if (dy) *dy = w->a1 * c1 / w->l1 + w->a2 * c2 / w->l2;
  return w->a1 * s1 + w->a2 * s2;
}

static void bxs_world_init(bxs_world* w, uint64_t seed, uint32_t nboxes, double xmin, double xmax,
                           double half_width) {
  bxs_rng r = bxs_seed(seed, 4);
  w->seed = seed;
  w->a1 = bxs_uniform(&r, 3.0, 8.0);
  w->l1 = bxs_uniform(&r, 25.0, 45.0);
  w->f1 = bxs_uniform(&r, 0.0, 2.0 * BXS_PI);
  w->a2 = bxs_uniform(&r, 1.0, 3.0);
  w->l2 = bxs_uniform(&r, 8.0, 15.0);
  w->f2 = bxs_uniform(&r, 0.0, 2.0 * BXS_PI);
  w->nboxes = nboxes;
  w->boxes = (bxs_box*)malloc(sizeof(bxs_box) * (nboxes ? nboxes : 1));
  for (uint32_t b = 0; b < nboxes; ++b) {
    bxs_box* bx = &w->boxes[b];
    const double ex = bxs_uniform(&r, 0.5, 5.0), ey = bxs_uniform(&r, 0.5, 5.0),
                 ez = bxs_uniform(&r, 0.5, 5.0);
    const double cx = bxs_uniform(&r, xmin, xmax);
    /* keep a 3 m corridor around the path free of boxes */
    double off = bxs_uniform(&r, 3.0 + 0.5 * ey, half_width);
    if (bxs_u01(&r) < 0.5) off = -off;
    const double cy = bxs_path_y(w, cx, NULL) + off;
    bx->lo[0] = cx - 0.5 * ex;
    bx->hi[0] = cx + 0.5 * ex;
    bx->lo[1] = cy - 0.5 * ey;
    bx->hi[1] = cy + 0.5 * ey;
    bx->lo[2] = 0.0;
    bx->hi[2] = ez;
  }
}

/* Sensor pose at arc length s metres along the path (height h), with heading
 * cos/sin taken from the path tangent. */
static void bxs_pose(const bxs_world* w, double s, double h, double p[3], double* ch, double* sh) {
  /* integrate arc length with 0.05 m steps in x */
  double x = 0.0, acc = 0.0;
  const double dx = 0.05;
  for (;;) {
    double dy;
    bxs_path_y(w, x + 0.5 * dx, &dy);
    const double ds = dx * sqrt(1.0 + dy * dy);
    if (acc + ds >= s) {
      x += dx * (s - acc) / ds;
      break;
    }
    acc += ds;
    x += dx;
  }
  double dy;
  p[0] = x;
  p[1] = bxs_path_y(w, x, &dy);
  p[2] = h;
  const double nrm = sqrt(1.0 + dy * dy);
  *ch = 1.0 / nrm;
  *sh = dy / nrm;
}

/* distance along unit direction d from p to the first surface (ground z=0 or
 * a box), or -1 if nothing is hit within tmax */
static double bxs_cast(const bxs_box* const* cand, uint32_t nc, const double p[3], const double d[3],
                       double tmax) {
  double best = tmax;
  int hit = 0;
  if (d[2] < 0.0) {
    const double t = -p[2] / d[2];
    if (t < best) {
      best = t;
      hit = 1;
    }
  }
  for (uint32_t k = 0; k < nc; ++k) {
    const bxs_box* b = cand[k];
    double t0 = 0.0, t1 = best;
    int ok = 1;
    for (int a = 0; a < 3 && ok; ++a) {
      if (d[a] == 0.0) {
        if (p[a] < b->lo[a] || p[a] > b->hi[a]) ok = 0;
        continue;
      }
      const double inv = 1.0 / d[a];
      double ta = (b->lo[a] - p[a]) * inv, tb = (b->hi[a] - p[a]) * inv;
      if (ta > tb) {
        const double t = ta;
        ta = tb;
        tb = t;
      }
      if (ta > t0) t0 = ta;
      if (tb < t1) t1 = tb;
      if (t0 > t1) ok = 0;
    }
    if (ok && t0 > 0.0 && t0 < best) {
      best = t0;
      hit = 1;
    }
  }
  return hit ? best : -1.0;
}

/* One spinning-LiDAR scan: nbeams elevations linearly spaced in
 * [elev_lo, elev_hi] degrees times nazim azimuths, ray-cast on the world of
 * `world_seed` (ground plane plus nboxes boxes) from the pose at arc length
 * `s` metres.  A return gets N(0, range_sigma) range noise; a ray with no
 * return within cast_range, and a dropout (probability dropout), is emitted
 * at range far_range (beyond the integration max range, so it becomes a
 * clipped free-space ray).  Point i = beam * nazim + azimuth. */
BXS_EXPORT void bxs_lidar_scan(uint64_t world_seed, uint32_t nboxes, uint64_t scan_seed, double s,
                               uint32_t nbeams, uint32_t nazim, double elev_lo, double elev_hi,
                               double cast_range, double far_range, double range_sigma,
                               double dropout, double sensor_h, float* out, double origin[3]) {
  bxs_world w;
  bxs_world_init(&w, world_seed, nboxes, -20.0, 260.0, 45.0);
  bxs_rng r = bxs_seed(scan_seed, 5);
  double p[3], ch, sh;
  bxs_pose(&w, s, sensor_h, p, &ch, &sh);
  origin[0] = p[0];
  origin[1] = p[1];
  origin[2] = p[2];

  /* candidate boxes: footprint within cast_range of the sensor */
  const bxs_box** cand = (const bxs_box**)malloc(sizeof(bxs_box*) * (w.nboxes ? w.nboxes : 1));
  uint32_t nc = 0;
  for (uint32_t b = 0; b < w.nboxes; ++b) {
    const bxs_box* bx = &w.boxes[b];
    const double qx = p[0] < bx->lo[0] ? bx->lo[0] : (p[0] > bx->hi[0] ? bx->hi[0] : p[0]);
    const double qy = p[1] < bx->lo[1] ? bx->lo[1] : (p[1] > bx->hi[1] ? bx->hi[1] : p[1]);
    const double ddx = qx - p[0], ddy = qy - p[1];
    if (ddx * ddx + ddy * ddy <= cast_range * cast_range) cand[nc++] = bx;
  }
  /* per azimuth column, the candidates whose footprint the column's vertical
   * half-plane crosses (a conservative test: some corner on each side, or the
   * box straddles the sensor) */
  const bxs_box*** col = (const bxs_box***)malloc(sizeof(bxs_box**) * nazim);
  uint32_t* ncol = (uint32_t*)calloc(nazim, sizeof(uint32_t));
  const bxs_box** pool = (const bxs_box**)malloc(sizeof(bxs_box*) * ((uint64_t)nazim * (nc ? nc : 1)));
  double* az_c = (double*)malloc(sizeof(double) * nazim);
  double* az_s = (double*)malloc(sizeof(double) * nazim);
  for (uint32_t a = 0; a < nazim; ++a) {
    double sa, ca;
    bxs_sincos(2.0 * BXS_PI * (double)a / (double)nazim, &sa, &ca);
    /* world-frame horizontal direction: rotate by the heading */
    az_c[a] = ca * ch - sa * sh;
    az_s[a] = ca * sh + sa * ch;
    col[a] = pool + (uint64_t)a * (nc ? nc : 1);
    for (uint32_t k = 0; k < nc; ++k) {
      const bxs_box* bx = cand[k];
      int pos = 0, neg = 0, front = 0;
      for (int cx = 0; cx < 2; ++cx)
        for (int cy = 0; cy < 2; ++cy) {
          const double vx = (cx ? bx->hi[0] : bx->lo[0]) - p[0];
          const double vy = (cy ? bx->hi[1] : bx->lo[1]) - p[1];
          const double cr = az_c[a] * vy - az_s[a] * vx;
          const double dt = az_c[a] * vx + az_s[a] * vy;
          pos |= cr >= 0.0;
          neg |= cr <= 0.0;
          front |= dt >= 0.0;
        }
      const int inside = p[0] >= bx->lo[0] && p[0] <= bx->hi[0] && p[1] >= bx->lo[1] && p[1] <= bx->hi[1];
      if ((pos && neg && front) || inside) col[a][ncol[a]++] = bx;
    }
  }
  for (uint32_t e = 0; e < nbeams; ++e) {
    const double el = (elev_lo + (elev_hi - elev_lo) * (nbeams > 1 ? (double)e / (double)(nbeams - 1) : 0.5)) *
                      (BXS_PI / 180.0);
    double se, ce;
    bxs_sincos(el, &se, &ce);
    for (uint32_t a = 0; a < nazim; ++a) {
      const double d[3] = {ce * az_c[a], ce * az_s[a], se};
      double range = bxs_cast(col[a], ncol[a], p, d, cast_range);
      const double drop = bxs_u01(&r);
      const double noise = bxs_normal(&r) * range_sigma;
      if (range < 0.0 || drop < dropout)
        range = far_range;
      else
        range += noise;
      const uint64_t i = (uint64_t)e * nazim + a;
      out[3 * i] = (float)(p[0] + d[0] * range);
      out[3 * i + 1] = (float)(p[1] + d[1] * range);
      out[3 * i + 2] = (float)(p[2] + d[2] * range);
    }
  }
  free(az_c);
  free(az_s);
  free(pool);
  free(ncol);
  free(col);
  free(cand);
  free(w.boxes);
}

/* ------------------------------------------------------- depth frames */

/* A width x height pinhole depth frame (fx, fy, cx, cy) back-projected to
 * world points.  The scene is a closed room with half extents U[2.5, 4] x
 * U[2, 3.5] m and height U[2.5, 3.2] m holding nboxes boxes; the camera sits
 * in the middle half of the room, 1.0-1.8 m high, looking horizontally along
 * the room's x axis away from the nearer end wall, turned by up to 60 deg.  Depth is the ray-cast distance along
 * the pixel ray (every ray ends on a wall, the floor, the ceiling or a box). */
BXS_EXPORT void bxs_depth_frame(uint64_t seed, uint32_t width, uint32_t height, double fx,
                                double fy, double cx, double cy, uint32_t nboxes, float* out,
                                double origin[3]) {
  bxs_rng r = bxs_seed(seed, 6);
  const double hx = bxs_uniform(&r, 2.5, 4.0), hy = bxs_uniform(&r, 2.0, 3.5),
               hz = bxs_uniform(&r, 2.5, 3.2);
  bxs_box* boxes = (bxs_box*)malloc(sizeof(bxs_box) * (nboxes ? nboxes : 1));
  const bxs_box** cand = (const bxs_box**)malloc(sizeof(bxs_box*) * (nboxes ? nboxes : 1));
  for (uint32_t b = 0; b < nboxes; ++b) {
    const double ex = bxs_uniform(&r, 0.2, 1.2), ey = bxs_uniform(&r, 0.2, 1.2),
                 ez = bxs_uniform(&r, 0.2, 1.5);
    const double bx = bxs_uniform(&r, -hx + 0.5 * ex, hx - 0.5 * ex);
    const double by = bxs_uniform(&r, -hy + 0.5 * ey, hy - 0.5 * ey);
    boxes[b].lo[0] = bx - 0.5 * ex;
    boxes[b].hi[0] = bx + 0.5 * ex;
    boxes[b].lo[1] = by - 0.5 * ey;
    boxes[b].hi[1] = by + 0.5 * ey;
    boxes[b].lo[2] = 0.0;
    boxes[b].hi[2] = ez;
  }
  double p[3];
  uint32_t nc = 0;
  for (int tries = 0;; ++tries) {
    p[0] = bxs_uniform(&r, -0.5 * hx, 0.5 * hx);
    p[1] = bxs_uniform(&r, -0.5 * hy, 0.5 * hy);
    p[2] = bxs_uniform(&r, 1.0, 1.8);
    int inside = 0;
    for (uint32_t b = 0; b < nboxes; ++b)
      inside |= p[0] >= boxes[b].lo[0] - 0.2 && p[0] <= boxes[b].hi[0] + 0.2 &&
                p[1] >= boxes[b].lo[1] - 0.2 && p[1] <= boxes[b].hi[1] + 0.2 && p[2] <= boxes[b].hi[2] + 0.2;
    if (!inside || tries > 64) break;
  }
  for (uint32_t b = 0; b < nboxes; ++b) cand[nc++] = &boxes[b];
  /* heading: away from the nearer end wall along the room's long axis,
   * turned by up to +-60 degrees */
  double sy, cyaw, st, ct;
  bxs_sincos(bxs_uniform(&r, -BXS_PI / 3.0, BXS_PI / 3.0), &st, &ct);
  const double ax = p[0] > 0.0 ? -1.0 : 1.0;
  cyaw = ax * ct;
  sy = ax * st;
  origin[0] = p[0];
  origin[1] = p[1];
  origin[2] = p[2];
  /* camera frame: z forward (heading), x right, y down */
  const double fwd[3] = {cyaw, sy, 0.0}, right[3] = {sy, -cyaw, 0.0}, down[3] = {0.0, 0.0, -1.0};
  for (uint32_t v = 0; v < height; ++v) {
    for (uint32_t u = 0; u < width; ++u) {
      const double xc = ((double)u - cx) / fx, yc = ((double)v - cy) / fy;
      double d[3];
      for (int a = 0; a < 3; ++a) d[a] = fwd[a] + xc * right[a] + yc * down[a];
      const double nrm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      for (int a = 0; a < 3; ++a) d[a] /= nrm;
      /* room: the exit distance through the six walls */
      double t = 1e30;
      for (int a = 0; a < 3; ++a) {
        const double lo = a == 0 ? -hx : (a == 1 ? -hy : 0.0);
        const double hi = a == 0 ? hx : (a == 1 ? hy : hz);
        if (d[a] > 0.0) {
          const double ta = (hi - p[a]) / d[a];
          if (ta < t) t = ta;
        } else if (d[a] < 0.0) {
          const double ta = (lo - p[a]) / d[a];
          if (ta < t) t = ta;
        }
      }
      const double tb = bxs_cast(cand, nc, p, d, t);
      if (tb > 0.0 && tb < t) t = tb;
      const uint64_t i = (uint64_t)v * width + u;
      out[3 * i] = (float)(p[0] + d[0] * t);
      out[3 * i + 1] = (float)(p[1] + d[1] * t);
      out[3 * i + 2] = (float)(p[2] + d[2] * t);
    }
  }
  free(cand);
  free(boxes);
}

/* FNV-1a 64 of a byte buffer (input digests for the fixtures) */
BXS_EXPORT uint64_t bxs_fnv1a(const void* data, uint64_t n) {
  const uint8_t* p = (const uint8_t*)data;
  uint64_t h = 0xCBF29CE484222325ull;
  for (uint64_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 0x100000001B3ull;
  }
  return h;
}
