// Seeded synthetic inputs for tests and benchmarks (host C++, built with g++).
//
// Scene: ground plane z = 0 over a square of side `extent` centred on the
// origin, axis-aligned box buildings (width U(5,30) m per axis, height
// U(4,20) m, uniform placement), vertical poles (radius 0.15 m, height
// U(4,8) m).  Map points are sampled uniformly by surface area (ground under
// building footprints excluded) with isotropic Gaussian noise.  Scans come from
// a ring lidar ray-cast against the same primitives, in the sensor frame,
// followed by a centroid voxel downsample (the reference's voxelize_pcd,
// src/ndt_cuda/src/test_utils.rs:76-110, with output sorted by cell key so
// the result is deterministic).
//
// Every random draw is a pure function of (seed, stream, counter) through
// SplitMix64, so results do not depend on thread count or call order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Rng {
  uint64_t seed, stream, ctr = 0;
  Rng(uint64_t s, uint64_t st) : seed(s), stream(st) {}
  uint64_t next() { return splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + (ctr++)); }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double normal() {  // Box-Muller
    double u1 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

struct Box {
  double x0, y0, x1, y1, h;
};
struct Pole {
  double x, y, r, h;
};

struct Scene {
  double extent;
  std::vector<Box> boxes;
  std::vector<Pole> poles;
  // 2D bucket grid over building footprints (for ground rejection)
  double cell = 32.0;
  int gdim = 1;
  std::vector<std::vector<int>> buckets;
};

bool inside_any_box(const Scene& s, double x, double y) {
  const double half = s.extent * 0.5;
  const int cx = std::min(s.gdim - 1, std::max(0, (int)((x + half) / s.cell)));
  const int cy = std::min(s.gdim - 1, std::max(0, (int)((y + half) / s.cell)));
  for (int bi : s.buckets[(size_t)cy * s.gdim + cx]) {
    const Box& b = s.boxes[bi];
    if (x >= b.x0 && x <= b.x1 && y >= b.y0 && y <= b.y1) return true;
  }
  return false;
}

// Ray (o + t d), returns nearest t in (0, tmax) or -1.
double ray_box(const double* o, const double* d, const Box& b, double tmax) {
  const double lo[3] = {b.x0, b.y0, 0.0}, hi[3] = {b.x1, b.y1, b.h};
  double t0 = 0.0, t1 = tmax;
  for (int a = 0; a < 3; ++a) {
    if (std::fabs(d[a]) < 1e-12) {
      if (o[a] < lo[a] || o[a] > hi[a]) return -1.0;
      continue;
    }
    double ta = (lo[a] - o[a]) / d[a], tb = (hi[a] - o[a]) / d[a];
    if (ta > tb) std::swap(ta, tb);
    t0 = std::max(t0, ta);
    t1 = std::min(t1, tb);
    if (t0 > t1) return -1.0;
  }
  return t0 > 1e-6 ? t0 : -1.0;
}

double ray_pole(const double* o, const double* d, const Pole& p, double tmax) {
  const double ox = o[0] - p.x, oy = o[1] - p.y;
  const double a = d[0] * d[0] + d[1] * d[1];
  if (a < 1e-12) return -1.0;
  const double b = 2.0 * (ox * d[0] + oy * d[1]);
  const double c = ox * ox + oy * oy - p.r * p.r;
  const double disc = b * b - 4 * a * c;
  if (disc < 0) return -1.0;
  const double t = (-b - std::sqrt(disc)) / (2 * a);
  if (t <= 1e-6 || t >= tmax) return -1.0;
  const double z = o[2] + t * d[2];
  if (z < 0.0 || z > p.h) return -1.0;
  return t;
}

}  // namespace

extern "C" {

void* synth_scene_create(uint64_t seed, double extent, uint32_t n_boxes, uint32_t n_poles) {
  auto* s = new Scene();
  s->extent = extent;
  const double half = extent * 0.5;
  Rng rb(seed, 1), rp(seed, 2);
  for (uint32_t i = 0; i < n_boxes; ++i) {
    const double wx = rb.uniform(5.0, 30.0), wy = rb.uniform(5.0, 30.0), h = rb.uniform(4.0, 20.0);
    const double cx = rb.uniform(-half + wx * 0.5, half - wx * 0.5), cy = rb.uniform(-half + wy * 0.5, half - wy * 0.5);
    s->boxes.push_back(Box{cx - wx * 0.5, cy - wy * 0.5, cx + wx * 0.5, cy + wy * 0.5, h});
  }
  for (uint32_t i = 0; i < n_poles; ++i) {
    const double x = rp.uniform(-half, half), y = rp.uniform(-half, half);
    s->poles.push_back(Pole{x, y, 0.15, rp.uniform(4.0, 8.0)});
  }
  s->gdim = std::max(1, (int)std::ceil(extent / s->cell));
  s->buckets.assign((size_t)s->gdim * s->gdim, {});
  for (size_t bi = 0; bi < s->boxes.size(); ++bi) {
    const Box& b = s->boxes[bi];
    const int x0 = std::max(0, (int)((b.x0 + half) / s->cell)), x1 = std::min(s->gdim - 1, (int)((b.x1 + half) / s->cell));
    const int y0 = std::max(0, (int)((b.y0 + half) / s->cell)), y1 = std::min(s->gdim - 1, (int)((b.y1 + half) / s->cell));
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) s->buckets[(size_t)y * s->gdim + x].push_back((int)bi);
  }
  return s;
}

void synth_scene_destroy(void* h) { delete static_cast<Scene*>(h); }

// Samples n map points; returns the count written (== n).
size_t synth_scene_sample(void* h, uint64_t seed, size_t n, double noise, float* out) {
  const Scene& s = *static_cast<Scene*>(h);
  const double half = s.extent * 0.5;
  // Surface table: 0 = ground, then 5 faces per box (4 walls + roof), then poles.
  std::vector<double> area;
  area.push_back(s.extent * s.extent);
  for (const Box& b : s.boxes) {
    const double wx = b.x1 - b.x0, wy = b.y1 - b.y0;
    area.push_back(wx * b.h);
    area.push_back(wx * b.h);
    area.push_back(wy * b.h);
    area.push_back(wy * b.h);
    area.push_back(wx * wy);
  }
  for (const Pole& p : s.poles) area.push_back(2.0 * M_PI * p.r * p.h);
  std::vector<double> cum(area.size());
  double tot = 0.0;
  for (size_t i = 0; i < area.size(); ++i) cum[i] = (tot += area[i]);
  const size_t nb = s.boxes.size();
  size_t k = 0;
  for (uint64_t i = 0; k < n; ++i) {
    Rng r(seed, 1000 + i);
    const double u = r.uniform() * tot;
    const size_t surf = std::upper_bound(cum.begin(), cum.end(), u) - cum.begin();
    double x, y, z;
    if (surf == 0) {
      x = r.uniform(-half, half);
      y = r.uniform(-half, half);
      z = 0.0;
      if (inside_any_box(s, x, y)) continue;  // ground hidden under a building
    } else if (surf <= 5 * nb) {
      const Box& b = s.boxes[(surf - 1) / 5];
      const int f = (int)((surf - 1) % 5);
      const double a = r.uniform(), c = r.uniform();
      if (f == 0) { x = b.x0 + a * (b.x1 - b.x0); y = b.y0; z = c * b.h; }
      else if (f == 1) { x = b.x0 + a * (b.x1 - b.x0); y = b.y1; z = c * b.h; }
      else if (f == 2) { x = b.x0; y = b.y0 + a * (b.y1 - b.y0); z = c * b.h; }
      else if (f == 3) { x = b.x1; y = b.y0 + a * (b.y1 - b.y0); z = c * b.h; }
      else { x = b.x0 + a * (b.x1 - b.x0); y = b.y0 + c * (b.y1 - b.y0); z = b.h; }
    } else {
      const Pole& p = s.poles[surf - 1 - 5 * nb];
      const double th = r.uniform(0.0, 2.0 * M_PI);
      x = p.x + p.r * std::cos(th);
      y = p.y + p.r * std::sin(th);
      z = r.uniform() * p.h;
    }
    out[3 * k + 0] = (float)(x + noise * r.normal());
    out[3 * k + 1] = (float)(y + noise * r.normal());
    out[3 * k + 2] = (float)(z + noise * r.normal());
    ++k;
  }
  return k;
}

// Ring lidar at sensor pose (x, y, z, yaw): `beams` elevation angles evenly
// spaced over [vmin, vmax] degrees, `az_steps` azimuths over 360 degrees.
// Writes hits in the sensor frame; returns the count (<= beams * az_steps).
size_t synth_scene_lidar(void* h, uint64_t seed, const double* pose4, uint32_t beams, double vmin_deg, double vmax_deg,
                         uint32_t az_steps, double max_range, double noise, float* out) {
  const Scene& s = *static_cast<Scene*>(h);
  const double sx = pose4[0], sy = pose4[1], sz = pose4[2], yaw = pose4[3];
  const double cyw = std::cos(yaw), syw = std::sin(yaw);
  std::vector<int> near_boxes, near_poles;
  for (size_t i = 0; i < s.boxes.size(); ++i) {
    const Box& b = s.boxes[i];
    const double dx = std::max({b.x0 - sx, 0.0, sx - b.x1}), dy = std::max({b.y0 - sy, 0.0, sy - b.y1});
    if (dx * dx + dy * dy <= max_range * max_range) near_boxes.push_back((int)i);
  }
  for (size_t i = 0; i < s.poles.size(); ++i) {
    const double dx = s.poles[i].x - sx, dy = s.poles[i].y - sy;
    if (dx * dx + dy * dy <= (max_range + 1.0) * (max_range + 1.0)) near_poles.push_back((int)i);
  }
  size_t k = 0;
  const double o[3] = {sx, sy, sz};
  for (uint32_t bi = 0; bi < beams; ++bi) {
    const double el = (beams > 1 ? vmin_deg + (vmax_deg - vmin_deg) * bi / (beams - 1) : vmin_deg) * M_PI / 180.0;
    for (uint32_t ai = 0; ai < az_steps; ++ai) {
      const double az = 2.0 * M_PI * ai / az_steps;
      // direction in the sensor frame, then rotated by yaw into the map frame
      const double ls[3] = {std::cos(el) * std::cos(az), std::cos(el) * std::sin(az), std::sin(el)};
      const double d[3] = {cyw * ls[0] - syw * ls[1], syw * ls[0] + cyw * ls[1], ls[2]};
      double best = max_range;
      bool hit = false;
      if (d[2] < -1e-9) {
        const double t = -sz / d[2];
        if (t > 0 && t < best) {
          const double gx = sx + t * d[0], gy = sy + t * d[1];
          const double half = s.extent * 0.5;
          if (std::fabs(gx) <= half && std::fabs(gy) <= half) {
            best = t;
            hit = true;
          }
        }
      }
      for (int i : near_boxes) {
        const double t = ray_box(o, d, s.boxes[i], best);
        if (t > 0 && t < best) {
          best = t;
          hit = true;
        }
      }
      for (int i : near_poles) {
        const double t = ray_pole(o, d, s.poles[i], best);
        if (t > 0 && t < best) {
          best = t;
          hit = true;
        }
      }
      if (!hit) continue;
      Rng r(seed, ((uint64_t)bi << 32) | ai);
      const double range = best + noise * r.normal();
      out[3 * k + 0] = (float)(range * ls[0]);
      out[3 * k + 1] = (float)(range * ls[1]);
      out[3 * k + 2] = (float)(range * ls[2]);
      ++k;
    }
  }
  return k;
}

// Centroid voxel downsample (test_utils.rs:76-110: key = floor(p * (1/leaf))
// in f32, f64 sums, skip non-finite points); output sorted by key.
size_t synth_voxel_downsample(const float* pts, size_t n, float leaf, float* out) {
  struct Acc {
    double x = 0, y = 0, z = 0;
    size_t c = 0;
  };
  const float inv = 1.0f / leaf;
  std::unordered_map<uint64_t, Acc> m;
  m.reserve(n / 2 + 16);
  for (size_t i = 0; i < n; ++i) {
    const float* p = pts + 3 * i;
    if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) continue;
    const int32_t k[3] = {(int32_t)std::floor(p[0] * inv), (int32_t)std::floor(p[1] * inv),
                          (int32_t)std::floor(p[2] * inv)};
    const uint64_t key = ((uint64_t)(uint32_t)(k[0] + (1 << 20)) << 42) | ((uint64_t)(uint32_t)(k[1] + (1 << 20)) << 21) |
                         (uint64_t)(uint32_t)(k[2] + (1 << 20));
    Acc& a = m[key];
    a.x += p[0];
    a.y += p[1];
    a.z += p[2];
    a.c += 1;
  }
  std::vector<uint64_t> ks;
  ks.reserve(m.size());
  for (auto& kv : m) ks.push_back(kv.first);
  std::sort(ks.begin(), ks.end());
  size_t j = 0;
  for (uint64_t k : ks) {
    const Acc& a = m[k];
    const double c = (double)a.c;
    out[3 * j + 0] = (float)(a.x / c);
    out[3 * j + 1] = (float)(a.y / c);
    out[3 * j + 2] = (float)(a.z / c);
    ++j;
  }
  return j;
}

}  // extern "C"
