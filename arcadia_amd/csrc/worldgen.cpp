// Seeded synthetic worlds and ray sets for the traversal workloads (C ABI in arcadia_world.h).
//
// Worlds are value-noise fBm terrains in the style of the reference generator
// (crates/arcadia-world/src/generator.rs:656-694: smoothstep-faded value noise, amplitude
// halving per octave) with the seed folded into the lattice hash. All generator arithmetic
// is binary64 with no transcendental functions, so a world is bit-identical on every host
// that runs this code. Rays are binary32 and follow the reference's formulas.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "world_internal.h"

namespace arc {
namespace {

inline uint32_t hash32(uint32_t x) {  // integer avalanche hash (lowbias32 constants)
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint32_t seed32(uint64_t seed, uint32_t salt) {
  return (uint32_t)splitmix64(seed * 0x100000001B3ull + salt);
}

inline uint32_t lattice_hash(uint32_t s, int32_t x, int32_t y, int32_t z) {
  return hash32(((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349663u) ^ ((uint32_t)z * 83492791u) ^ s);
}

inline double u01(uint32_t h) { return (double)h * (1.0 / 4294967296.0); }

inline double fade(double t) { return t * t * (3.0 - 2.0 * t); }

// 2D value noise at (x, z) in lattice units.
inline double vnoise2(uint32_t s, double x, double z) {
  double fx = std::floor(x), fz = std::floor(z);
  int32_t ix = (int32_t)fx, iz = (int32_t)fz;
  double ux = fade(x - fx), uz = fade(z - fz);
  double v00 = u01(lattice_hash(s, ix, 0, iz)), v10 = u01(lattice_hash(s, ix + 1, 0, iz));
  double v01 = u01(lattice_hash(s, ix, 0, iz + 1)), v11 = u01(lattice_hash(s, ix + 1, 0, iz + 1));
  double a = v00 + ux * (v10 - v00), b = v01 + ux * (v11 - v01);
  return a + uz * (b - a);
}

// fBm in [0, 1): octave k has wavelength base/2^k and amplitude 0.5^k, normalised.
double fbm2(uint64_t seed, uint32_t octaves, double wavelength, double x, double z) {
  double sum = 0.0, norm = 0.0, amp = 1.0, lam = wavelength;
  for (uint32_t k = 0; k < octaves; ++k) {
    sum += amp * vnoise2(seed32(seed, 100 + k), x / lam, z / lam);
    norm += amp;
    amp *= 0.5;
    lam *= 0.5;
  }
  return sum / norm;
}

struct Gen {
  arc_gen_params p;
  uint32_t grid[3];
  struct Cuboid {
    uint32_t x0, y0, z0, x1, y1, z1;  // [x0, x1)
    uint8_t mat;
  };
  std::vector<Cuboid> cuboids;

  void init(const arc_gen_params& prm, const uint32_t g[3]) {
    p = prm;
    std::memcpy(grid, g, sizeof grid);
    cuboids.clear();
    if (p.kind == ARC_GEN_CAVES) {
      uint64_t st = splitmix64(p.seed ^ 0xC0B01D5ull);
      auto next = [&]() { st = splitmix64(st); return st; };
      uint32_t span = p.cuboid_max - p.cuboid_min + 1u;
      uint32_t mspan = p.mat_max - p.mat_min + 1u;
      for (uint32_t i = 0; i < p.n_cuboids; ++i) {
        uint32_t e[3], o[3];
        for (int a = 0; a < 3; ++a) e[a] = p.cuboid_min + (uint32_t)(next() % span);
        for (int a = 0; a < 3; ++a) {
          uint32_t dim = grid[a] * 256u;
          o[a] = (uint32_t)(next() % (uint64_t)(dim - e[a] + 1u));
        }
        uint8_t m = (uint8_t)(p.mat_min + (uint32_t)(next() % mspan));
        cuboids.push_back({o[0], o[1], o[2], o[0] + e[0], o[1] + e[1], o[2] + e[2], m});
      }
    }
  }

  inline uint8_t rand_mat(uint32_t x, uint32_t y, uint32_t z) const {
    uint32_t span = p.mat_max - p.mat_min + 1u;
    uint32_t h = hash32(lattice_hash(seed32(p.seed, 7), (int32_t)x, (int32_t)y, (int32_t)z));
    return (uint8_t)(p.mat_min + h % span);
  }

  double height(uint32_t x, uint32_t z) const {
    return fbm2(p.seed, p.octaves, p.wavelength, (double)x, (double)z);
  }

  // Caves density band: solid iff fbm3 + (hs - y)/band > threshold.
  static constexpr double kBand = 192.0;

  // Fill one chunk's dense array (x fastest).
  void fill_chunk(uint32_t cx, uint32_t cy, uint32_t cz, uint8_t* dense, std::vector<double>& scratch) const {
    const uint32_t X0 = cx * 256u, Y0 = cy * 256u, Z0 = cz * 256u;
    std::memset(dense, 0, (size_t)256 * 256 * 256);
    if (p.kind == ARC_GEN_HEIGHTMAP || p.kind == ARC_GEN_SPARSE) {
      double span = (double)p.height_max - (double)p.height_min;
      for (uint32_t lz = 0; lz < 256; ++lz)
        for (uint32_t lx = 0; lx < 256; ++lx) {
          double f = height(X0 + lx, Z0 + lz);
          int64_t h = (int64_t)p.height_min + (int64_t)std::floor(span * f);
          for (uint32_t ly = 0; ly < 256; ++ly) {
            int64_t y = (int64_t)Y0 + ly;
            if (y >= h) break;
            dense[((size_t)lz * 256 + ly) * 256 + lx] = rand_mat(X0 + lx, (uint32_t)y, Z0 + lz);
          }
        }
    } else {
      fill_caves(X0, Y0, Z0, dense, scratch);
    }
    for (const Cuboid& c : cuboids) {
      uint32_t x0 = std::max(c.x0, X0), x1 = std::min(c.x1, X0 + 256u);
      uint32_t y0 = std::max(c.y0, Y0), y1 = std::min(c.y1, Y0 + 256u);
      uint32_t z0 = std::max(c.z0, Z0), z1 = std::min(c.z1, Z0 + 256u);
      if (x0 >= x1 || y0 >= y1 || z0 >= z1) continue;
      for (uint32_t z = z0; z < z1; ++z)
        for (uint32_t y = y0; y < y1; ++y)
          std::memset(&dense[((size_t)(z - Z0) * 256 + (y - Y0)) * 256 + (x0 - X0)], c.mat, x1 - x0);
    }
  }

  void fill_caves(uint32_t X0, uint32_t Y0, uint32_t Z0, uint8_t* dense, std::vector<double>& lat) const {
    const double span = (double)p.height_max - (double)p.height_min;
    const double thr = p.density_threshold;
    // surface height per column
    std::vector<double> hs(256 * 256);
    double hmin = 1e30, hmax = -1e30;
    for (uint32_t lz = 0; lz < 256; ++lz)
      for (uint32_t lx = 0; lx < 256; ++lx) {
        double h = (double)p.height_min + span * height(X0 + lx, Z0 + lz);
        hs[lz * 256 + lx] = h;
        hmin = std::min(hmin, h);
        hmax = std::max(hmax, h);
      }
    // y range where the 3D noise decides: (hs - y)/band in (thr - 1, thr]
    const double y_solid_below = hmin - thr * kBand;          // y < this: density > thr always
    const double y_air_above = hmax + (1.0 - thr) * kBand;    // y >= this: never solid
    int64_t ya = (int64_t)std::floor(y_solid_below) - 1, yb = (int64_t)std::ceil(y_air_above) + 1;
    (void)ya;
    int64_t ly_hi = std::min<int64_t>(256, yb - (int64_t)Y0);
    // 3D fBm lattice per octave (wavelength 128 / 2^k voxels: divides the chunk edge)
    const uint32_t oct3 = 4;
    const uint32_t lam0 = 128;
    struct Oct { uint32_t lam, n; size_t off; double amp; };
    Oct oc[8];
    size_t total = 0;
    double norm = 0.0, amp = 1.0;
    for (uint32_t k = 0; k < oct3; ++k) {
      uint32_t lam = lam0 >> k;
      uint32_t n = 256u / lam + 3u;
      oc[k] = {lam, n, total, amp};
      total += (size_t)n * n * n;
      norm += amp;
      amp *= 0.5;
    }
    lat.resize(total);
    for (uint32_t k = 0; k < oct3; ++k) {
      uint32_t s = seed32(p.seed, 300 + k);
      int32_t bx = (int32_t)(X0 / oc[k].lam), by = (int32_t)(Y0 / oc[k].lam), bz = (int32_t)(Z0 / oc[k].lam);
      uint32_t n = oc[k].n;
      for (uint32_t k2 = 0; k2 < n; ++k2)
        for (uint32_t j = 0; j < n; ++j)
          for (uint32_t i = 0; i < n; ++i)
            lat[oc[k].off + ((size_t)k2 * n + j) * n + i] =
                u01(lattice_hash(s, bx + (int32_t)i, by + (int32_t)j, bz + (int32_t)k2));
    }
    for (int64_t ly = 0; ly < 256; ++ly) {
      const int64_t y = (int64_t)Y0 + ly;
      for (uint32_t lz = 0; lz < 256; ++lz) {
        for (uint32_t lx = 0; lx < 256; ++lx) {
          const double h = hs[lz * 256 + lx];
          const double base = (h - (double)y) / kBand;
          bool solid;
          if (ly >= ly_hi) solid = false;                     // above every column's band
          else if (base > thr) solid = true;                  // fbm3 >= 0
          else if (base <= thr - 1.0) solid = false;          // fbm3 < 1
          else {
            double f = 0.0;
            const uint32_t x = X0 + lx, yy = (uint32_t)y, z = Z0 + lz;
            for (uint32_t k = 0; k < oct3; ++k) {
              const uint32_t lam = oc[k].lam;
              const uint32_t gx = x / lam - X0 / lam, gy = yy / lam - Y0 / lam, gz = z / lam - Z0 / lam;
              const double tx = fade((double)(x % lam) / lam), ty = fade((double)(yy % lam) / lam),
                           tz = fade((double)(z % lam) / lam);
              const uint32_t n = oc[k].n;
              const double* L = &lat[oc[k].off];
              auto v = [&](uint32_t i, uint32_t j, uint32_t k2) { return L[((size_t)k2 * n + j) * n + i]; };
              double c00 = v(gx, gy, gz) + tx * (v(gx + 1, gy, gz) - v(gx, gy, gz));
              double c10 = v(gx, gy + 1, gz) + tx * (v(gx + 1, gy + 1, gz) - v(gx, gy + 1, gz));
              double c01 = v(gx, gy, gz + 1) + tx * (v(gx + 1, gy, gz + 1) - v(gx, gy, gz + 1));
              double c11 = v(gx, gy + 1, gz + 1) + tx * (v(gx + 1, gy + 1, gz + 1) - v(gx, gy + 1, gz + 1));
              double c0 = c00 + ty * (c10 - c00), c1 = c01 + ty * (c11 - c01);
              f += oc[k].amp * (c0 + tz * (c1 - c0));
            }
            f /= norm;
            solid = f + base > thr;
          }
          if (solid) dense[((size_t)lz * 256 + (size_t)ly) * 256 + lx] = rand_mat(X0 + lx, (uint32_t)y, Z0 + lz);
        }
      }
    }
  }

  // Quick per-chunk emptiness test (sky chunks skip the dense fill entirely).
  bool chunk_maybe_solid(uint32_t cx, uint32_t cy, uint32_t cz) const {
    const uint32_t X0 = cx * 256u, Y0 = cy * 256u, Z0 = cz * 256u;
    for (const Cuboid& c : cuboids)
      if (c.x0 < X0 + 256 && c.x1 > X0 && c.y0 < Y0 + 256 && c.y1 > Y0 && c.z0 < Z0 + 256 && c.z1 > Z0) return true;
    double top = (double)p.height_max;
    if (p.kind == ARC_GEN_CAVES) top += (1.0 - p.density_threshold) * kBand + 2.0;
    return (double)Y0 < top;
  }
};

}  // namespace
}  // namespace arc

using namespace arc;

static bool valid_params(const arc_gen_params* p) {
  if (!p) return false;
  if (p->kind > ARC_GEN_SPARSE) return false;
  if (p->octaves == 0 || p->octaves > 12) return false;
  if (!(p->wavelength >= 1.0f)) return false;
  if (!(p->height_max >= p->height_min)) return false;
  if (p->mat_min == 0 || p->mat_max < p->mat_min || p->mat_max > 255) return false;
  if (p->kind == ARC_GEN_CAVES && p->n_cuboids && (p->cuboid_min == 0 || p->cuboid_max < p->cuboid_min))
    return false;
  return true;
}

extern "C" {

int arc_world_generate(arc_world_host* wh, const arc_gen_params* params, int nthreads) {
  World* w = reinterpret_cast<World*>(wh);
  if (!w || !valid_params(params)) return ARC_ERR_INVALID;
  if (params->kind == ARC_GEN_CAVES)
    for (int a = 0; a < 3; ++a)
      if (params->cuboid_max > w->grid[a] * 256u) return ARC_ERR_INVALID;
  Gen g;
  g.init(*params, w->grid);
  const uint32_t n = w->grid[0] * w->grid[1] * w->grid[2];
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  std::atomic<uint32_t> next{0};
  auto work = [&]() {
    std::vector<uint8_t> dense((size_t)256 * 256 * 256);
    std::vector<double> scratch;
    for (;;) {
      uint32_t i = next.fetch_add(1);
      if (i >= n) break;
      uint32_t cx = i % w->grid[0], cy = (i / w->grid[0]) % w->grid[1], cz = i / (w->grid[0] * w->grid[1]);
      if (!g.chunk_maybe_solid(cx, cy, cz)) {
        w->chunks[i] = ChunkTree();
        w->chunks[i].nodes.push_back(arc_node64{0, 0, 0, 0});
        continue;
      }
      g.fill_chunk(cx, cy, cz, dense.data(), scratch);
      build_chunk_dense(dense.data(), w->chunks[i]);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nthreads; ++t) th.emplace_back(work);
  work();
  for (auto& t : th) t.join();
  w->packed = false;
  return ARC_OK;
}

uint32_t arc_gen_voxel(const arc_world_host* wh, const arc_gen_params* params, uint32_t x, uint32_t y, uint32_t z) {
  const World* w = reinterpret_cast<const World*>(wh);
  if (!w || !valid_params(params)) return 0;
  if (x >= w->grid[0] * 256u || y >= w->grid[1] * 256u || z >= w->grid[2] * 256u) return 0;
  Gen g;
  g.init(*params, w->grid);
  std::vector<uint8_t> dense((size_t)256 * 256 * 256);
  std::vector<double> scratch;
  g.fill_chunk(x / 256, y / 256, z / 256, dense.data(), scratch);
  return dense[((size_t)(z % 256) * 256 + (y % 256)) * 256 + (x % 256)];
}

}  // extern "C"
