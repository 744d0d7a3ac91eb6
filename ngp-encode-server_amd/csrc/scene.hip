// scene.hip -- seeded synthetic scene and camera helpers (host code).
//
// The reference's renderer loads a trained instant-ngp snapshot; this library
// takes the scene as three blobs (hash table, MLP weights, occupancy bits).
// For benchmarks and tests the scene is synthesised from a counter-based
// splitmix64 stream so that every consumer (host, device, tests) can
// regenerate it bit for bit.
#include <cmath>
#include <algorithm>
#include <cstring>
#include <thread>

#include "nes_internal.h"

namespace nes {

namespace {
enum : uint64_t {
  kStreamTable = 1,
  kStreamOcc = 2,
  kStreamBlob = 3,
  kStreamWeight0 = 10,
};
constexpr int kLayerIn[5] = {32, 64, 31, 64, 64};
constexpr int kLayerOut[5] = {64, 16, 64, 64, 3};
constexpr int kLayerOff[5] = {kW1, kW2, kW3, kW4, kW5};
// Occupancy: a Gaussian blob exp(-r^2 / (2 s^2)) with s = 0.2, centre jittered
// by +-0.05 per axis, thresholded at 0.115 after a per-cell factor in
// [0.75, 1.25) -- about 30% of the 128^3 cells.
constexpr double kBlobSigma = 0.2;
constexpr double kBlobTau = 0.115;

double normal(uint64_t seed, uint64_t stream, uint64_t i) {
  double u1 = rand_uniform(seed, stream, 2 * i);
  double u2 = rand_uniform(seed, stream, 2 * i + 1);
  return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
}
}  // namespace

}  // namespace nes

using namespace nes;

extern "C" {

int nes_scene_synthesize(uint64_t seed, int table_init, float *table,
                         float *weights, uint32_t *occupancy) {
  if (table_init != 0 && table_init != 1) return NES_EINVAL;
  if (table) {
    const uint64_t n = (uint64_t)kTableEntries * 2;
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t)
      pool.emplace_back([=] {
        for (uint64_t j = n * t / nt; j < n * (t + 1) / nt; ++j) {
          if (table_init == 0)
            table[j] = (float)(0.1 * normal(seed, kStreamTable, j));
          else
            table[j] = (float)((2.0 * rand_uniform(seed, kStreamTable, j) - 1.0) * 1e-4);
        }
      });
    for (auto &th : pool) th.join();
  }
  if (weights) {
    for (int l = 0; l < 5; ++l) {
      const double a = std::sqrt(6.0 / (double)(kLayerIn[l] + kLayerOut[l]));
      const int n = kLayerIn[l] * kLayerOut[l];
      for (int j = 0; j < n; ++j)
        weights[kLayerOff[l] + j] =
            (float)((2.0 * rand_uniform(seed, kStreamWeight0 + l + 1, (uint64_t)j) - 1.0) * a);
    }
  }
  if (occupancy) {
    double c[3];
    for (int a = 0; a < 3; ++a) c[a] = 0.5 + 0.1 * (rand_uniform(seed, kStreamBlob, (uint64_t)a) - 0.5);
    std::memset(occupancy, 0, sizeof(uint32_t) * NES_OCC_WORDS);
    for (int k = 0; k < kOccRes; ++k)
      for (int j = 0; j < kOccRes; ++j)
        for (int i = 0; i < kOccRes; ++i) {
          const uint32_t cell = (uint32_t)(i + kOccRes * (j + kOccRes * k));
          const double x = (i + 0.5) / kOccRes - c[0];
          const double y = (j + 0.5) / kOccRes - c[1];
          const double z = (k + 0.5) / kOccRes - c[2];
          const double rho = std::exp(-(x * x + y * y + z * z) / (2.0 * kBlobSigma * kBlobSigma));
          const double u = rand_uniform(seed, kStreamOcc, cell);
          if (rho * (0.75 + 0.5 * u) > kBlobTau) occupancy[cell >> 5] |= 1u << (cell & 31);
        }
  }
  return NES_OK;
}

int nes_orbit_camera(double azimuth_deg, double elevation_deg, double radius,
                     double eye_offset, float c2w[12]) {
  if (!c2w || !(radius > 0.0)) return NES_EINVAL;
  const double pi = 3.14159265358979323846;
  const double az = azimuth_deg * pi / 180.0, el = elevation_deg * pi / 180.0;
  const double p[3] = {0.5 + radius * std::cos(el) * std::sin(az), 0.5 + radius * std::sin(el),
                       0.5 + radius * std::cos(el) * std::cos(az)};
  double f[3] = {0.5 - p[0], 0.5 - p[1], 0.5 - p[2]};
  const double fn = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
  for (double &x : f) x /= fn;
  double r[3] = {-f[2], 0.0, f[0]};  // forward x up(0, 1, 0)
  const double rn = std::sqrt(r[0] * r[0] + r[2] * r[2]);
  for (double &x : r) x /= rn;
  const double dn[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2],
                        f[0] * r[1] - f[1] * r[0]};  // forward x right
  for (int i = 0; i < 3; ++i) {
    c2w[4 * i + 0] = (float)r[i];
    c2w[4 * i + 1] = (float)dn[i];
    c2w[4 * i + 2] = (float)f[i];
    c2w[4 * i + 3] = (float)(p[i] + eye_offset * r[i]);
  }
  return NES_OK;
}

}  // extern "C"
