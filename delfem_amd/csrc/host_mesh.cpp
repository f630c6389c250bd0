// Seeded structured meshes and counter-based random draws (host side; synthetic inputs of the benchmarks).
//
// RNG: SplitMix64 finaliser mix(); u64(seed, stream, idx) = mix(mix(seed ^ stream*G) + (idx+1)*G).
// Drawn on the host only and uploaded, so that every consumer sees bit-identical inputs.
#include <cmath>

#include "dfm_internal.hpp"

namespace dfm {

static const uint64_t kG = 0x9E3779B97F4A7C15ull;

static inline uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t idx) {
  const uint64_t k = mix(seed ^ (stream * kG));
  return mix(k + (idx + 1) * kG);
}

double rng_u01(uint64_t seed, uint64_t stream, uint64_t idx) {
  return (double)(rng_u64(seed, stream, idx) >> 11) * 0x1.0p-53;
}

static double rng_u01_open(uint64_t seed, uint64_t stream, uint64_t idx) {
  return ((double)(rng_u64(seed, stream, idx) >> 11) + 1.0) * 0x1.0p-53;
}

double rng_normal(uint64_t seed, uint64_t stream, uint64_t idx) {
  const double a = rng_u01_open(seed, stream, 2 * idx);
  const double b = rng_u01(seed, stream, 2 * idx + 1);
  return std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * M_PI * b);
}

}  // namespace dfm

using namespace dfm;

extern "C" int dfm_rng_uniform(unsigned long long seed, unsigned long long stream, size_t n, double* out) {
  DFM_CHECK(out || n == 0, DFM_ERR_ARG, "dfm_rng_uniform: null output");
  for (size_t i = 0; i < n; i++) out[i] = rng_u01(seed, stream, i);
  return DFM_OK;
}

extern "C" int dfm_rng_normal(unsigned long long seed, unsigned long long stream, size_t n, double* out) {
  DFM_CHECK(out || n == 0, DFM_ERR_ARG, "dfm_rng_normal: null output");
  for (size_t i = 0; i < n; i++) out[i] = rng_normal(seed, stream, i);
  return DFM_OK;
}

extern "C" int dfm_mesh_rect_tri(int nx, int ny, double lx, double ly, unsigned long long seed, double jitter,
                                 double* coords, int* tri) {
  DFM_CHECK(nx > 0 && ny > 0 && coords && tri, DFM_ERR_ARG, "dfm_mesh_rect_tri: bad arguments");
  const double hx = lx / nx, hy = ly / ny;
  for (int iy = 0; iy <= ny; iy++) {
    for (int ix = 0; ix <= nx; ix++) {
      const long long id = (long long)iy * (nx + 1) + ix;
      double x = (double)ix * hx, y = (double)iy * hy;
      const bool inner = ix > 0 && ix < nx && iy > 0 && iy < ny;
      if (inner && jitter != 0.0) {
        x = x + (rng_u01(seed, 1, id) - 0.5) * (2.0 * jitter) * hx;
        y = y + (rng_u01(seed, 2, id) - 0.5) * (2.0 * jitter) * hy;
      }
      coords[id * 2 + 0] = x;
      coords[id * 2 + 1] = y;
    }
  }
  for (int qy = 0; qy < ny; qy++) {
    for (int qx = 0; qx < nx; qx++) {
      const long long q = (long long)qy * nx + qx;
      const int n00 = qy * (nx + 1) + qx, n10 = n00 + 1, n01 = n00 + (nx + 1), n11 = n01 + 1;
      int* t = &tri[q * 6];
      t[0] = n00; t[1] = n10; t[2] = n11;
      t[3] = n00; t[4] = n11; t[5] = n01;
    }
  }
  return DFM_OK;
}

extern "C" int dfm_mesh_box_tet(int nx, int ny, int nz, double lx, double ly, double lz,
                                unsigned long long seed, double jitter, double* coords, int* tet) {
  DFM_CHECK(nx > 0 && ny > 0 && nz > 0 && coords && tet, DFM_ERR_ARG, "dfm_mesh_box_tet: bad arguments");
  const double h[3] = {lx / nx, ly / ny, lz / nz};
  for (int iz = 0; iz <= nz; iz++)
    for (int iy = 0; iy <= ny; iy++)
      for (int ix = 0; ix <= nx; ix++) {
        const long long id = ((long long)iz * (ny + 1) + iy) * (nx + 1) + ix;
        double x[3] = {(double)ix * h[0], (double)iy * h[1], (double)iz * h[2]};
        const bool inner = ix > 0 && ix < nx && iy > 0 && iy < ny && iz > 0 && iz < nz;
        if (inner && jitter != 0.0)
          for (int d = 0; d < 3; d++) x[d] = x[d] + (rng_u01(seed, 1 + d, id) - 0.5) * (2.0 * jitter) * h[d];
        for (int d = 0; d < 3; d++) coords[id * 3 + d] = x[d];
      }
  // six tetrahedra per hex along the paths (0,0,0) -> (1,1,1) of the axis permutations; odd permutations
  // swap their last two nodes so that every volume is positive
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  static const bool odd[6] = {false, true, true, false, false, true};
  auto nid = [&](int x, int y, int z) { return (z * (ny + 1) + y) * (nx + 1) + x; };
  for (int hz = 0; hz < nz; hz++)
    for (int hy = 0; hy < ny; hy++)
      for (int hx = 0; hx < nx; hx++) {
        const long long hidx = ((long long)hz * ny + hy) * nx + hx;
        for (int t = 0; t < 6; t++) {
          int c[3] = {0, 0, 0};
          int path[4];
          path[0] = nid(hx, hy, hz);
          for (int k = 0; k < 2; k++) {
            c[perms[t][k]] = 1;
            path[k + 1] = nid(hx + c[0], hy + c[1], hz + c[2]);
          }
          path[3] = nid(hx + 1, hy + 1, hz + 1);
          if (odd[t]) { const int tmp = path[2]; path[2] = path[3]; path[3] = tmp; }
          int* o = &tet[(hidx * 6 + t) * 4];
          for (int k = 0; k < 4; k++) o[k] = path[k];
        }
      }
  return DFM_OK;
}
