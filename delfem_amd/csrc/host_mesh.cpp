// Seeded structured meshes and counter-based random draws (host side; synthetic inputs of the benchmarks).
//
// RNG: SplitMix64 finaliser mix(); u64(seed, stream, idx) = mix(mix(seed ^ stream*G) + (idx+1)*G).
// Drawn on the host only and uploaded, so that every consumer sees bit-identical inputs.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

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

// GiD .msh reader (CMesher2D::ReadFromFile_GiDMsh, src/msh/mesher2d.cpp:2463-2620; Msh::CMesh3D::ReadFromFile_GiDMsh,
// src/msh/mesh3d.cpp:601-760): a header "MESH dimension D ElemType T Nnode K", "Coordinates", numbered node lines
// up to "end coordinates", a blank line, "Elements", numbered element lines (1-based nodes) up to "end elements".
// Triangles: with D == 3 the points must lie in a plane (the smallest per-axis variance at most 1e-4 of their sum,
// :2507-2515) and keep the two axes that follow the dropped one cyclically (:2516-2519).
extern "C" int dfm_mesh_read_gid(const char* path, int* ndim, int* nnode, int* nelem, int* npoel, double* coords,
                                 int* lnods) {
  using dfm::set_error;
  if (!path || !ndim || !nnode || !nelem || !npoel) {
    set_error("dfm_mesh_read_gid: null argument");
    return DFM_ERR_ARG;
  }
  FILE* fp = std::fopen(path, "r");
  if (!fp) {
    set_error(std::string("dfm_mesh_read_gid: cannot open ") + path);
    return DFM_ERR_ARG;
  }
  auto fail = [&](const char* why) {
    std::fclose(fp);
    set_error(std::string("dfm_mesh_read_gid: ") + why);
    return DFM_ERR_ARG;
  };
  char line[512], w1[64], w2[64], w3[64], type[64], w5[64];
  int dim = 0, nnoel = 0;
  if (!std::fgets(line, sizeof line, fp) ||
      std::sscanf(line, "%63s %63s %d %63s %63s %63s %d", w1, w2, &dim, w3, type, w5, &nnoel) != 7)
    return fail("bad header");
  const bool tri = std::strcmp(type, "Triangle") == 0, tet = std::strcmp(type, "Tetrahedra") == 0;
  if (!(tri && nnoel == 3) && !(tet && nnoel == 4)) return fail("only P1 triangles and tetrahedra are supported");
  if (tet && dim != 3) return fail("tetrahedra need dimension 3");
  if (dim != 2 && dim != 3) return fail("bad dimension");
  if (!std::fgets(line, sizeof line, fp)) return fail("missing Coordinates");
  std::vector<double> xyz;
  for (int count = 0;; count++) {
    if (!std::fgets(line, sizeof line, fp)) return fail("missing end coordinates");
    if (std::strncmp(line, "end coordinates", 15) == 0) break;
    int id = 0;
    double c[3] = {0.0, 0.0, 0.0};
    const int got = std::sscanf(line, "%d%lf%lf%lf", &id, &c[0], &c[1], &c[2]);
    if (got < 1 + dim || id != count + 1) return fail("bad coordinate line");
    xyz.insert(xyz.end(), c, c + 3);
  }
  const int nn = (int)(xyz.size() / 3);
  if (nn == 0) return fail("no nodes");
  // triangles: the plane's two axes
  int a0 = 0, a1 = 1;
  if (tri && dim == 3) {
    double m[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    for (int i = 0; i < nn; i++)
      for (int d = 0; d < 3; d++) {
        m[d] += xyz[(size_t)i * 3 + d];
        v[d] += xyz[(size_t)i * 3 + d] * xyz[(size_t)i * 3 + d];
      }
    for (int d = 0; d < 3; d++) {
      m[d] /= nn;
      v[d] = v[d] / nn - m[d] * m[d];
    }
    int dmin = 0;
    if (v[1] < v[dmin]) dmin = 1;
    if (v[2] < v[dmin]) dmin = 2;
    if (v[dmin] > (v[0] + v[1] + v[2]) * 0.0001) return fail("triangles not in a plane");
    a0 = (dmin + 1) % 3;
    a1 = (dmin + 2) % 3;
  }
  if (!std::fgets(line, sizeof line, fp) || !std::fgets(line, sizeof line, fp)) return fail("missing Elements");
  std::vector<int> con;
  for (int count = 0;; count++) {
    if (!std::fgets(line, sizeof line, fp)) return fail("missing end elements");
    if (std::strncmp(line, "end elements", 12) == 0) break;
    int id = 0, v[4] = {0, 0, 0, 0};
    const int got = std::sscanf(line, "%d%d%d%d%d", &id, &v[0], &v[1], &v[2], &v[3]);
    if (got < 1 + nnoel || id != count + 1) return fail("bad element line");
    for (int k = 0; k < nnoel; k++) {
      if (v[k] < 1 || v[k] > nn) return fail("element node out of range");
      con.push_back(v[k] - 1);
    }
  }
  std::fclose(fp);
  const int outdim = tri ? 2 : 3;
  *ndim = outdim;
  *nnode = nn;
  *nelem = (int)(con.size() / nnoel);
  *npoel = nnoel;
  if (coords) {
    for (int i = 0; i < nn; i++) {
      if (tri) {
        coords[(size_t)i * 2] = xyz[(size_t)i * 3 + a0];
        coords[(size_t)i * 2 + 1] = xyz[(size_t)i * 3 + a1];
      } else {
        for (int d = 0; d < 3; d++) coords[(size_t)i * 3 + d] = xyz[(size_t)i * 3 + d];
      }
    }
  }
  if (lnods) std::memcpy(lnods, con.data(), sizeof(int) * con.size());
  return DFM_OK;
}
