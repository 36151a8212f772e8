// field.hip -- gfx950 kernel and host tables of the deformed-field re-bake
// (field.h).
//
// Reference (paths relative to the Talus tree):
//   build_probe     src/talus/deformable/contact/contact.cc:886-907 (cell
//                   keys), 1080-1150 (deformedProbe with an infinite reach:
//                   boundary faces from the rest cells' missing neighbours,
//                   crack faces added), 1152-1285 (placeDeformedProbe: per
//                   cell bounds, centre, centre Jacobian inverse, the padded
//                   uniform grid as a CSR in cell order)
//   invert_hex      contact.cc:909-957
//   evaluate_hex    contact.cc:959-1027 (the distance and crack faces)
//   screened        contact.cc:1029-1065
//   probe_deformed  contact.cc:1287-1425 (the inversions of the first pass
//                   are recomputed, not cached, in the second: invertHex is
//                   a pure function of the cell and the point)
//   sdf_probe       src/talus/field/sampledsdf/sampledsdf.cc:234-266,
//                   293-304, 432-440
//   k_bake          src/talus/deformable/deformedfield/deformedfield.cc:
//                   54-80 (one thread per sample; the changed box by atomic
//                   min/max, which is order-free)
#include "field.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "device.h"

namespace tmi {

namespace {

constexpr int NT = 256;
constexpr double DEFORMED_TOLERANCE = 1.0e-10;  // contact.h:96
constexpr int DEFORMED_ITERATIONS = 8;          // contact.h:97
constexpr double DEFORMED_PAD = 1.5;            // contact.h:98
constexpr double DEFORMED_SCREEN = 0.5;         // contact.h:101
constexpr float UNCHANGED = 1.0e-7f;            // deformedfield.cc:7

struct ProbeDev {
  const uint32_t *hex;
  const double *pos, *rest;
  const uint8_t *boundary, *cracked;
  const double *low, *high, *centre, *inverse;
  const uint32_t *first, *members;
  double cell, pad_cells, pad;
  double minimum[3], maximum[3];
  uint32_t cells[3];
  // the rest field
  const float *dist;
  float origin[3], voxel[3];
  uint32_t res[3];
  double model[12];  // [R | t], row-major
};

__host__ __device__ inline void hex_shape(const double *l, double *shape) {
  for (uint32_t c = 0; c < 8; ++c) {
    double product = 1.0;
    for (int a = 0; a < 3; ++a) product *= ((c >> a) & 1u) ? l[a] : 1.0 - l[a];
    shape[c] = product;
  }
}

__host__ __device__ inline void hex_slope(const double *l, double *slope) {
  for (uint32_t c = 0; c < 8; ++c) {
    double f[3], s[3];
    for (int a = 0; a < 3; ++a) {
      const bool high = ((c >> a) & 1u) != 0u;
      f[a] = high ? l[a] : 1.0 - l[a];
      s[a] = high ? 1.0 : -1.0;
    }
    slope[3 * c + 0] = s[0] * f[1] * f[2];
    slope[3 * c + 1] = f[0] * s[1] * f[2];
    slope[3 * c + 2] = f[0] * f[1] * s[2];
  }
}

__host__ __device__ inline double det3d(const double *a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

__host__ __device__ inline void inverse3d(const double *a, double *o) {
  const double d = det3d(a);
  o[0] = (a[4] * a[8] - a[5] * a[7]) / d;
  o[1] = (a[2] * a[7] - a[1] * a[8]) / d;
  o[2] = (a[1] * a[5] - a[2] * a[4]) / d;
  o[3] = (a[5] * a[6] - a[3] * a[8]) / d;
  o[4] = (a[0] * a[8] - a[2] * a[6]) / d;
  o[5] = (a[2] * a[3] - a[0] * a[5]) / d;
  o[6] = (a[3] * a[7] - a[4] * a[6]) / d;
  o[7] = (a[1] * a[6] - a[0] * a[7]) / d;
  o[8] = (a[0] * a[4] - a[1] * a[3]) / d;
}

// probe<double>(field, point).distance
__device__ double sdf_probe(const ProbeDev &p, const double *q) {
  double t[3], gap2 = 0.0;
  uint32_t base[3];
  for (int k = 0; k < 3; ++k) {
    const double lo = double(p.origin[k]);
    const double hi = double(p.origin[k] + p.voxel[k] * float(p.res[k] - 1u));
    const double c = q[k] < lo ? lo : (q[k] > hi ? hi : q[k]);
    gap2 += (q[k] - c) * (q[k] - c);
    const double lattice = (c - double(p.origin[k])) / double(p.voxel[k]);
    const double fl = floor(lattice);
    uint32_t b = fl < 0.0 ? 0u : uint32_t(fl);
    b = b > p.res[k] - 2u ? p.res[k] - 2u : b;
    base[k] = b;
    t[k] = lattice - double(b);
  }
  double c[8];
  for (uint32_t i = 0; i < 8; ++i) {
    const size_t x = base[0] + (i & 1u), y = base[1] + ((i >> 1) & 1u), z = base[2] + ((i >> 2) & 1u);
    c[i] = double(p.dist[x + size_t(p.res[0]) * (y + size_t(p.res[1]) * z)]);
  }
  const double x00 = c[0] * (1.0 - t[0]) + c[1] * t[0];
  const double x10 = c[2] * (1.0 - t[0]) + c[3] * t[0];
  const double x01 = c[4] * (1.0 - t[0]) + c[5] * t[0];
  const double x11 = c[6] * (1.0 - t[0]) + c[7] * t[0];
  const double y0 = x00 * (1.0 - t[1]) + x10 * t[1];
  const double y1 = x01 * (1.0 - t[1]) + x11 * t[1];
  const double inner = y0 * (1.0 - t[2]) + y1 * t[2];
  const double outside = sqrt(gap2);
  return outside <= 0.0 ? inner : inner + outside;
}

struct Inversion {
  double local[3];
  bool converged, inverted;
};

__device__ Inversion invert_hex(const ProbeDev &p, uint32_t e, const double *point) {
  Inversion f{{0.5, 0.5, 0.5}, false, false};
  double determinant = 0.0;
  for (int it = 0; it < DEFORMED_ITERATIONS; ++it) {
    double shape[8], slope[24], mapped[3] = {0.0, 0.0, 0.0}, J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Ji[9];
    hex_shape(f.local, shape);
    hex_slope(f.local, slope);
    for (int c = 0; c < 8; ++c) {
      const double *at = p.pos + 3 * size_t(p.hex[8 * size_t(e) + c]);
      for (int r = 0; r < 3; ++r) {
        mapped[r] += shape[c] * at[r];
        for (int s = 0; s < 3; ++s) J[3 * r + s] += at[r] * slope[3 * c + s];
      }
    }
    determinant = det3d(J);
    if (determinant == 0.0 || !isfinite(determinant)) break;
    inverse3d(J, Ji);
    double most = 0.0;
    for (int r = 0; r < 3; ++r) {
      const double step = Ji[3 * r] * (mapped[0] - point[0]) + Ji[3 * r + 1] * (mapped[1] - point[1]) +
                          Ji[3 * r + 2] * (mapped[2] - point[2]);
      f.local[r] -= step;
      most = fabs(step) > most ? fabs(step) : most;
    }
    if (most < DEFORMED_TOLERANCE) {
      f.converged = true;
      break;
    }
  }
  f.inverted = f.converged && !(determinant > 0.0);
  return f;
}

__device__ double evaluate_hex(const ProbeDev &p, uint32_t e, const double *local) {
  double shape[8], slope[24], rp[3] = {0.0, 0.0, 0.0}, R[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  hex_shape(local, shape);
  hex_slope(local, slope);
  for (int c = 0; c < 8; ++c) {
    const double *rc = p.rest + 3 * size_t(p.hex[8 * size_t(e) + c]);
    for (int r = 0; r < 3; ++r) {
      rp[r] += shape[c] * rc[r];
      for (int s = 0; s < 3; ++s) R[3 * r + s] += rc[r] * slope[3 * c + s];
    }
  }
  double distance = sdf_probe(p, rp);
  const uint8_t cracks = p.cracked[e];
  for (int a = 0; a < 3 && cracks; ++a)
    for (int side = 0; side < 2; ++side) {
      if (!(cracks & (1u << (2 * a + side)))) continue;
      const double extent = sqrt(R[a] * R[a] + R[3 + a] * R[3 + a] + R[6 + a] * R[6 + a]);
      const double depth = side == 0 ? local[a] : 1.0 - local[a];
      const double face = -depth * extent;
      if (face > distance) distance = face;
    }
  return distance;
}

__device__ bool screened(const ProbeDev &p, uint32_t e, const double *point, uint8_t faces) {
  const double *Mi = p.inverse + 9 * size_t(e);
  bool zero = true;
  for (int k = 0; k < 9; ++k) zero = zero && Mi[k] == 0.0;
  if (zero) return true;
  const double *c = p.centre + 3 * size_t(e);
  const double d[3] = {point[0] - c[0], point[1] - c[1], point[2] - c[2]};
  for (int a = 0; a < 3; ++a) {
    const double value = Mi[3 * a] * d[0] + Mi[3 * a + 1] * d[1] + Mi[3 * a + 2] * d[2] + 0.5;
    if (value < -DEFORMED_SCREEN) {
      if (!(faces & (1u << (2 * a))) || value < -DEFORMED_SCREEN - p.pad_cells) return false;
    } else if (value > 1.0 + DEFORMED_SCREEN) {
      if (!(faces & (1u << (2 * a + 1))) || value > 1.0 + DEFORMED_SCREEN + p.pad_cells) return false;
    }
  }
  return true;
}

__device__ bool inside(const double *q, const double *lo, const double *hi, double margin) {
  for (int k = 0; k < 3; ++k)
    if (!(q[k] >= lo[k] - margin && q[k] <= hi[k] + margin)) return false;
  return true;
}

__device__ bool probe_deformed(const ProbeDev &p, const double *point, double *distance) {
  if (p.cells[0] == 0u || !inside(point, p.minimum, p.maximum, 0.0)) return false;
  uint32_t cell[3];
  for (int k = 0; k < 3; ++k) {
    const uint32_t c = uint32_t(floor((point[k] - p.minimum[k]) / p.cell));
    cell[k] = c < p.cells[k] - 1u ? c : p.cells[k] - 1u;
  }
  const size_t index = cell[0] + p.cells[0] * (size_t(cell[1]) + size_t(p.cells[1]) * cell[2]);
  const double margin = DEFORMED_TOLERANCE * p.cell;
  const uint32_t begin = p.first[index], end = p.first[index + 1];
  for (uint32_t at = begin; at != end; ++at) {
    const uint32_t e = p.members[at];
    if (!inside(point, p.low + 3 * size_t(e), p.high + 3 * size_t(e), margin) || !screened(p, e, point, 0u))
      continue;
    const Inversion f = invert_hex(p, e, point);
    if (!f.converged || f.inverted) continue;
    const double *l = f.local;
    if (l[0] >= -DEFORMED_TOLERANCE && l[1] >= -DEFORMED_TOLERANCE && l[2] >= -DEFORMED_TOLERANCE &&
        l[0] <= 1.0 + DEFORMED_TOLERANCE && l[1] <= 1.0 + DEFORMED_TOLERANCE && l[2] <= 1.0 + DEFORMED_TOLERANCE) {
      *distance = evaluate_hex(p, e, l);
      return true;
    }
  }
  bool accepted = false;
  uint32_t chosen = 0;
  double chosen_local[3] = {0.0, 0.0, 0.0}, chosen_excess = INFINITY;
  for (uint32_t at = begin; at != end; ++at) {
    const uint32_t e = p.members[at];
    const uint8_t faces = p.boundary[e];
    if (faces == 0 || !inside(point, p.low + 3 * size_t(e), p.high + 3 * size_t(e), p.pad) ||
        !screened(p, e, point, faces))
      continue;
    const Inversion f = invert_hex(p, e, point);
    if (!f.converged || f.inverted) continue;
    bool across = true;
    double excess = 0.0;
    for (int a = 0; a < 3 && across; ++a) {
      const double value = f.local[a];
      if (value < -DEFORMED_TOLERANCE) {
        across = (faces & (1u << (2 * a))) != 0;
        excess = -value > excess ? -value : excess;
      } else if (value > 1.0 + DEFORMED_TOLERANCE) {
        across = (faces & (1u << (2 * a + 1))) != 0;
        excess = value - 1.0 > excess ? value - 1.0 : excess;
      }
    }
    if (across && excess < chosen_excess) {
      accepted = true;
      chosen = e;
      for (int k = 0; k < 3; ++k) chosen_local[k] = f.local[k];
      chosen_excess = excess;
    }
  }
  if (!accepted) return false;
  *distance = evaluate_hex(p, chosen, chosen_local);
  return true;
}

// one sample per thread; box = {min x, y, z, max x, y, z}
__global__ __launch_bounds__(NT) void k_bake(ProbeDev p, const float *current, float *baked, int carve,
                                             uint32_t *box) {
  const size_t count = size_t(p.res[0]) * p.res[1] * p.res[2];
  const size_t index = size_t(blockIdx.x) * NT + threadIdx.x;
  if (index >= count) return;
  const uint32_t x = uint32_t(index % p.res[0]);
  const uint32_t y = uint32_t((index / p.res[0]) % p.res[1]);
  const uint32_t z = uint32_t(index / (size_t(p.res[0]) * p.res[1]));
  const uint32_t xyz[3] = {x, y, z};
  double local[3], world[3], distance = 0.0;
  for (int k = 0; k < 3; ++k) local[k] = double(p.origin[k]) + double(p.voxel[k]) * double(xyz[k]);
  for (int r = 0; r < 3; ++r)
    world[r] = p.model[4 * r] * local[0] + p.model[4 * r + 1] * local[1] + p.model[4 * r + 2] * local[2] +
               p.model[4 * r + 3];
  const float rest = p.dist[index];
  const float unplaced = carve ? (rest > float(p.pad) ? rest : float(p.pad)) : rest;
  const float value = probe_deformed(p, world, &distance) ? float(distance) : unplaced;
  const float now = current[index];
  if (fabsf(value - now) > UNCHANGED) {
    baked[index] = value;
    for (int k = 0; k < 3; ++k) {
      atomicMin(&box[k], xyz[k]);
      atomicMax(&box[3 + k], xyz[k]);
    }
  } else {
    baked[index] = now;
  }
}

uint64_t cell_key(const double *rest, const double *smallest, const double *voxel) {
  uint64_t key = 0;
  for (int k = 0; k < 3; ++k) {
    const double scaled = std::round((rest[k] - smallest[k]) / voxel[k]);
    const uint64_t packed = uint64_t(int64_t(scaled) + int64_t(1ull << 20)) & 0x1fffffull;
    key |= packed << (21 * k);
  }
  return key;
}

template <class T>
T *upload(const std::vector<T> &h, std::vector<void *> &owned) {
  T *ptr = nullptr;
  HIPCHECK(hipMalloc(&ptr, std::max<size_t>(h.size(), 1) * sizeof(T)));
  owned.push_back(ptr);
  if (!h.empty()) HIPCHECK(hipMemcpy(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return ptr;
}

}  // namespace

bool bake_deformed_field(const LatticeBody &b, const double *x, const double *model, const talus_mi_sdf &rest,
                         bool carve, talus_mi_sdf &current, talus_mi_sample_box &changed, hipStream_t stream) {
  for (int k = 0; k < 3; ++k)
    if (current.resolution[k] != rest.resolution[k] || current.origin[k] != rest.origin[k] ||
        current.voxel_size[k] != rest.voxel_size[k])
      return false;
  const size_t E = b.hexes, N = b.nodes;
  ProbeDev p{};
  // deformedProbe (infinite reach: DEFORMED_PAD cells of pad)
  p.cell = double(std::max({b.voxel[0], b.voxel[1], b.voxel[2]}));
  p.pad_cells = DEFORMED_PAD;
  p.pad = p.pad_cells * p.cell;
  const double voxel[3] = {double(b.voxel[0]), double(b.voxel[1]), double(b.voxel[2])};
  double smallest[3] = {INFINITY, INFINITY, INFINITY};
  for (size_t n = 0; n < N; ++n)
    for (int k = 0; k < 3; ++k) smallest[k] = std::min(smallest[k], b.rest[3 * n + k]);
  std::vector<uint64_t> keys(E);
  for (size_t e = 0; e < E; ++e) keys[e] = cell_key(&b.rest[3 * size_t(b.hex[8 * e])], smallest, voxel);
  std::sort(keys.begin(), keys.end());
  std::vector<uint8_t> boundary(E, 0), cracked(E, 0);
  for (size_t e = 0; e < E; ++e) {
    const double *corner = &b.rest[3 * size_t(b.hex[8 * e])];
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side) {
        double nb[3] = {corner[0], corner[1], corner[2]};
        nb[a] += side == 0 ? -voxel[a] : voxel[a];
        if (!std::binary_search(keys.begin(), keys.end(), cell_key(nb, smallest, voxel)))
          boundary[e] |= uint8_t(1u << (2 * a + side));
      }
    if (e < b.crack.size()) {
      cracked[e] = b.crack[e];
      boundary[e] |= b.crack[e];
    }
  }
  // placeDeformedProbe
  std::vector<double> low(3 * E), high(3 * E), centre(3 * E), inverse(9 * E, 0.0);
  double middle[24];
  const double half[3] = {0.5, 0.5, 0.5};
  hex_slope(half, middle);
  for (size_t e = 0; e < E; ++e) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double c[3] = {0.0, 0.0, 0.0}, J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 8; ++k) {
      const double *at = &x[3 * size_t(b.hex[8 * e + k])];
      for (int r = 0; r < 3; ++r) {
        lo[r] = std::min(lo[r], at[r]);
        hi[r] = std::max(hi[r], at[r]);
        c[r] += 0.125 * at[r];
        for (int s = 0; s < 3; ++s) J[3 * r + s] += at[r] * middle[3 * k + s];
      }
    }
    for (int r = 0; r < 3; ++r) {
      low[3 * e + r] = lo[r];
      high[3 * e + r] = hi[r];
      centre[3 * e + r] = c[r];
    }
    const double d = det3d(J);
    if (d > 0.0 && std::isfinite(d)) inverse3d(J, &inverse[9 * e]);
  }
  for (int k = 0; k < 3; ++k) {
    p.minimum[k] = INFINITY;
    p.maximum[k] = -INFINITY;
  }
  for (size_t e = 0; e < E; ++e) {
    const double reach = boundary[e] ? p.pad : 0.0;
    for (int k = 0; k < 3; ++k) {
      p.minimum[k] = std::min(p.minimum[k], low[3 * e + k] - reach);
      p.maximum[k] = std::max(p.maximum[k], high[3 * e + k] + reach);
    }
  }
  std::vector<uint32_t> first(1, 0), members;
  if (E) {
    for (int k = 0; k < 3; ++k) p.cells[k] = uint32_t(std::floor((p.maximum[k] - p.minimum[k]) / p.cell)) + 1u;
    const size_t total = size_t(p.cells[0]) * p.cells[1] * p.cells[2];
    first.assign(total + 1, 0);
    auto cell_of = [&](const double *q, uint32_t *out) {
      for (int k = 0; k < 3; ++k) {
        const double s = std::max(std::floor((q[k] - p.minimum[k]) / p.cell), 0.0);
        out[k] = std::min(uint32_t(s), p.cells[k] - 1u);
      }
    };
    auto visit = [&](size_t e, auto &&action) {
      const double reach = boundary[e] ? p.pad : 0.0;
      double a[3], c[3];
      uint32_t from[3], to[3];
      for (int k = 0; k < 3; ++k) {
        a[k] = low[3 * e + k] - reach;
        c[k] = high[3 * e + k] + reach;
      }
      cell_of(a, from);
      cell_of(c, to);
      for (uint32_t z = from[2]; z <= to[2]; ++z)
        for (uint32_t y = from[1]; y <= to[1]; ++y)
          for (uint32_t xx = from[0]; xx <= to[0]; ++xx)
            action(xx + p.cells[0] * (size_t(y) + size_t(p.cells[1]) * z));
    };
    for (size_t e = 0; e < E; ++e) visit(e, [&](size_t i) { ++first[i + 1]; });
    for (size_t i = 0; i < total; ++i) first[i + 1] += first[i];
    members.resize(first[total]);
    std::vector<uint32_t> cursor(first.begin(), first.end() - 1);
    for (size_t e = 0; e < E; ++e) visit(e, [&](size_t i) { members[cursor[i]++] = uint32_t(e); });
  }
  // device tables and the bake
  const size_t count = size_t(rest.resolution[0]) * rest.resolution[1] * rest.resolution[2];
  std::vector<void *> owned;
  try {
    p.hex = upload(b.hex, owned);
    p.pos = upload(std::vector<double>(x, x + 3 * N), owned);
    p.rest = upload(b.rest, owned);
    p.boundary = upload(boundary, owned);
    p.cracked = upload(cracked, owned);
    p.low = upload(low, owned);
    p.high = upload(high, owned);
    p.centre = upload(centre, owned);
    p.inverse = upload(inverse, owned);
    p.first = upload(first, owned);
    p.members = upload(members, owned);
    p.dist = upload(std::vector<float>(rest.distances, rest.distances + count), owned);
    for (int k = 0; k < 3; ++k) {
      p.origin[k] = rest.origin[k];
      p.voxel[k] = rest.voxel_size[k];
      p.res[k] = rest.resolution[k];
    }
    std::memcpy(p.model, model, sizeof p.model);
    float *cur = upload(std::vector<float>(current.distances, current.distances + count), owned);
    float *baked = upload(std::vector<float>(count, 0.0f), owned);
    const uint32_t init[6] = {UINT32_MAX, UINT32_MAX, UINT32_MAX, 0u, 0u, 0u};
    uint32_t *box = upload(std::vector<uint32_t>(init, init + 6), owned);
    if (count)
      hipLaunchKernelGGL(k_bake, dim3(unsigned((count + NT - 1) / NT)), dim3(NT), 0, stream, p, cur, baked,
                         carve ? 1 : 0, box);
    HIPCHECK(hipGetLastError());
    uint32_t hbox[6];
    HIPCHECK(hipMemcpyAsync(current.distances, baked, count * sizeof(float), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipMemcpyAsync(hbox, box, sizeof hbox, hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    const bool any = hbox[0] != UINT32_MAX;
    for (int k = 0; k < 3; ++k) {
      changed.minimum[k] = any ? hbox[k] : 0u;
      changed.extent[k] = any ? hbox[3 + k] - hbox[k] + 1u : 0u;
    }
  } catch (...) {
    for (void *ptr : owned) (void)hipFree(ptr);
    throw;
  }
  for (void *ptr : owned) (void)hipFree(ptr);
  return true;
}

}  // namespace tmi
