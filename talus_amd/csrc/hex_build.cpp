// hex_build.cpp -- host side of the strainstep hex lattices: the hex mesh
// builder, the lattice body set-up and the 18-row element build.
//
// Reference (under src/talus/):
//   deformable/solver/solver.cc:2185-2316   buildMesh, VolumeElementType::Hex
//   dynamics/bodyset/bodyset.cc:13-86       makeLattice
//   dynamics/strainstep/element.cc:22-99    springOf, symmetricInverse
//   dynamics/strainstep/element.cc:162-342  signs, elementMaterial,
//                                           cellFrequency, cornerRest,
//                                           rowJacobian, makeElement
//   dynamics/strainstep/strainstep.cc:61-149, 551-567  pinned nodes, substep
//                                           rule, inverse masses
//   dynamics/strainstep/plasticflow.cc:62-72, 342-389  authoredRest, the
//                                           lattice turn
// Elements are set-up work (built once per body and substep length, as the
// reference caches them in StrainLattice); the per-substep update and solves
// run on the GPU (lattice.hip).  Compiled without floating-point
// contraction, so a build equals the oracle's operation for operation.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <thread>
#include <unordered_map>

#include "hexlattice.h"
#include "host.h"

namespace tmi {

namespace {

constexpr double MAX_CORRECTIVE_SPEED = 3.0;
constexpr double RESOLVED = 1.5;
constexpr uint32_t MOST_SUBSTEPS = 64;
constexpr double INF = std::numeric_limits<double>::infinity();

void signs_of(uint32_t c, double *s) {
  s[0] = (c & 1u) ? 1.0 : -1.0;
  s[1] = (c & 2u) ? 1.0 : -1.0;
  s[2] = (c & 4u) ? 1.0 : -1.0;
}

double hg_sign(uint32_t mode, uint32_t corner) {
  double s[3];
  signs_of(corner, s);
  switch (mode) {
    case 0: return s[0] * s[1];
    case 1: return s[1] * s[2];
    case 2: return s[2] * s[0];
    default: return s[0] * s[1] * s[2];
  }
}

double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

double det3(const double *a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

void inverse3(const double *a, double *inv) {
  const double s = 1.0 / det3(a);
  inv[0] = (a[4] * a[8] - a[5] * a[7]) * s;
  inv[1] = (a[2] * a[7] - a[1] * a[8]) * s;
  inv[2] = (a[1] * a[5] - a[2] * a[4]) * s;
  inv[3] = (a[5] * a[6] - a[3] * a[8]) * s;
  inv[4] = (a[0] * a[8] - a[2] * a[6]) * s;
  inv[5] = (a[2] * a[3] - a[0] * a[5]) * s;
  inv[6] = (a[3] * a[7] - a[4] * a[6]) * s;
  inv[7] = (a[1] * a[6] - a[0] * a[7]) * s;
  inv[8] = (a[0] * a[4] - a[1] * a[3]) * s;
}

// Jacobian of one row with respect to one corner's velocity in the element frame
void row_jacobian(const HexElement &e, uint32_t corner, uint32_t row, double *j) {
  const double *c = e.gradient + 3 * corner;
  j[0] = j[1] = j[2] = 0.0;
  switch (row) {
    case 0: j[0] = c[0]; return;
    case 1: j[1] = c[1]; return;
    case 2: j[2] = c[2]; return;
    case 3: j[0] = 0.5 * c[1]; j[1] = 0.5 * c[0]; return;
    case 4: j[1] = 0.5 * c[2]; j[2] = 0.5 * c[1]; return;
    case 5: j[0] = 0.5 * c[2]; j[2] = 0.5 * c[0]; return;
    default: j[(row - 6) % 3] = e.bend * hg_sign((row - 6) / 3, corner); return;
  }
}

void spring_of(double omega, double zeta, double h, double theta, double *rate, double *softness) {
  const double b = h * omega;
  const double a = theta * b * b + 2.0 * zeta * b;
  if (a > 0.0) {
    *rate = h * omega * omega / a;
    *softness = 1.0 / a;
  } else {
    *rate = 0.0;
    *softness = 0.0;
  }
}

// Gauss-Jordan inverse of the SPD row block without pivoting; inert rows are
// decoupled before and zeroed after; a non-positive pivot gives zero.
void symmetric_inverse(const double *block, const bool *inert, double *out) {
  constexpr uint32_t N = HEX_ROWS;
  double m[N * N];
  std::memcpy(m, block, sizeof m);
  std::memset(out, 0, N * N * sizeof(double));
  for (uint32_t r = 0; r < N; ++r) {
    if (inert[r]) {
      for (uint32_t k = 0; k < N; ++k) m[N * r + k] = m[N * k + r] = 0.0;
      m[(N + 1) * r] = 1.0;
    }
    out[(N + 1) * r] = 1.0;
  }
  for (uint32_t col = 0; col < N; ++col) {
    const double pivot = m[(N + 1) * col];
    if (!(pivot > 0.0)) {
      std::memset(out, 0, N * N * sizeof(double));
      return;
    }
    for (uint32_t k = 0; k < N; ++k) {
      m[N * col + k] /= pivot;
      out[N * col + k] /= pivot;
    }
    for (uint32_t r = 0; r < N; ++r) {
      const double f = m[N * r + col];
      if (r == col || f == 0.0) continue;
      for (uint32_t k = 0; k < N; ++k) {
        m[N * r + k] -= f * m[N * col + k];
        out[N * r + k] -= f * out[N * col + k];
      }
    }
  }
  for (uint32_t r = 0; r < N; ++r)
    if (inert[r])
      for (uint32_t k = 0; k < N; ++k) out[N * r + k] = out[N * k + r] = 0.0;
}

void quat_normalize(double *q) {
  const double len = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (len <= 0.0) {
    q[0] = 1.0;
    q[1] = q[2] = q[3] = 0.0;
    return;
  }
  const double s = 1.0 / len;
  for (int k = 0; k < 4; ++k) q[k] *= s;
}

void quat_matrix(const double *q, double *R) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  const double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  const double wx = w * x, wy = w * y, wz = w * z;
  R[0] = 1.0 - 2.0 * (yy + zz);
  R[1] = 2.0 * (xy - wz);
  R[2] = 2.0 * (xz + wy);
  R[3] = 2.0 * (xy + wz);
  R[4] = 1.0 - 2.0 * (xx + zz);
  R[5] = 2.0 * (yz - wx);
  R[6] = 2.0 * (xz - wy);
  R[7] = 2.0 * (yz + wx);
  R[8] = 1.0 - 2.0 * (xx + yy);
}

double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

constexpr double GAUSS_OFFSET = 0.5773502691896258;  // 1 / sqrt(3)

void mm3(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

void mv3(const double *M, const double *x, double *y) {
  for (int r = 0; r < 3; ++r) y[r] = M[3 * r] * x[0] + M[3 * r + 1] * x[1] + M[3 * r + 2] * x[2];
}

// fittedJacobian (plasticflow.cc:83-92): the rest Jacobian of a fitted cell
// at Gauss point `point`
void fitted_jacobian(const double *shape, const double *hg, uint32_t point, double *J) {
  double s[3];
  signs_of(point, s);
  const double at[3] = {GAUSS_OFFSET * s[0], GAUSS_OFFSET * s[1], GAUSS_OFFSET * s[2]};
  const double slopes[12] = {at[1], at[0], 0.0, 0.0, at[2], at[1], at[2], 0.0, at[0],
                             at[1] * at[2], at[0] * at[2], at[0] * at[1]};
  std::memcpy(J, shape, 9 * sizeof(double));
  for (uint32_t m = 0; m < 4; ++m)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) J[3 * i + j] += hg[3 * m + i] * slopes[3 * m + j] / 8.0;
}

double remaining(const double *before, const double *after, int n) {
  double was = 0.0, now = 0.0;
  for (int k = 0; k < n; ++k) {
    was += before[k] * before[k];
    now += after[k] * after[k];
  }
  return was > 0.0 ? std::min(1.0, std::sqrt(now / was)) : 1.0;
}

}  // namespace

bool hex_fit_holds(const double *shape, const double *hg) {
  for (uint32_t p = 0; p < 8; ++p) {
    double J[9];
    fitted_jacobian(shape, hg, p, J);
    if (!(det3(J) > 0.0)) return false;
  }
  return true;
}

void hex_strain(HexElement &e, const double *x, double *strain) {
  double F[9] = {}, hgv[12] = {};
  for (uint32_t c = 0; c < 8; ++c) {
    const double *xc = x + 3 * size_t(e.nodes[c]);
    const double *g = e.gradient + 3 * c;
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) F[3 * r + k] += xc[r] * g[k];
    for (uint32_t m = 0; m < 4; ++m) {
      const double sg = hg_sign(m, c);
      for (int r = 0; r < 3; ++r) hgv[3 * m + r] += sg * xc[r];
    }
  }
  hex_rotation_of(F, e.quat, 16, e.quat);
  double T[9];
  quat_matrix(e.quat, T);
  double L[9];
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) L[3 * r + k] = T[r] * F[k] + T[3 + r] * F[3 + k] + T[6 + r] * F[6 + k];
  const bool inverted = det3(F) < 0.0;
  double raw[HEX_ROWS];
  raw[0] = L[0] - 1.0;
  raw[1] = L[4] - 1.0;
  raw[2] = L[8] - 1.0;
  raw[3] = 0.5 * (L[1] + L[3]);
  raw[4] = 0.5 * (L[5] + L[7]);
  raw[5] = 0.5 * (L[6] + L[2]);
  for (uint32_t m = 0; m < 4; ++m)
    for (int k = 0; k < 3; ++k) {
      const double b = T[k] * hgv[3 * m] + T[3 + k] * hgv[3 * m + 1] + T[6 + k] * hgv[3 * m + 2];
      raw[6 + 3 * m + k] = (b - e.rest_hg[3 * m + k]) * e.bend;
    }
  for (uint32_t r = 0; r < HEX_ROWS; ++r) {
    const double cap = inverted ? e.cap[r] : std::max(e.cap[r], 1.0);
    strain[r] = clampd(raw[r], -cap, cap);
  }
}

void hex_corner_rest(const double *corners, double *shape, double *hourglass) {
  std::memset(shape, 0, 9 * sizeof(double));
  std::memset(hourglass, 0, 12 * sizeof(double));
  for (uint32_t c = 0; c < 8; ++c) {
    double s[3];
    signs_of(c, s);
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) shape[3 * r + k] += corners[3 * c + r] * s[k] / 8.0;
    for (uint32_t m = 0; m < 4; ++m) {
      const double g = hg_sign(m, c);
      for (int r = 0; r < 3; ++r) hourglass[3 * m + r] += g * corners[3 * c + r];
    }
  }
}

void hex_material(float youngs, float poisson, double *shear, double *first) {
  const double nu = clampd(double(poisson), 0.0, 0.499);
  const double e = std::max(0.0, double(youngs));
  *shear = e / (2.0 * (1.0 + nu));
  *first = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
}

double hex_cell_frequency(float youngs, float poisson, float density, double cell) {
  const double nu = clampd(double(poisson), 0.0, 0.499);
  const double e = std::max(0.0, double(youngs));
  return std::sqrt(e / ((1.0 + nu) * double(density))) / cell;
}

// rowDamping (element.cc:29-40)
double row_damping(const HexDamping &d, double omega) {
  if (!std::isfinite(d.target) || !std::isfinite(d.lowest) || !(d.lowest > 0.0)) return 1.0;
  return clampd(d.target * omega / d.lowest - d.own * omega / 2.0, 0.0, 1.0);
}

bool hex_make(HexElement &e, uint32_t cell, const uint32_t *nodes, const double *inv_mass,
              const double *shape, const double *hourglass, double volume, double shear,
              double first, double h, const HexDamping *damping) {
  const HexDamping undamped{};
  const HexDamping &dmp = damping ? *damping : undamped;
  const double det = det3(shape);
  if (!(std::fabs(det) > 0.0) || !(shear > 0.0) || !(volume > 0.0)) return false;
  e = HexElement{};
  inverse3(shape, e.shape_inverse);
  e.cell = cell;
  e.size = std::cbrt(8.0 * std::fabs(det));
  e.bend = 1.0 / (4.0 * e.size);
  std::memcpy(e.rest_hg, hourglass, sizeof e.rest_hg);
  const double *inv = e.shape_inverse;
  for (uint32_t c = 0; c < 8; ++c) {
    double s[3];
    signs_of(c, s);
    e.nodes[c] = nodes[c];
    e.inv_mass[c] = inv_mass[c];
    for (int r = 0; r < 3; ++r)
      e.gradient[3 * c + r] = (inv[r] * s[0] + inv[3 + r] * s[1] + inv[6 + r] * s[2]) / 8.0;
  }
  double block[HEX_INVERSE] = {};
  double vweight = 0.0;
  for (uint32_t c = 0; c < 8; ++c) {
    const double w = e.inv_mass[c];
    if (w == 0.0) continue;
    vweight += w * dot3(e.gradient + 3 * c, e.gradient + 3 * c);
    for (uint32_t r = 0; r < HEX_ROWS; ++r) {
      double jl[3];
      row_jacobian(e, c, r, jl);
      for (uint32_t k = 0; k < HEX_ROWS; ++k) {
        double jr[3];
        row_jacobian(e, c, k, jr);
        block[HEX_ROWS * r + k] += w * dot3(jl, jr);
      }
    }
  }
  if (!(vweight > 0.0)) return false;
  const double hg = shear * (3.0 * first + 2.0 * shear) / (3.0 * (first + shear));
  bool inert[HEX_ROWS] = {};
  for (uint32_t r = 0; r < HEX_ROWS; ++r) {
    const double modulus = r < 3 ? 2.0 * shear : (r < 6 ? 4.0 * shear : hg);
    const double weight = block[(HEX_ROWS + 1) * r];
    const double omega = std::sqrt(modulus * volume * weight);
    e.stiffness[r] = modulus * volume;
    if (omega > 0.0) {
      double rate, soft;
      spring_of(omega, row_damping(dmp, omega), h, dmp.theta, &rate, &soft);
      e.rate[r] = rate;
      e.softness[r] = weight * soft;
      block[(HEX_ROWS + 1) * r] += e.softness[r];
      e.cap[r] = MAX_CORRECTIVE_SPEED / (rate * e.size);
    } else {
      inert[r] = true;
      e.cap[r] = INF;
    }
  }
  symmetric_inverse(block, inert, e.inverse.data());
  e.vol_stiffness = first * volume;
  e.vol_cap = INF;
  const double vomega = std::sqrt(e.vol_stiffness * vweight);
  if (vomega > 0.0) {
    double rate, soft;
    spring_of(vomega, row_damping(dmp, vomega), h, dmp.theta, &rate, &soft);
    e.vol_rate = rate;
    e.vol_softness = soft;
    e.vol_cap = rate > 0.0 ? MAX_CORRECTIVE_SPEED / (rate * e.size) : INF;
    e.volumetric = true;
  }
  return true;
}

void hex_rotation_of(const double *F, const double *guess, uint32_t iterations, double *out) {
  double q[4] = {guess[0], guess[1], guess[2], guess[3]};
  quat_normalize(q);
  for (uint32_t it = 0; it < iterations; ++it) {
    double T[9];
    quat_matrix(q, T);
    double torque[3] = {0.0, 0.0, 0.0}, tr = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double tc[3] = {T[k], T[3 + k], T[6 + k]};
      const double fc[3] = {F[k], F[3 + k], F[6 + k]};
      const double cr[3] = {tc[1] * fc[2] - fc[1] * tc[2], tc[2] * fc[0] - fc[2] * tc[0],
                            tc[0] * fc[1] - fc[0] * tc[1]};
      if (k == 0) {
        torque[0] = cr[0];
        torque[1] = cr[1];
        torque[2] = cr[2];
        tr = dot3(tc, fc);
      } else {
        torque[0] += cr[0];
        torque[1] += cr[1];
        torque[2] += cr[2];
        tr += dot3(tc, fc);
      }
    }
    const double scale = std::fabs(tr) + 1.0e-9;
    const double om[3] = {torque[0] / scale, torque[1] / scale, torque[2] / scale};
    const double angle = std::sqrt(dot3(om, om));
    if (!(angle > 1.0e-12)) break;
    const double half = angle * 0.5, s = std::sin(half);
    const double p[4] = {std::cos(half), om[0] / angle * s, om[1] / angle * s, om[2] / angle * s};
    const double n[4] = {p[0] * q[0] - p[1] * q[1] - p[2] * q[2] - p[3] * q[3],
                         p[0] * q[1] + p[1] * q[0] + p[2] * q[3] - p[3] * q[2],
                         p[0] * q[2] + p[2] * q[0] + p[3] * q[1] - p[1] * q[3],
                         p[0] * q[3] + p[3] * q[0] + p[1] * q[2] - p[2] * q[1]};
    std::memcpy(q, n, sizeof q);
    quat_normalize(q);
    if (angle <= 1.0e-9) break;
  }
  std::memcpy(out, q, sizeof q);
}

// ---- hex lattice mesh (buildMesh, hex elements) ---------------------------

int build_hex_lattice(const talus_mi_lattice *L, const talus_mi_material *materials, uint32_t lanes,
                      const talus_mi_support *supports, uint32_t nsupports, talus_mi_hex_mesh *out) {
  std::memset(out, 0, sizeof *out);
  for (int k = 0; k < 3; ++k)
    if (L->resolution[k] < 2) return fail(TALUS_MI_ERR_ARGUMENT, "lattice resolution below 2");
  const uint32_t cx = L->resolution[0] - 1, cy = L->resolution[1] - 1, cz = L->resolution[2] - 1;
  const size_t rx = L->resolution[0], ry = L->resolution[1];
  const size_t samples = rx * ry * L->resolution[2];
  auto sample_index = [&](uint32_t x, uint32_t y, uint32_t z) { return x + rx * (y + ry * size_t(z)); };
  std::vector<uint32_t> node_of(samples, UINT32_MAX), sample_of;
  size_t hexes = 0;
  for (uint32_t z = 0; z < cz; ++z)
    for (uint32_t y = 0; y < cy; ++y)
      for (uint32_t x = 0; x < cx; ++x) {
        if (L->solid && !L->solid[x + size_t(cx) * (y + size_t(cy) * z)]) continue;
        ++hexes;
        for (uint32_t c = 0; c < 8; ++c) {
          const size_t s = sample_index(x + (c & 1u), y + ((c >> 1) & 1u), z + ((c >> 2) & 1u));
          if (node_of[s] == UINT32_MAX) {
            node_of[s] = uint32_t(sample_of.size());
            sample_of.push_back(uint32_t(s));
          }
        }
      }
  const uint32_t nodes = uint32_t(sample_of.size());
  if (nodes == 0) return fail(TALUS_MI_ERR_ARGUMENT, "lattice has no solid cell");
  std::vector<float> restf(3 * size_t(nodes));
  for (uint32_t n = 0; n < nodes; ++n) {
    const size_t s = sample_of[n];
    for (int k = 0; k < 3; ++k) {
      if (L->sample_positions) {
        restf[3 * size_t(n) + k] = L->sample_positions[3 * s + k];
      } else {
        const size_t idx = k == 0 ? s % rx : (k == 1 ? (s / rx) % ry : s / (rx * ry));
        restf[3 * size_t(n) + k] = L->origin[k] + L->voxel_size[k] * float(idx);
      }
    }
  }
  std::vector<uint8_t> anchored(nodes, 0);
  for (uint32_t n = 0; n < nodes; ++n)
    for (uint32_t v = 0; v < nsupports; ++v) {
      const float *M = supports[v].world_to_support;
      bool inside = true;
      for (int r = 0; r < 3 && inside; ++r) {
        const float l = M[r * 4 + 0] * restf[3 * size_t(n)] + M[r * 4 + 1] * restf[3 * size_t(n) + 1] +
                        M[r * 4 + 2] * restf[3 * size_t(n) + 2] + M[r * 4 + 3];
        if (!(std::fabs(l) <= supports[v].half_extents[r])) inside = false;
      }
      if (inside) {
        anchored[n] = 1;
        break;
      }
    }
  out->rest_positions = (float *)std::malloc(restf.size() * sizeof(float));
  out->node_mass = (float *)std::calloc(nodes, sizeof(float));
  out->anchored = (uint8_t *)std::malloc(nodes);
  out->hex_nodes = (uint32_t *)std::malloc(8 * hexes * sizeof(uint32_t));
  out->volume = (float *)std::malloc(hexes * sizeof(float));
  out->hex_material = (uint32_t *)std::malloc(hexes * sizeof(uint32_t));
  if (!out->rest_positions || !out->node_mass || !out->anchored || !out->hex_nodes ||
      !out->volume || !out->hex_material) {
    hex_mesh_free(out);
    return fail(TALUS_MI_ERR_MEMORY, "out of host memory building a hex mesh");
  }
  out->nodes = nodes;
  out->hexes = uint32_t(hexes);
  for (int k = 0; k < 3; ++k) out->voxel_size[k] = L->voxel_size[k];
  std::memcpy(out->rest_positions, restf.data(), restf.size() * sizeof(float));
  std::memcpy(out->anchored, anchored.data(), nodes);
  size_t h = 0;
  for (uint32_t z = 0; z < cz; ++z)
    for (uint32_t y = 0; y < cy; ++y)
      for (uint32_t x = 0; x < cx; ++x) {
        const size_t cell = x + size_t(cx) * (y + size_t(cy) * z);
        if (L->solid && !L->solid[cell]) continue;
        const uint32_t lane = L->cell_lane ? L->cell_lane[cell] : 0u;
        const float density = lane_material(materials, lanes, lane).density;
        const float vol = 1.0f * L->voxel_size[0] * L->voxel_size[1] * L->voxel_size[2];
        const float share = density * vol * 0.125f;
        for (uint32_t c = 0; c < 8; ++c) {
          const uint32_t n = node_of[sample_index(x + (c & 1u), y + ((c >> 1) & 1u), z + ((c >> 2) & 1u))];
          out->hex_nodes[8 * h + c] = n;
          out->node_mass[n] += share;
        }
        out->volume[h] = vol;
        out->hex_material[h] = lane;
        ++h;
      }
  return TALUS_MI_OK;
}

void hex_mesh_free(talus_mi_hex_mesh *m) {
  if (!m) return;
  std::free(m->rest_positions);
  std::free(m->node_mass);
  std::free(m->anchored);
  std::free(m->hex_nodes);
  std::free(m->volume);
  std::free(m->hex_material);
  std::memset(m, 0, sizeof *m);
}

// ---- lattice bodies --------------------------------------------------------

const talus_mi_material &LatticeBody::material_of(uint32_t cell_index) const {
  const uint32_t l = lane[cell_index];
  return materials[std::min<size_t>(l, materials.size() - 1)];
}

bool LatticeBody::pinned(size_t node) const { return !(mass[node] > 0.0f) || anchored[node] != 0; }

double LatticeBody::stiffest() const {
  double f = 0.0;
  for (uint32_t c = 0; c < hexes; ++c) {
    const talus_mi_material &m = material_of(c);
    f = std::max(f, hex_cell_frequency(m.youngs_modulus, m.poissons_ratio, m.density, cell));
  }
  return f;
}

uint32_t LatticeBody::island_substeps(double dt, uint32_t softstep) const {
  const double f = stiffest();
  uint32_t needed = 1;
  if (f > 0.0 && std::isfinite(f)) {
    needed = uint32_t(clampd(std::ceil(dt * f / RESOLVED), 1.0, double(MOST_SUBSTEPS)));
    while (needed < MOST_SUBSTEPS && dt / double(needed) * f > RESOLVED) ++needed;
  }
  return std::max({std::max(softstep, 1u), substeps, needed});
}

void LatticeBody::build_elements(double step) {
  std::vector<double> inv(nodes, 0.0);
  for (size_t i = 0; i < nodes; ++i)
    if (!pinned(i) && mass[i] > 0.0f) inv[i] = 1.0 / double(mass[i]);
  // lattice turn: the proper rotation of the mass-weighted moment of the
  // positions about their centre against the rest positions
  double q[4] = {1.0, 0.0, 0.0, 0.0};
  {
    double c[3] = {0, 0, 0}, rc[3] = {0, 0, 0}, total = 0.0;
    for (size_t i = 0; i < nodes; ++i) {
      const double m = mass[i];
      for (int k = 0; k < 3; ++k) {
        c[k] += m * x[3 * i + k];
        rc[k] += m * rest[3 * i + k];
      }
      total += m;
    }
    if (total > 0.0) {
      for (int k = 0; k < 3; ++k) {
        c[k] /= total;
        rc[k] /= total;
      }
      double M[9] = {};
      for (size_t i = 0; i < nodes; ++i) {
        const double m = mass[i];
        for (int r = 0; r < 3; ++r)
          for (int k = 0; k < 3; ++k) M[3 * r + k] += m * (x[3 * i + r] - c[r]) * (rest[3 * i + k] - rc[k]);
      }
      const double id[4] = {1.0, 0.0, 0.0, 0.0};
      hex_rotation_of(M, id, 64, q);
    }
  }
  // a damped lattice keeps the lowest frequency it was first built with
  // (strainstep.cc:300-331)
  if (damped && !(lowest > 0.0 && std::isfinite(lowest))) lowest = lowest_angular(step);
  std::vector<HexElement> built(hexes);
  std::vector<uint8_t> ok(hexes, 0);
  const bool kept = rest_vol.size() == hexes;
  std::vector<double> shapes(9 * size_t(hexes)), hgs(12 * size_t(hexes)), vols(hexes);
  auto work = [&](uint32_t first, uint32_t last) {
    for (uint32_t cell_index = first; cell_index < last; ++cell_index) {
      const uint32_t *nd = &hex[8 * size_t(cell_index)];
      double w[8], shear, firstl;
      double *shape = &shapes[9 * size_t(cell_index)], *hgv = &hgs[12 * size_t(cell_index)];
      double &vol = vols[cell_index];
      for (uint32_t c = 0; c < 8; ++c) w[c] = inv[nd[c]];
      // restOf (strainstep.cc:128-136): the plastic rest, or the rest kept
      // from the last build when the plastic rest's fit does not hold
      plastic_rest(cell_index, shape, hgv, &vol);
      if (!hex_fit_holds(shape, hgv) && kept && rest_vol[cell_index] > 0.0) {
        std::memcpy(shape, &rest_shape[9 * size_t(cell_index)], 9 * sizeof(double));
        std::memcpy(hgv, &rest_hg[12 * size_t(cell_index)], 12 * sizeof(double));
        vol = rest_vol[cell_index];
      }
      const talus_mi_material &m = material_of(cell_index);
      hex_material(m.youngs_modulus, m.poissons_ratio, &shear, &firstl);
      HexElement &e = built[cell_index];
      const HexDamping damping = cell_damping(cell_index, step);
      if (hex_make(e, cell_index, nd, w, shape, hgv, vol, shear, firstl, step, &damping)) {
        std::memcpy(e.quat, q, sizeof q);
        ok[cell_index] = 1;
      }
    }
  };
  const uint32_t threads = hexes >= 4096 ? std::min(16u, std::max(1u, std::thread::hardware_concurrency())) : 1u;
  if (threads <= 1) {
    work(0, hexes);
  } else {
    std::vector<std::thread> pool;
    const uint32_t chunk = (hexes + threads - 1) / threads;
    for (uint32_t t = 0; t < threads; ++t) {
      const uint32_t a = std::min(hexes, t * chunk), b = std::min(hexes, a + chunk);
      if (a < b) pool.emplace_back(work, a, b);
    }
    for (auto &t : pool) t.join();
  }
  elements.clear();
  for (uint32_t c = 0; c < hexes; ++c)
    if (ok[c]) elements.push_back(built[c]);
  h = step;
  // kept rests: those whose fit holds
  rest_shape.swap(shapes);
  rest_hg.swap(hgs);
  rest_vol.swap(vols);
  for (uint32_t c = 0; c < hexes; ++c)
    if (!hex_fit_holds(&rest_shape[9 * size_t(c)], &rest_hg[12 * size_t(c)])) rest_vol[c] = 0.0;
  if (!plastic.empty()) cached = plastic;
}

void LatticeBody::ensure_plastic() {
  if (!plastic.empty()) return;
  plastic.assign(72 * size_t(hexes), 0.0);
  for (size_t i = 0; i < 8 * size_t(hexes); ++i) plastic[9 * i] = plastic[9 * i + 4] = plastic[9 * i + 8] = 1.0;
  hardening.assign(8 * size_t(hexes), 0.0);
  cached = plastic;
}

void LatticeBody::plastic_rest(uint32_t cell_index, double *shape, double *hg, double *vol) const {
  double corners[24], ashape[9], ahg[12];
  const uint32_t *nd = &hex[8 * size_t(cell_index)];
  for (uint32_t c = 0; c < 8; ++c)
    for (int k = 0; k < 3; ++k) corners[3 * c + k] = rest[3 * size_t(nd[c]) + k];
  hex_corner_rest(corners, ashape, ahg);
  const double avol = double(volume[cell_index]);
  if (plastic.empty()) {
    std::memcpy(shape, ashape, sizeof ashape);
    std::memcpy(hg, ahg, sizeof ahg);
    *vol = avol;
    return;
  }
  double stretch[8][9], mean[9] = {}, v = 0.0;
  for (uint32_t p = 0; p < 8; ++p) {
    inverse3(&plastic[9 * (8 * size_t(cell_index) + p)], stretch[p]);
    for (int k = 0; k < 9; ++k) mean[k] += stretch[p][k] / 8.0;
    v += det3(stretch[p]) / 8.0;
  }
  const double factor = std::sqrt(3.0) / 2.0;
  mm3(mean, ashape, shape);
  *vol = avol * v;
  std::memset(hg, 0, 12 * sizeof(double));
  for (uint32_t p = 0; p < 8; ++p) {
    double target[9], s[3], t[3];
    mm3(stretch[p], ashape, target);
    signs_of(p, s);
    const double d[4][3] = {{s[1], s[0], 0.0}, {0.0, s[2], s[1]}, {s[2], 0.0, s[0]},
                            {s[1] * s[2], s[0] * s[2], s[0] * s[1]}};
    for (uint32_t m = 0; m < 4; ++m) {
      mv3(target, d[m], t);
      for (int k = 0; k < 3; ++k) hg[3 * m + k] += m < 3 ? factor * t[k] : t[k];
    }
  }
  double fitted = 0.0;
  for (uint32_t p = 0; p < 8; ++p) {
    double J[9];
    fitted_jacobian(shape, hg, p, J);
    fitted += det3(J) / 8.0;
  }
  const double ratio = v * det3(ashape) / fitted;
  const double size = std::isfinite(ratio) && ratio > 0.0 ? std::cbrt(ratio) : 1.0;
  for (int k = 0; k < 9; ++k) shape[k] *= size;
  for (int k = 0; k < 12; ++k) hg[k] *= size;
}

void LatticeBody::write_fitted(uint32_t cell_index, const double *shape, const double *hg) {
  if (plastic.empty()) return;
  double corners[24], ashape[9], ahg[12];
  const uint32_t *nd = &hex[8 * size_t(cell_index)];
  for (uint32_t c = 0; c < 8; ++c)
    for (int k = 0; k < 3; ++k) corners[3 * c + k] = rest[3 * size_t(nd[c]) + k];
  hex_corner_rest(corners, ashape, ahg);
  for (uint32_t p = 0; p < 8; ++p) {
    double J[9], Ji[9];
    fitted_jacobian(shape, hg, p, J);
    inverse3(J, Ji);
    mm3(ashape, Ji, &plastic[9 * (8 * size_t(cell_index) + p)]);
  }
}

uint32_t LatticeBody::refresh_flowed() {
  if (plastic.empty() || h == 0.0) return 0;
  std::vector<double> inv(nodes, 0.0);
  for (size_t i = 0; i < nodes; ++i)
    if (!pinned(i) && mass[i] > 0.0f) inv[i] = 1.0 / double(mass[i]);
  uint32_t rebuilt = 0;
  for (HexElement &old : elements) {
    const uint32_t cell_index = old.cell;
    const size_t at = 72 * size_t(cell_index);
    if (!std::memcmp(&plastic[at], &cached[at], 72 * sizeof(double))) continue;
    double shape[9], hgv[12], vol;
    plastic_rest(cell_index, shape, hgv, &vol);
    if (!hex_fit_holds(shape, hgv)) continue;
    const uint32_t *nd = &hex[8 * size_t(cell_index)];
    double w[8], shear, firstl;
    for (uint32_t c = 0; c < 8; ++c) w[c] = inv[nd[c]];
    const talus_mi_material &m = material_of(cell_index);
    hex_material(m.youngs_modulus, m.poissons_ratio, &shear, &firstl);
    HexElement fresh;
    const HexDamping damping = cell_damping(cell_index, h);
    if (!hex_make(fresh, cell_index, nd, w, shape, hgv, vol, shear, firstl, h, &damping)) continue;
    std::memcpy(fresh.quat, old.quat, sizeof fresh.quat);
    std::memcpy(fresh.impulse, old.impulse, sizeof fresh.impulse);
    fresh.vol_impulse = old.vol_impulse;
    double before[HEX_ROWS], after[HEX_ROWS];
    hex_strain(old, x.data(), before);
    hex_strain(fresh, x.data(), after);
    const double strain = remaining(before, after, 6);
    const double bend = remaining(before + 6, after + 6, HEX_ROWS - 6);
    for (uint32_t r = 0; r < HEX_ROWS; ++r) fresh.impulse[r] *= r < 6 ? strain : bend;
    write_fitted(cell_index, shape, hgv);
    std::memcpy(&cached[at], &plastic[at], 72 * sizeof(double));
    std::memcpy(&rest_shape[9 * size_t(cell_index)], shape, sizeof shape);
    std::memcpy(&rest_hg[12 * size_t(cell_index)], hgv, sizeof hgv);
    rest_vol[cell_index] = vol;
    old = fresh;
    ++rebuilt;
  }
  return rebuilt;
}


// ---- lowest mode (strainstep/lowestmode.cc) --------------------------------

namespace {

double mass_dot(const std::vector<double> &a, const std::vector<double> &b, const std::vector<double> &m) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * m[i] * b[i];
  return s;
}

void deflate(std::vector<double> &x, const std::vector<std::vector<double>> &rigid, const std::vector<double> &m) {
  for (const auto &mode : rigid) {
    const double c = mass_dot(mode, x, m);
    for (size_t i = 0; i < x.size(); ++i) x[i] -= c * mode[i];
  }
}

}  // namespace

// the order the free coordinates are numbered in: nodes by rest position
// along the longest extent, then the other axes, then id, so the band of
// the lowest-mode factor spans one cross-section of the body
std::vector<uint32_t> band_order(const std::vector<double> &rest, size_t nodes) {
  double lo[3] = {INF, INF, INF}, hi[3] = {-INF, -INF, -INF};
  for (size_t i = 0; i < nodes; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = rest[3 * i + k] < lo[k] ? rest[3 * i + k] : lo[k];
      hi[k] = rest[3 * i + k] > hi[k] ? rest[3 * i + k] : hi[k];
    }
  int axes[3] = {0, 1, 2};
  for (int k = 1; k < 3; ++k)
    if (hi[k] - lo[k] > hi[axes[0]] - lo[axes[0]]) axes[0] = k;
  for (int k = 0, at = 1; k < 3; ++k)
    if (k != axes[0]) axes[at++] = k;
  std::vector<uint32_t> order(nodes);
  for (size_t i = 0; i < nodes; ++i) order[i] = uint32_t(i);
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    for (int k = 0; k < 3; ++k) {
      const double x = rest[3 * size_t(a) + axes[k]], y = rest[3 * size_t(b) + axes[k]];
      if (x < y) return true;
      if (y < x) return false;
    }
    return a < b;
  });
  return order;
}

double hex_lowest_angular(const std::vector<HexElement> &el, const std::vector<double> &mass,
                          const std::vector<uint8_t> &pinned_node, const std::vector<double> &rest) {
  const size_t nodes = mass.size();
  std::vector<int64_t> index(3 * nodes, -1);
  std::vector<double> w;
  std::vector<size_t> owner;
  bool held = false;
  for (const uint32_t node : band_order(rest, nodes)) {
    if (pinned_node[node]) {
      held = true;
      continue;
    }
    if (!(mass[node] > 0.0)) continue;
    for (int a = 0; a < 3; ++a) {
      index[3 * node + a] = int64_t(w.size());
      w.push_back(mass[node]);
      owner.push_back(3 * node + size_t(a));
    }
  }
  const size_t n = w.size();
  if (n == 0) return INF;
  size_t bw = 0;
  for (const HexElement &e : el)
    for (int i = 0; i < 24; ++i) {
      const int64_t p = index[3 * size_t(e.nodes[i / 3]) + i % 3];
      if (p < 0) continue;
      for (int j = 0; j < 24; ++j) {
        const int64_t q = index[3 * size_t(e.nodes[j / 3]) + j % 3];
        if (q >= 0) bw = std::max(bw, size_t(p > q ? p - q : q - p));
      }
    }
  const size_t W = bw + 1;
  std::vector<double> K(n * W, 0.0);
  for (const HexElement &e : el)
    for (uint32_t row = 0; row < HEX_ROWS + 1; ++row) {
      const double st = row < HEX_ROWS ? e.stiffness[row] : e.vol_stiffness;
      if (!(st > 0.0)) continue;
      double line[24];
      for (uint32_t c = 0; c < 8; ++c) {
        if (row < HEX_ROWS)
          row_jacobian(e, c, row, line + 3 * c);
        else
          std::memcpy(line + 3 * c, e.gradient + 3 * c, 3 * sizeof(double));
      }
      for (int i = 0; i < 24; ++i) {
        const int64_t p = index[3 * size_t(e.nodes[i / 3]) + i % 3];
        if (p < 0 || line[i] == 0.0) continue;
        for (int j = 0; j < 24; ++j) {
          const int64_t q = index[3 * size_t(e.nodes[j / 3]) + j % 3];
          if (q >= 0 && q <= p && line[j] != 0.0) K[size_t(p) * W + size_t(p - q)] += st * line[i] * line[j];
        }
      }
    }
  std::vector<std::vector<double>> rigid;
  if (!held) {
    double centre[3] = {0.0, 0.0, 0.0}, total = 0.0;
    for (size_t i = 0; i < n; i += 3) {
      const size_t node = owner[i] / 3;
      for (int a = 0; a < 3; ++a) centre[a] += w[i] * rest[3 * node + a];
      total += w[i];
    }
    for (int a = 0; a < 3; ++a) centre[a] /= total;
    for (int mode = 0; mode < 6; ++mode) {
      std::vector<double> shape(n);
      for (size_t i = 0; i < n; ++i) {
        const size_t node = owner[i] / 3;
        const int axis = int(owner[i] % 3);
        double dir[3] = {0.0, 0.0, 0.0};
        if (mode < 3) {
          dir[mode] = 1.0;
        } else {
          double spin[3] = {0.0, 0.0, 0.0}, r[3];
          spin[mode - 3] = 1.0;
          for (int a = 0; a < 3; ++a) r[a] = rest[3 * node + a] - centre[a];
          dir[0] = spin[1] * r[2] - spin[2] * r[1];
          dir[1] = spin[2] * r[0] - spin[0] * r[2];
          dir[2] = spin[0] * r[1] - spin[1] * r[0];
        }
        shape[i] = dir[axis];
      }
      deflate(shape, rigid, w);
      const double norm = std::sqrt(mass_dot(shape, shape, w));
      if (norm > 0.0) {
        for (double &v : shape) v /= norm;
        rigid.push_back(std::move(shape));
      }
    }
  }
  double traceK = 0.0, sumw = 0.0;
  for (size_t i = 0; i < n; ++i) {
    traceK += K[i * W];
    sumw += w[i];
  }
  const double shift = held ? 0.0 : 1.0e-6 * traceK / sumw;
  std::vector<double> L(K), D(n);
  for (size_t i = 0; i < n; ++i) L[i * W] += shift * w[i];
  for (size_t j = 0; j < n; ++j) {  // band L D L^T, in place
    const size_t lo = j > bw ? j - bw : 0;
    double d = L[j * W];
    for (size_t k = lo; k < j; ++k) d -= L[j * W + (j - k)] * L[j * W + (j - k)] * D[k];
    if (d == 0.0) return 0.0;
    D[j] = d;
    const size_t hi = std::min(j + bw, n - 1);
    for (size_t i = j + 1; i <= hi; ++i) {
      double v = L[i * W + (i - j)];
      for (size_t k = std::max(i > bw ? i - bw : 0, lo); k < j; ++k) v -= L[i * W + (i - k)] * L[j * W + (j - k)] * D[k];
      L[i * W + (i - j)] = v / d;
    }
  }
  std::vector<double> x(n), z(n), kz(n);
  for (size_t i = 0; i < n; ++i) x[i] = std::sin(1.7 * double(i) + 0.3) + 0.5;
  double previous = INF, value = INF;
  for (uint32_t it = 0; it < 300; ++it) {  // ITERATIONS, TOLERANCE (lowestmode.cc:16-17)
    deflate(x, rigid, w);
    const double norm = std::sqrt(mass_dot(x, x, w));
    for (double &v : x) v /= norm;
    for (size_t i = 0; i < n; ++i) {
      double v = x[i] * w[i];
      for (size_t k = i > bw ? i - bw : 0; k < i; ++k) v -= L[i * W + (i - k)] * z[k];
      z[i] = v;
    }
    for (size_t i = 0; i < n; ++i) z[i] /= D[i];
    for (size_t i = n; i-- > 0;) {
      double v = z[i];
      const size_t hi = std::min(i + bw, n - 1);
      for (size_t k = i + 1; k <= hi; ++k) v -= L[k * W + (k - i)] * z[k];
      z[i] = v;
    }
    deflate(z, rigid, w);
    std::fill(kz.begin(), kz.end(), 0.0);
    for (size_t p = 0; p < n; ++p) {
      kz[p] += K[p * W] * z[p];
      for (size_t k = 1; k <= bw && k <= p; ++k) {
        const double a = K[p * W + k];
        kz[p] += a * z[p - k];
        kz[p - k] += a * z[p];
      }
    }
    double zkz = 0.0;
    for (size_t i = 0; i < n; ++i) zkz += z[i] * kz[i];
    value = zkz / mass_dot(z, z, w);
    x = z;
    if (std::fabs(value - previous) <= 1.0e-12 * std::fabs(value)) break;
    previous = value;
  }
  return std::sqrt(value > 0.0 ? value : 0.0);
}

HexDamping LatticeBody::cell_damping(uint32_t cell_index, double step) const {
  const double target = material_of(cell_index).damping_ratio;
  HexDamping d{target, std::isfinite(lowest) ? lowest : 0.0, 1.0, 0.0};
  const bool exact = !(step * stiffest() > 1.5);
  if (exact && lowest > 0.0 && std::isfinite(lowest) && std::isfinite(target)) {
    d.theta = clampd(2.0 * target / (step * lowest), 0.5, 1.0);  // MIDPOINT (strainstep.cc:58)
    d.own = d.theta * step;
  }
  return d;
}

double LatticeBody::lowest_angular(double step) const {
  std::vector<double> inv(nodes, 0.0), m(nodes);
  std::vector<uint8_t> held(nodes);
  for (size_t i = 0; i < nodes; ++i) {
    if (!pinned(i) && mass[i] > 0.0f) inv[i] = 1.0 / double(mass[i]);
    m[i] = mass[i];
    held[i] = pinned(i) ? 1 : 0;
  }
  const bool kept = rest_vol.size() == hexes;
  std::vector<HexElement> plain;
  plain.reserve(hexes);
  for (uint32_t c = 0; c < hexes; ++c) {
    double shape[9], hgv[12], vol, w[8], shear, firstl;
    plastic_rest(c, shape, hgv, &vol);
    if (!hex_fit_holds(shape, hgv) && kept && rest_vol[c] > 0.0) {
      std::memcpy(shape, &rest_shape[9 * size_t(c)], 9 * sizeof(double));
      std::memcpy(hgv, &rest_hg[12 * size_t(c)], 12 * sizeof(double));
      vol = rest_vol[c];
    }
    const uint32_t *nd = &hex[8 * size_t(c)];
    for (uint32_t k = 0; k < 8; ++k) w[k] = inv[nd[k]];
    const talus_mi_material &mat = material_of(c);
    hex_material(mat.youngs_modulus, mat.poissons_ratio, &shear, &firstl);
    HexElement e;
    if (hex_make(e, c, nd, w, shape, hgv, vol, shear, firstl, step)) plain.push_back(e);
  }
  return hex_lowest_angular(plain, m, held, rest);
}

// ---- cell tearing (strainstep/tearing.cc) ---------------------------------

namespace {

constexpr uint32_t MIN_PIECE_CELLS = 2;  // tearing.h:13

// FACES (tearing.cc:27-34): the corners of face 2 axis + side
constexpr uint32_t FACES[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5},
                                  {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};

using Face = std::array<uint32_t, 4>;

uint32_t root_of(std::vector<uint32_t> &parent, uint32_t at) {
  while (parent[at] != at) {
    parent[at] = parent[parent[at]];
    at = parent[at];
  }
  return at;
}

// the mesh being torn: the body's corner lists with, per node, the cells
// holding it in ascending order (cellsAt, kept current as nodes split)
struct Tearing {
  LatticeBody &b;
  std::vector<std::vector<uint32_t>> cells_of;
  std::vector<std::pair<uint32_t, uint32_t>> copies;  // (copy, original)

  explicit Tearing(LatticeBody &body) : b(body), cells_of(body.nodes) {
    for (uint32_t c = 0; c < b.hexes; ++c)
      for (int k = 0; k < 8; ++k) {
        auto &list = cells_of[b.hex[8 * size_t(c) + k]];
        if (list.empty() || list.back() != c) list.push_back(c);
      }
  }
  const uint32_t *nodes_of(uint32_t cell) const { return &b.hex[8 * size_t(cell)]; }
  bool has(uint32_t cell, uint32_t node) const {
    const uint32_t *nd = nodes_of(cell);
    for (int k = 0; k < 8; ++k)
      if (nd[k] == node) return true;
    return false;
  }
  Face face_of(uint32_t cell, uint32_t face) const {
    Face f;
    for (int k = 0; k < 4; ++k) f[k] = nodes_of(cell)[FACES[face][k]];
    std::sort(f.begin(), f.end());
    return f;
  }
  uint32_t shared_nodes(uint32_t x, uint32_t y) const {
    uint32_t n = 0;
    for (int k = 0; k < 8; ++k) n += has(y, nodes_of(x)[k]) ? 1u : 0u;
    return n;
  }
  // groupsOf (tearing.cc:92-137)
  uint32_t groups_of(const std::vector<uint32_t> &fan, const std::vector<uint8_t> &sides,
                     std::vector<uint32_t> &group) const {
    const uint32_t n = uint32_t(fan.size());
    std::vector<uint32_t> parent(n), number(n, UINT32_MAX);
    for (uint32_t i = 0; i < n; ++i) parent[i] = i;
    for (uint32_t i = 0; i < n; ++i)
      for (uint32_t j = i + 1; j < n; ++j)
        if (sides[i] == sides[j] && shared_nodes(fan[i], fan[j]) >= 4) {
          const uint32_t x = root_of(parent, i), y = root_of(parent, j);
          if (x != y) parent[std::max(x, y)] = std::min(x, y);
        }
    group.assign(n, 0);
    uint32_t count = 0;
    for (uint8_t side = 0; side < 2; ++side)
      for (uint32_t i = 0; i < n; ++i) {
        if (sides[i] != side) continue;
        const uint32_t r = root_of(parent, i);
        if (number[r] == UINT32_MAX) number[r] = count++;
        group[i] = number[r];
      }
    return count;
  }
  // reachable (tearing.cc:139-180): cells the seeds reach through faces that
  // do not hold the pivot, stopping once limit are seen
  size_t reachable(const std::vector<uint32_t> &seeds, uint32_t pivot, size_t limit) const {
    std::vector<uint32_t> seen(seeds.begin(), seeds.end()), stack(seeds.begin(), seeds.end());
    std::sort(seen.begin(), seen.end());
    seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
    auto saw = [&](uint32_t c) { return std::binary_search(seen.begin(), seen.end(), c); };
    while (!stack.empty() && seen.size() < limit) {
      const uint32_t cell = stack.back();
      stack.pop_back();
      for (uint32_t face = 0; face < 6; ++face) {
        const Face f = face_of(cell, face);
        if (f[0] == pivot || f[1] == pivot || f[2] == pivot || f[3] == pivot) continue;
        for (const uint32_t other : cells_of[f[0]]) {
          if (saw(other)) continue;
          bool match = false;
          for (uint32_t g = 0; g < 6 && !match; ++g) match = face_of(other, g) == f;
          if (match) {
            seen.insert(std::upper_bound(seen.begin(), seen.end(), other), other);
            stack.push_back(other);
          }
        }
      }
    }
    return seen.size();
  }
  // splitNode (tearing.cc:182-212)
  uint32_t split_node(uint32_t node) {
    const uint32_t copy = b.nodes;
    for (int k = 0; k < 3; ++k) {
      b.rest.push_back(b.rest[3 * size_t(node) + k]);
      b.x.push_back(b.x[3 * size_t(node) + k]);
      b.v.push_back(b.v[3 * size_t(node) + k]);
    }
    b.mass.push_back(0.0f);
    b.anchored.push_back(0);
    b.node_colour.push_back(b.node_colour[node]);
    cells_of.emplace_back();
    ++b.nodes;
    return copy;
  }
  // split (tearing.cc:214-249)
  bool split(uint32_t node, const std::vector<uint32_t> &fan, const std::vector<uint32_t> &group, uint32_t count) {
    if (count < 2 || b.anchored[node]) return false;
    for (uint32_t piece = 1; piece < count; ++piece) {
      const uint32_t copy = split_node(node);
      copies.emplace_back(copy, node);
      for (size_t at = 0; at < fan.size(); ++at) {
        if (group[at] != piece) continue;
        uint32_t *nd = &b.hex[8 * size_t(fan[at])];
        for (int k = 0; k < 8; ++k)
          if (nd[k] == node) nd[k] = copy;
        auto &list = cells_of[node];
        list.erase(std::find(list.begin(), list.end(), fan[at]));
        cells_of[copy].push_back(fan[at]);  // fan is ascending
      }
    }
    return true;
  }
  // crackThroughCell (tearing.cc:302-376)
  bool crack_through_cell(uint32_t cell) {
    double s, across[3], centre[3] = {0.0, 0.0, 0.0};
    b.cell_stretch(cell, &s, across);
    uint32_t corners[8];
    std::copy(nodes_of(cell), nodes_of(cell) + 8, corners);
    for (int k = 0; k < 8; ++k)
      for (int a = 0; a < 3; ++a) centre[a] += b.rest[3 * size_t(corners[k]) + a] / 8.0;
    std::array<std::pair<double, uint32_t>, 8> order;
    for (int k = 0; k < 8; ++k) {
      double r[3];
      for (int a = 0; a < 3; ++a) r[a] = b.rest[3 * size_t(corners[k]) + a] - centre[a];
      order[k] = {std::fabs(dot3(r, across)), corners[k]};
    }
    std::sort(order.begin(), order.end());
    std::vector<uint32_t> group, seeds;
    std::vector<uint8_t> sides;
    for (const auto &entry : order) {
      const uint32_t node = entry.second;
      if (b.anchored[node]) continue;
      const std::vector<uint32_t> fan = cells_of[node];
      sides.assign(fan.size(), 0);
      bool low = false, high = false;
      const double *rn = &b.rest[3 * size_t(node)];
      for (size_t at = 0; at < fan.size(); ++at) {
        double offset = 0.0;
        for (int c = 0; c < 8; ++c) {
          const double *ro = &b.rest[3 * size_t(nodes_of(fan[at])[c])];
          const double r[3] = {ro[0] - rn[0], ro[1] - rn[1], ro[2] - rn[2]};
          offset += dot3(r, across);
        }
        sides[at] = offset > 0.0 ? 1 : 0;
        low = low || sides[at] == 0;
        high = high || sides[at] == 1;
      }
      if (!low || !high) continue;
      const uint32_t count = groups_of(fan, sides, group);
      bool whole = true;
      for (uint32_t piece = 0; piece < count && whole; ++piece) {
        seeds.clear();
        for (size_t at = 0; at < fan.size(); ++at)
          if (group[at] == piece) seeds.push_back(fan[at]);
        whole = reachable(seeds, node, MIN_PIECE_CELLS) >= MIN_PIECE_CELLS;
      }
      if (whole && split(node, fan, group, count)) return true;
    }
    return false;
  }
  // splitPinched (tearing.cc:378-408)
  void split_pinched() {
    std::vector<uint8_t> near(b.nodes, 0), candidate(b.nodes, 0);
    for (const auto &cp : copies) near[cp.first] = near[cp.second] = 1;
    for (uint32_t c = 0; c < b.hexes; ++c) {
      const uint32_t *nd = nodes_of(c);
      bool touches = false;
      for (int k = 0; k < 8; ++k) touches = touches || near[nd[k]];
      if (touches)
        for (int k = 0; k < 8; ++k) candidate[nd[k]] = 1;
    }
    const uint32_t count = b.nodes;
    std::vector<uint32_t> group;
    for (uint32_t node = 0; node < count; ++node) {
      if (!candidate[node] || cells_of[node].size() < 2) continue;
      const std::vector<uint32_t> fan = cells_of[node];
      const std::vector<uint8_t> sides(fan.size(), 0);
      const uint32_t groups = groups_of(fan, sides, group);
      (void)split(node, fan, group, groups);
    }
  }
  // recomputeMasses (tearing.cc:410-424), in single precision
  void recompute_masses() {
    std::fill(b.mass.begin(), b.mass.end(), 0.0f);
    for (uint32_t c = 0; c < b.hexes; ++c) {
      const float density = b.lane[c] < b.materials.size() ? b.materials[b.lane[c]].density : 1000.0f;
      const float share = density * b.volume[c] * 0.125f;
      for (int k = 0; k < 8; ++k) b.mass[nodes_of(c)[k]] += share;
    }
  }
  // markCracks (tearing.cc:426-484): crack faces, their nodes put on the
  // surface, then each copy of a surface node
  void mark_cracks() {
    b.surface.resize(b.nodes, 0);
    float lo[3] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                   std::numeric_limits<float>::max()};
    for (uint32_t i = 0; i < b.nodes; ++i)
      for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], float(b.rest[3 * size_t(i) + k]));
    constexpr int64_t OFFSET = 1 << 20;
    auto key_of = [](const int32_t *at) {
      return uint64_t(at[0] + OFFSET) | uint64_t(at[1] + OFFSET) << 21 | uint64_t(at[2] + OFFSET) << 42;
    };
    std::vector<int32_t> index(3 * size_t(b.hexes));
    std::unordered_map<uint64_t, uint32_t> cell_at;
    for (uint32_t c = 0; c < b.hexes; ++c) {
      for (int k = 0; k < 3; ++k)
        index[3 * size_t(c) + k] =
            int32_t(std::round((float(b.rest[3 * size_t(nodes_of(c)[0]) + k]) - lo[k]) / b.voxel[k]));
      cell_at.emplace(key_of(&index[3 * size_t(c)]), c);
    }
    b.crack.assign(b.hexes, 0);
    for (uint32_t c = 0; c < b.hexes; ++c)
      for (uint32_t axis = 0; axis < 3; ++axis)
        for (uint32_t side = 0; side < 2; ++side) {
          int32_t next[3] = {index[3 * size_t(c)], index[3 * size_t(c) + 1], index[3 * size_t(c) + 2]};
          next[axis] += side == 0 ? -1 : 1;
          const auto found = cell_at.find(key_of(next));
          if (found == cell_at.end()) continue;
          const Face mine = face_of(c, 2 * axis + side);
          if (mine != face_of(found->second, 2 * axis + (1 - side))) {
            b.crack[c] = uint8_t(b.crack[c] | (1u << (2 * axis + side)));
            for (const uint32_t n : mine) b.surface[n] = 1;
          }
        }
    for (const auto &cp : copies)
      if (b.surface[cp.second]) b.surface[cp.first] = 1;
  }
};

}  // namespace

void sym_eigen3(const double *A, double *w, double *V) {
  double a[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      a[r][c] = A[3 * r + c];
      V[3 * r + c] = r == c ? 1.0 : 0.0;
    }
  constexpr int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (!(off > 1.0e-40 * diag)) break;
    for (int k = 0; k < 3; ++k) {
      const int p = P[k], q = Q[k];
      const double apq = a[p][q];
      if (apq == 0.0) continue;
      const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
      const double t = std::fabs(theta) > 1.0e150
                           ? 0.5 / theta
                           : (theta < 0.0 ? -1.0 : 1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
      const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
      const int r = 3 - p - q;
      const double arp = a[r][p], arq = a[r][q];
      a[p][p] -= t * apq;
      a[q][q] += t * apq;
      a[p][q] = a[q][p] = 0.0;
      a[r][p] = a[p][r] = c * arp - s * arq;
      a[r][q] = a[q][r] = s * arp + c * arq;
      for (int i = 0; i < 3; ++i) {
        const double vp = V[3 * i + p], vq = V[3 * i + q];
        V[3 * i + p] = c * vp - s * vq;
        V[3 * i + q] = s * vp + c * vq;
      }
    }
  }
  for (int k = 0; k < 3; ++k) w[k] = a[k][k];
}

double LatticeBody::tear_limit(uint32_t cell_index) const {
  return 1.0 + double(material_of(cell_index).tear_strain);
}

void LatticeBody::cell_stretch(uint32_t cell_index, double *stretch, double *direction) const {
  const uint32_t *nd = &hex[8 * size_t(cell_index)];
  double best = 0.0, dir[3] = {1.0, 0.0, 0.0};
  for (uint32_t corner = 0; corner < 8; ++corner) {
    const uint32_t base = nd[corner];
    double Rm[9], N[9];
    for (uint32_t e = 0; e < 3; ++e) {
      const uint32_t other = nd[corner ^ (1u << e)];
      for (int k = 0; k < 3; ++k) {
        Rm[3 * k + e] = rest[3 * size_t(other) + k] - rest[3 * size_t(base) + k];
        N[3 * k + e] = x[3 * size_t(other) + k] - x[3 * size_t(base) + k];
      }
    }
    const double det = Rm[0] * (Rm[4] * Rm[8] - Rm[5] * Rm[7]) - Rm[1] * (Rm[3] * Rm[8] - Rm[5] * Rm[6]) +
                       Rm[2] * (Rm[3] * Rm[7] - Rm[4] * Rm[6]);
    if (!(std::fabs(det) > 0.0)) continue;
    double Ri[9], F[9], C[9], w[3], V[9];
    inverse3(Rm, Ri);
    mm3(N, Ri, F);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) C[3 * r + c] = F[r] * F[c] + F[3 + r] * F[3 + c] + F[6 + r] * F[6 + c];
    sym_eigen3(C, w, V);
    int top = 0;
    for (int k = 1; k < 3; ++k)
      if (w[k] > w[top]) top = k;
    const double s = std::sqrt(w[top] > 0.0 ? w[top] : 0.0);
    if (s > best) {
      const double v[3] = {V[top], V[3 + top], V[6 + top]};
      const double len = std::sqrt(dot3(v, v));
      int big = 0;
      for (int k = 1; k < 3; ++k)
        if (std::fabs(v[k]) > std::fabs(v[big])) big = k;
      const double sign = v[big] < 0.0 ? -1.0 : 1.0;
      for (int k = 0; k < 3; ++k) dir[k] = sign * v[k] / len;
      best = s;
    }
  }
  *stretch = best;
  std::copy(dir, dir + 3, direction);
}

void LatticeBody::ensure_surface() {
  if (!surface.empty()) return;
  std::map<Face, uint32_t> faces;
  for (uint32_t c = 0; c < hexes; ++c)
    for (uint32_t face = 0; face < 6; ++face) {
      Face f;
      for (int k = 0; k < 4; ++k) f[k] = hex[8 * size_t(c) + FACES[face][k]];
      std::sort(f.begin(), f.end());
      ++faces[f];
    }
  surface.assign(nodes, 0);
  for (const auto &f : faces)
    if (f.second == 1)
      for (const uint32_t n : f.first) surface[n] = 1;
}

uint32_t LatticeBody::tear_cells(const uint32_t *cells, uint32_t count) {
  if (count == 0) return 0;
  ensure_surface();
  Tearing t(*this);
  for (uint32_t i = 0; i < count; ++i)
    if (cells[i] < hexes) (void)t.crack_through_cell(cells[i]);
  if (t.copies.empty()) return 0;
  t.split_pinched();
  t.recompute_masses();
  t.mark_cracks();
  retopo = true;
  return uint32_t(t.copies.size());
}

uint32_t LatticeBody::pieces() const {
  std::vector<uint32_t> parent(hexes), owner(nodes, UINT32_MAX);
  for (uint32_t c = 0; c < hexes; ++c) parent[c] = c;
  for (uint32_t c = 0; c < hexes; ++c)
    for (int k = 0; k < 8; ++k) {
      const uint32_t node = hex[8 * size_t(c) + k];
      if (owner[node] == UINT32_MAX) {
        owner[node] = c;
        continue;
      }
      const uint32_t x0 = root_of(parent, owner[node]), y0 = root_of(parent, c);
      if (x0 != y0) parent[std::max(x0, y0)] = std::min(x0, y0);
    }
  uint32_t n = 0;
  for (uint32_t c = 0; c < hexes; ++c) n += root_of(parent, c) == c ? 1u : 0u;
  return n;
}

void LatticeBody::rebuild_torn() {
  std::vector<HexElement> old;
  old.swap(elements);
  build_elements(h);
  if (old.size() == elements.size())
    for (size_t i = 0; i < old.size(); ++i)
      if (elements[i].cell == old[i].cell) {
        std::copy(old[i].impulse, old[i].impulse + HEX_ROWS, elements[i].impulse);
        elements[i].vol_impulse = old[i].vol_impulse;
        std::copy(old[i].quat, old[i].quat + 4, elements[i].quat);
      }
  retopo = false;
}

int make_lattice(LatticeBody &b, const talus_mi_hex_mesh *mesh, const talus_mi_material *materials,
                 uint32_t lanes, const double *offset, const double *velocity) {
  if (!mesh || !mesh->rest_positions || !mesh->node_mass || !mesh->anchored || !mesh->hex_nodes ||
      !mesh->volume || !mesh->hex_material)
    return fail(TALUS_MI_ERR_ARGUMENT, "incomplete hex mesh");
  if (mesh->nodes == 0 || mesh->hexes == 0) return fail(TALUS_MI_ERR_ARGUMENT, "empty hex mesh");
  if (!materials || lanes == 0)
    return fail(TALUS_MI_ERR_ARGUMENT, "a strainstep lattice needs its material lanes");
  for (uint32_t c = 0; c < 8 * mesh->hexes; ++c)
    if (mesh->hex_nodes[c] >= mesh->nodes) return fail(TALUS_MI_ERR_ARGUMENT, "hex node out of range");
  for (int k = 0; k < 3; ++k)
    if (!(mesh->voxel_size[k] > 0.0f)) return fail(TALUS_MI_ERR_ARGUMENT, "voxel size must be positive");
  b = LatticeBody{};
  b.nodes = mesh->nodes;
  b.hexes = mesh->hexes;
  const size_t n = b.nodes;
  b.rest.resize(3 * n);
  for (size_t i = 0; i < 3 * n; ++i) b.rest[i] = double(mesh->rest_positions[i]);
  b.mass.assign(mesh->node_mass, mesh->node_mass + n);
  b.anchored.assign(mesh->anchored, mesh->anchored + n);
  b.hex.assign(mesh->hex_nodes, mesh->hex_nodes + 8 * size_t(b.hexes));
  b.volume.assign(mesh->volume, mesh->volume + b.hexes);
  b.lane.assign(mesh->hex_material, mesh->hex_material + b.hexes);
  b.materials.assign(materials, materials + lanes);
  const double vox[3] = {mesh->voxel_size[0], mesh->voxel_size[1], mesh->voxel_size[2]};
  for (int k = 0; k < 3; ++k) b.voxel[k] = mesh->voxel_size[k];
  b.cell = std::min({vox[0], vox[1], vox[2]});
  double lo[3] = {INF, INF, INF};
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], b.rest[3 * i + k]);
  b.x.resize(3 * n);
  b.v.resize(3 * n);
  b.node_colour.resize(n);
  for (size_t i = 0; i < n; ++i) {
    uint32_t colour = 0;
    for (int k = 0; k < 3; ++k) {
      b.x[3 * i + k] = b.rest[3 * i + k] + (offset ? offset[k] : 0.0);
      b.v[3 * i + k] = b.anchored[i] ? 0.0 : (velocity ? velocity[k] : 0.0);
      const double index = std::round((b.rest[3 * i + k] - lo[k]) / vox[k]);
      colour |= (uint32_t(int64_t(index)) & 1u) << k;
    }
    b.node_colour[i] = uint8_t(colour);
  }
  for (uint32_t c = 0; c < b.hexes; ++c) b.yields = b.yields || std::isfinite(b.material_of(c).yield_stress);
  for (uint32_t c = 0; c < b.hexes; ++c) b.tears = b.tears || std::isfinite(b.tear_limit(c));
  for (uint32_t c = 0; c < b.hexes; ++c) b.damped = b.damped || std::isfinite(b.material_of(c).damping_ratio);
  if (b.damped) {
    // the lowest-mode solve factors a band of the lattice's stiffness: its
    // cost (free coordinates x band width squared) and size are bounded
    std::vector<int64_t> index(n, -1);
    int64_t free_nodes = 0;
    for (const uint32_t i : band_order(b.rest, n))
      if (!b.pinned(i) && b.mass[i] > 0.0f) index[i] = free_nodes++;
    int64_t band = 0;
    for (uint32_t c = 0; c < b.hexes; ++c)
      for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
          const int64_t p = index[b.hex[8 * size_t(c) + i]], q = index[b.hex[8 * size_t(c) + j]];
          if (p >= 0 && q >= 0) band = std::max(band, 3 * std::abs(p - q) + 2);
        }
    const double coords = 3.0 * double(free_nodes), width = double(band + 1);
    if (coords * width * width > 4.0e10 || coords * width > 1.0e8)
      return fail(TALUS_MI_ERR_ARGUMENT,
                  "a damped lattice this large exceeds the banded lowest-mode solve (damping ratio must be infinite)");
  }
  b.hex_colour.resize(b.hexes);
  for (uint32_t c = 0; c < b.hexes; ++c) b.hex_colour[c] = b.node_colour[b.hex[8 * size_t(c)]];
  // a colour's hexes are solved at once: no two of them may share a node
  {
    std::vector<uint32_t> seen(size_t(n) * 8, UINT32_MAX);
    for (uint32_t c = 0; c < b.hexes; ++c)
      for (uint32_t k = 0; k < 8; ++k) {
        uint32_t &slot = seen[size_t(b.hex[8 * size_t(c) + k]) * 8 + b.hex_colour[c]];
        if (slot != UINT32_MAX && slot != c)
          return fail(TALUS_MI_ERR_ARGUMENT,
                      "two hexes of one colour share a node: the mesh is not a lattice of its voxel size");
        slot = c;
      }
  }
  return TALUS_MI_OK;
}

}  // namespace tmi

extern "C" {

int talus_mi_build_hex_lattice(const talus_mi_lattice *lattice, const talus_mi_material *materials,
                               uint32_t lanes, const talus_mi_support *supports, uint32_t nsupports,
                               talus_mi_hex_mesh *out) {
  if (!lattice || !out || (nsupports && !supports))
    return tmi::fail(TALUS_MI_ERR_ARGUMENT, "null argument");
  try {
    return tmi::build_hex_lattice(lattice, materials, lanes, supports, nsupports, out);
  } catch (const std::bad_alloc &) {
    return tmi::fail(TALUS_MI_ERR_MEMORY, "out of host memory");
  }
}

void talus_mi_hex_mesh_free(talus_mi_hex_mesh *mesh) { tmi::hex_mesh_free(mesh); }

int talus_mi_hex_element_make(const double *corners, const double *inverse_mass, double volume,
                              float youngs_modulus, float poissons_ratio, double h,
                              talus_mi_hex_element *out) {
  if (!corners || !inverse_mass || !out) return tmi::fail(TALUS_MI_ERR_ARGUMENT, "null argument");
  double shape[9], hg[12], shear, first;
  tmi::hex_corner_rest(corners, shape, hg);
  tmi::hex_material(youngs_modulus, poissons_ratio, &shear, &first);
  const uint32_t nodes[8] = {0, 1, 2, 3, 4, 5, 6, 7};
  tmi::HexElement e;
  if (!tmi::hex_make(e, 0, nodes, inverse_mass, shape, hg, volume, shear, first, h))
    return tmi::fail(TALUS_MI_ERR_ARGUMENT, "the hex builds no element");
  std::memset(out, 0, sizeof *out);
  std::memcpy(out->gradient, e.gradient, sizeof out->gradient);
  std::memcpy(out->inverse_mass, e.inv_mass, sizeof out->inverse_mass);
  std::memcpy(out->rest_hourglass, e.rest_hg, sizeof out->rest_hourglass);
  out->bend = e.bend;
  out->size = e.size;
  std::memcpy(out->stiffness, e.stiffness, sizeof out->stiffness);
  std::memcpy(out->rate, e.rate, sizeof out->rate);
  std::memcpy(out->softness, e.softness, sizeof out->softness);
  std::memcpy(out->cap, e.cap, sizeof out->cap);
  std::memcpy(out->inverse, e.inverse.data(), sizeof out->inverse);
  out->volume_stiffness = e.vol_stiffness;
  out->volume_rate = e.vol_rate;
  out->volume_softness = e.vol_softness;
  out->volume_cap = e.vol_cap;
  out->volumetric = e.volumetric ? 1u : 0u;
  return TALUS_MI_OK;
}

}  // extern "C"
