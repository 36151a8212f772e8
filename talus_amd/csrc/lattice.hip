// lattice.hip -- gfx950 kernels of the strainstep hex-lattice solve and the
// device set of lattice bodies they step.
//
// Reference (under src/talus/dynamics/strainstep/):
//   element.cc:113-158  addRows / rowForce
//   element.cc:344-372  rotationOf
//   element.cc:374-459  updateElement
//   element.cc:461-523  applyElement / solveElement
//   strainstep.cc:61-84 integrateNodes / advanceNodes
//   strainstep.cc:708-801 the substep loop: integrate, update and warm-start
//                       every element, PASSES = 4 Gauss-Seidel passes of
//                       solveElement, advance
//
// Layout in HBM.  All bodies are concatenated.  Node state x, v is FP64
// xyz-interleaved.  Elements are stored in slot order: colour 0..7, then
// body, then cell, so one colour is one contiguous slot range and one
// launch.  Per-element data are element-major, field[slot * width + k], and
// 18 lanes work on one element (lane r its row r, lanes 0..7 also its
// corners), so the lanes of an element touch consecutive words.  The 18x18
// row-block inverses are interned: equal inverses (same rest shape, corner
// masses, material and substep) share one table entry, indexed per element.
// An entry is the upper triangle of the inverse, row by row (171 doubles,
// 1368 bytes instead of 2592): the row block is symmetric, so entry (k, r)
// with k < r stands for (r, k), which the Gauss-Jordan elimination gives
// equal up to rounding.  An element's entry is staged in LDS by its 18 lanes
// with coalesced loads before the rows use it.  Gradients are not stored:
// corner c's gradient is (shape^-T s_c) / 8 with s_c = +-1, recomputed
// exactly from the 9 entries of shape^-1.
//
// Order.  The reference sweeps a body's elements in cell order, one after
// another (Gauss-Seidel).  Here the passes sweep colour by colour: the hex
// colour is the parity of its corner-0 lattice index (makeLattice's node
// colour, bodyset.cc:30-37), two hexes of one colour share no node, so every
// element of a colour is solved at once with no write conflict, and the
// result is that of the sequential sweep in (colour, cell) order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "device.h"
#include "hexlattice.h"
#include "host.h"
#include "tmath.h"

namespace tmi {

namespace {

constexpr int NODE_BLOCK = 256;
constexpr uint32_t EBLOCK = 64;         // one wave per element block
constexpr uint32_t EPW = 3;             // elements per wave: 18 lanes each
constexpr uint32_t PASSES = 4;
constexpr uint32_t ROTATION_ITERATIONS = 16;
constexpr double MAX_SPEED = 400.0;
constexpr uint32_t R = HEX_ROWS;
constexpr uint32_t PACKED = R * (R + 1) / 2;  // upper triangle of an 18x18 inverse

// position of entry (i, j), i <= j, in the row-major upper triangle
__device__ __forceinline__ uint32_t packed_at(uint32_t i, uint32_t j) {
  return i * R - (i * (i - 1)) / 2 + (j - i);
}

// an element's packed inverse goes to LDS in two steps, so its loads are
// issued with the element's other loads: lane r loads entries r, r + 18, ...
constexpr uint32_t STAGE = (PACKED + R - 1) / R;
__device__ __forceinline__ void load_inverse(const double *packed, uint32_t r, double *reg) {
#pragma unroll
  for (uint32_t k = 0; k < STAGE; ++k) {
    const uint32_t i = r + k * R;
    reg[k] = i < PACKED ? packed[i] : 0.0;
  }
}
__device__ __forceinline__ void store_inverse(const double *reg, uint32_t r, double *dst) {
#pragma unroll
  for (uint32_t k = 0; k < STAGE; ++k) {
    const uint32_t i = r + k * R;
    if (i < PACKED) dst[i] = reg[k];
  }
}

// row r of Inv * x with Inv symmetric, from its upper triangle in LDS
__device__ __forceinline__ double inverse_row(const double *tri, uint32_t r, const double *x) {
  double t = 0.0;
#pragma unroll
  for (uint32_t k = 0; k < R; ++k) t += tri[k < r ? packed_at(k, r) : packed_at(r, k)] * x[k];
  return t;
}

struct LatDev {
  uint32_t N, E;
  double *x, *v;
  const uint32_t *node_body;
  const uint8_t *node_pinned;
  const double *body_h;
  const uint32_t *body_substeps;
  const uint8_t *body_exact;
  // elements, element-major [slot * width + k]; widths below
  const uint32_t *e_body;      // 1 plane
  const uint32_t *e_sub;       // the body's substeps, bit 31 = the body is exact
  const uint32_t *e_nodes;     // 8
  const uint32_t *e_inv_id;    // 1
  const double *e_shape_inv;   // 9
  const double *e_inv_mass;    // 8
  const double *e_bend;        // 1
  const double *e_rest_hg;     // 12
  const double *e_rate;        // 18
  const double *e_soft;        // 18
  const double *e_cap;         // 18
  const double *e_vol;         // 3: volume rate, softness, cap
  const uint8_t *e_volumetric; // 1
  const double *inv_table;     // [id][324]
  double *e_impulse;           // 18
  double *e_vimp;              // 1
  double *e_quat;              // 4
  double *e_turn;              // 9
  double *e_bias;              // 18: rate * strain
  double *e_vgrad;             // 24
  double *e_coupling;          // 18
  double *e_image;             // 18
  double *e_scal;              // 3: gain, rhs, 1 / schur (vweight + gain before the first pass)
  double *e_dv;                // 24: warm-start corner velocity changes
  // per node, its (slot * 8 + corner) entries with a free corner, in slot
  // (so colour) order: the warm start gathers them node by node
  const uint32_t *warm_start, *warm_entry;
  // plasticity, per cell of every body (bodies one after another, cells in
  // order, with or without an element): corner nodes, half extents, the
  // material (shear, bulk, H / 3 shear, yield stress, H, plastic cap, crush
  // strain, crush limit; yield infinite when the lane does not yield), and
  // per Gauss point P (9) and the plastic strain, and a moved flag
  uint32_t C;
  const uint32_t *c_nodes;  // 8
  const double *c_half;     // 3
  const double *c_pmat;     // 8
  double *c_plastic;        // 72
  double *c_hard;           // 8
  uint32_t *c_moved;        // 1
  // tearing, per cell of every body (the same order as the plasticity cells)
  // when a lane tears: corner nodes, corner rest positions, 1 + the lane's
  // tear strain (infinite when it never tears) and the overstretched flag
  uint32_t T;
  const uint32_t *t_nodes;  // 8
  const double *t_rest;     // 24
  const double *t_limit;    // 1
  uint32_t *t_torn;         // 1
};

__device__ __forceinline__ double corner_sign(uint32_t c, int axis) {
  return ((c >> axis) & 1u) ? 1.0 : -1.0;
}

__device__ __forceinline__ double hg_sign(uint32_t m, uint32_t c) {
  const double sx = corner_sign(c, 0), sy = corner_sign(c, 1), sz = corner_sign(c, 2);
  return m == 0 ? sx * sy : (m == 1 ? sy * sz : (m == 2 ? sz * sx : sx * sy * sz));
}

// corner gradient (shape^-T s_c) / 8; s_c = +-1 makes every product exact
__device__ __forceinline__ void grad_of(const double *inv, uint32_t c, double *g) {
  const double s0 = corner_sign(c, 0), s1 = corner_sign(c, 1), s2 = corner_sign(c, 2);
#pragma unroll
  for (int r = 0; r < 3; ++r) g[r] = (inv[r] * s0 + inv[3 + r] * s1 + inv[6 + r] * s2) / 8.0;
}

__device__ __forceinline__ void row_force(const double *imp, const double *g, double bend, uint32_t c,
                                          double *out) {
  out[0] = imp[0] * g[0] + 0.5 * (imp[3] * g[1] + imp[5] * g[2]);
  out[1] = imp[1] * g[1] + 0.5 * (imp[3] * g[0] + imp[4] * g[2]);
  out[2] = imp[2] * g[2] + 0.5 * (imp[4] * g[1] + imp[5] * g[0]);
#pragma unroll
  for (uint32_t m = 0; m < 4; ++m) {
    const double s = hg_sign(m, c) * bend;
    out[0] += s * imp[6 + 3 * m];
    out[1] += s * imp[7 + 3 * m];
    out[2] += s * imp[8 + 3 * m];
  }
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - b[1] * a[2];
  c[1] = a[2] * b[0] - b[2] * a[0];
  c[2] = a[0] * b[1] - b[0] * a[1];
}

__device__ __forceinline__ void quat_matrix(const double *q, double *T) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  const double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  const double wx = w * x, wy = w * y, wz = w * z;
  T[0] = 1.0 - 2.0 * (yy + zz);
  T[1] = 2.0 * (xy - wz);
  T[2] = 2.0 * (xz + wy);
  T[3] = 2.0 * (xy + wz);
  T[4] = 1.0 - 2.0 * (xx + zz);
  T[5] = 2.0 * (yz - wx);
  T[6] = 2.0 * (xz - wy);
  T[7] = 2.0 * (yz + wx);
  T[8] = 1.0 - 2.0 * (xx + yy);
}

__device__ __forceinline__ void quat_normalize(double *q) {
  const double len = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (len <= 0.0) {
    q[0] = 1.0;
    q[1] = q[2] = q[3] = 0.0;
    return;
  }
  const double s = 1.0 / len;
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] *= s;
}

// the rotation maximising tr(T^T F), iterated from the guess q
__device__ void rotation_of(const double *F, double *q) {
  quat_normalize(q);
  for (uint32_t it = 0; it < ROTATION_ITERATIONS; ++it) {
    double T[9];
    quat_matrix(q, T);
    double torque[3] = {0.0, 0.0, 0.0}, tr = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double t0 = T[k], t1 = T[3 + k], t2 = T[6 + k];
      const double f0 = F[k], f1 = F[3 + k], f2 = F[6 + k];
      torque[0] += t1 * f2 - f1 * t2;
      torque[1] += t2 * f0 - f2 * t0;
      torque[2] += t0 * f1 - f0 * t1;
      tr += t0 * f0 + t1 * f1 + t2 * f2;
    }
    const double scale = fabs(tr) + 1.0e-9;
    const double om0 = torque[0] / scale, om1 = torque[1] / scale, om2 = torque[2] / scale;
    const double angle = sqrt(om0 * om0 + om1 * om1 + om2 * om2);
    if (!(angle > 1.0e-12)) break;
    double s, c;
    sincos(angle * 0.5, &s, &c);
    const double p0 = c, p1 = om0 / angle * s, p2 = om1 / angle * s, p3 = om2 / angle * s;
    const double n0 = p0 * q[0] - p1 * q[1] - p2 * q[2] - p3 * q[3];
    const double n1 = p0 * q[1] + p1 * q[0] + p2 * q[3] - p3 * q[2];
    const double n2 = p0 * q[2] + p2 * q[0] + p3 * q[1] - p1 * q[3];
    const double n3 = p0 * q[3] + p3 * q[0] + p1 * q[2] - p2 * q[1];
    q[0] = n0;
    q[1] = n1;
    q[2] = n2;
    q[3] = n3;
    quat_normalize(q);
    if (angle <= 1.0e-9) break;
  }
}

__device__ __forceinline__ double clampd(double v, double lo, double hi) {
  return v < lo ? lo : (hi < v ? hi : v);
}

__global__ void __launch_bounds__(NODE_BLOCK) k_lat_integrate(LatDev d, uint32_t substep, double gx,
                                                             double gy, double gz) {
  const uint32_t i = blockIdx.x * NODE_BLOCK + threadIdx.x;
  if (i >= d.N) return;
  const uint32_t b = d.node_body[i];
  if (substep >= d.body_substeps[b]) return;
  const double h = d.body_h[b];
  double *vp = d.v + 3 * size_t(i);
  if (d.node_pinned[i]) {
    vp[0] = vp[1] = vp[2] = 0.0;
  } else {
    vp[0] = vp[0] + h * gx;
    vp[1] = vp[1] + h * gy;
    vp[2] = vp[2] + h * gz;
  }
}

// advanceNodes of substep `substep`, then integrateNodes of the next one
__global__ void __launch_bounds__(NODE_BLOCK) k_lat_advance(LatDev d, uint32_t substep, double gx, double gy,
                                                           double gz) {
  const uint32_t i = blockIdx.x * NODE_BLOCK + threadIdx.x;
  if (i >= d.N) return;
  const uint32_t b = d.node_body[i];
  const uint32_t subs = d.body_substeps[b];
  if (substep >= subs) return;
  const double h = d.body_h[b];
  double *vp = d.v + 3 * size_t(i);
  double *xp = d.x + 3 * size_t(i);
  double v0 = vp[0], v1 = vp[1], v2 = vp[2];
  const double speed = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
  if (speed > MAX_SPEED) {
    const double s = MAX_SPEED / speed;
    v0 *= s;
    v1 *= s;
    v2 *= s;
  }
  xp[0] += h * v0;
  xp[1] += h * v1;
  xp[2] += h * v2;
  if (substep + 1 < subs) {
    if (d.node_pinned[i]) {
      v0 = v1 = v2 = 0.0;
    } else {
      v0 = v0 + h * gx;
      v1 = v1 + h * gy;
      v2 = v2 + h * gz;
    }
  }
  vp[0] = v0;
  vp[1] = v1;
  vp[2] = v2;
}

// Element-major planes: row-indexed data of element e sit at [e * width + k],
// so the lanes of one element read and write consecutive words.
__device__ __forceinline__ size_t at(uint32_t e, uint32_t width, uint32_t k) {
  return size_t(e) * width + k;
}

// contribution of corner c's frame vector b (3) to row r of J b
__device__ __forceinline__ double row_term(uint32_t r, uint32_t c, const double *g, double bend,
                                           const double *b) {
  switch (r) {
    case 0: return b[0] * g[0];
    case 1: return b[1] * g[1];
    case 2: return b[2] * g[2];
    case 3: return 0.5 * (b[0] * g[1] + b[1] * g[0]);
    case 4: return 0.5 * (b[1] * g[2] + b[2] * g[1]);
    case 5: return 0.5 * (b[2] * g[0] + b[0] * g[2]);
    default: return hg_sign((r - 6) / 3, c) * bend * b[(r - 6) % 3];
  }
}

// updateElement at the current positions, one lane per element, for the
// elements [first, first + count) (of `order` when given) of the bodies whose
// exact flag equals
// `exact` (an exact body updates every element before the passes, a too-stiff
// body each colour right before its first solve).  Writes the rotation, the
// rate-scaled capped strain rows, the volume gradients, the coupling rows,
// gain, rhs and vweight + gain (the Schur complement's coupling term is
// folded in by the first pass's solve, which holds the row-block inverse),
// and the warm-start corner velocity changes applyElement makes with the
// accumulated impulses (k_lat_warm applies them colour by colour).
__global__ void __launch_bounds__(EBLOCK) k_lat_frame(LatDev d, const uint32_t *order, uint32_t first,
                                                      uint32_t count, uint32_t substep, uint32_t exact) {
  const uint32_t t = blockIdx.x * EBLOCK + threadIdx.x;
  if (t >= count) return;
  const uint32_t e = order ? order[first + t] : first + t;  // order: slots by descending substeps
  const uint32_t sub = d.e_sub[e];
  if (substep >= (sub & 0x7fffffffu) || (sub >> 31) != exact) return;
  double inv[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) inv[k] = d.e_shape_inv[at(e, 9, k)];
  const double bend = d.e_bend[e];
  double F[9] = {}, hg[12] = {};
#pragma unroll
  for (uint32_t c = 0; c < 8; ++c) {
    const uint32_t node = d.e_nodes[at(e, 8, c)];
    const double *xp = d.x + 3 * size_t(node);
    const double xc[3] = {xp[0], xp[1], xp[2]};
    double g[3];
    grad_of(inv, c, g);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) F[3 * i + j] += xc[i] * g[j];
#pragma unroll
    for (uint32_t m = 0; m < 4; ++m) {
      const double s = hg_sign(m, c);
#pragma unroll
      for (int i = 0; i < 3; ++i) hg[3 * m + i] += s * xc[i];
    }
  }
  double q[4], T[9];
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = d.e_quat[at(e, 4, k)];
  rotation_of(F, q);
  quat_matrix(q, T);
#pragma unroll
  for (int k = 0; k < 4; ++k) d.e_quat[at(e, 4, k)] = q[k];
#pragma unroll
  for (int k = 0; k < 9; ++k) d.e_turn[at(e, 9, k)] = T[k];
  const double det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                     F[2] * (F[3] * F[7] - F[4] * F[6]);
  const bool inverted = det < 0.0;
  {
    double L[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) L[3 * i + j] = T[i] * F[j] + T[3 + i] * F[3 + j] + T[6 + i] * F[6 + j];
    double raw[R];
    raw[0] = L[0] - 1.0;
    raw[1] = L[4] - 1.0;
    raw[2] = L[8] - 1.0;
    raw[3] = 0.5 * (L[1] + L[3]);
    raw[4] = 0.5 * (L[5] + L[7]);
    raw[5] = 0.5 * (L[6] + L[2]);
#pragma unroll
    for (uint32_t m = 0; m < 4; ++m)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double b = T[k] * hg[3 * m] + T[3 + k] * hg[3 * m + 1] + T[6 + k] * hg[3 * m + 2];
        raw[6 + 3 * m + k] = (b - d.e_rest_hg[at(e, 12, 3 * m + k)]) * bend;
      }
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) {
      const double cap0 = d.e_cap[at(e, R, r)];
      const double cap = inverted ? cap0 : (cap0 < 1.0 ? 1.0 : cap0);
      d.e_bias[at(e, R, r)] = d.e_rate[at(e, R, r)] * clampd(raw[r], -cap, cap);
    }
  }
  // cofactor columns F1 x F2, F2 x F0, F0 x F1
  double cof[9];
  {
    const double f0[3] = {F[0], F[3], F[6]}, f1[3] = {F[1], F[4], F[7]}, f2[3] = {F[2], F[5], F[8]};
    double c0[3], c1[3], c2[3];
    cross3(f1, f2, c0);
    cross3(f2, f0, c1);
    cross3(f0, f1, c2);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      cof[3 * i] = c0[i];
      cof[3 * i + 1] = c1[i];
      cof[3 * i + 2] = c2[i];
    }
  }
  double imp[R];
#pragma unroll
  for (uint32_t r = 0; r < R; ++r) imp[r] = d.e_impulse[at(e, R, r)];
  const double vimp = d.e_vimp[e];
  double coupling[R] = {};
  double vweight = 0.0;
#pragma unroll
  for (uint32_t c = 0; c < 8; ++c) {
    double g[3], vg[3];
    grad_of(inv, c, g);
#pragma unroll
    for (int i = 0; i < 3; ++i) vg[i] = cof[3 * i] * g[0] + cof[3 * i + 1] * g[1] + cof[3 * i + 2] * g[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) d.e_vgrad[at(e, 24, 3 * c + i)] = vg[i];
    const double w = d.e_inv_mass[at(e, 8, c)];
    vweight += w * (vg[0] * vg[0] + vg[1] * vg[1] + vg[2] * vg[2]);
    double l[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) l[k] = (T[k] * vg[0] + T[3 + k] * vg[1] + T[6 + k] * vg[2]) * w;
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) coupling[r] += row_term(r, c, g, bend, l);
    // warm start: this corner's velocity change from the accumulated impulses
    double dv[3] = {0.0, 0.0, 0.0};
    if (w != 0.0) {
      double lf[3];
      row_force(imp, g, bend, c, lf);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double f = T[3 * i] * lf[0] + T[3 * i + 1] * lf[1] + T[3 * i + 2] * lf[2];
        dv[i] = (f + vg[i] * vimp) * w;
      }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) d.e_dv[at(e, 24, 3 * c + i)] = dv[i];
  }
#pragma unroll
  for (uint32_t r = 0; r < R; ++r) d.e_coupling[at(e, R, r)] = coupling[r];
  const double vrate = d.e_vol[at(e, 3, 0)], vsoft = d.e_vol[at(e, 3, 1)], vcap0 = d.e_vol[at(e, 3, 2)];
  const double vcap = inverted ? vcap0 : (vcap0 < 1.0 ? 1.0 : vcap0);
  const double volume = clampd(det - 1.0, -vcap, vcap);
  const double gain = vweight * vsoft;
  d.e_scal[at(e, 3, 0)] = gain;
  d.e_scal[at(e, 3, 1)] = volume * vrate;
  d.e_scal[at(e, 3, 2)] = vweight + gain;  // the first pass turns it into 1 / schur
}

// applyElement's warm start for one colour's elements of the bodies whose
// exact flag equals `exact`: the corner velocity changes k_lat_frame made,
// one lane per corner, eight elements per wave.
__global__ void __launch_bounds__(EBLOCK) k_lat_warm(LatDev d, uint32_t first, uint32_t count, uint32_t substep,
                                                     uint32_t exact) {
  const uint32_t slot = threadIdx.x / 8, c = threadIdx.x % 8;
  const uint32_t t = blockIdx.x * (EBLOCK / 8) + slot;
  if (t >= count) return;
  const uint32_t e = first + t;
  const uint32_t sub = d.e_sub[e];
  if (substep >= (sub & 0x7fffffffu) || (sub >> 31) != exact) return;
  if (d.e_inv_mass[at(e, 8, c)] == 0.0) return;
  double *vp = d.v + 3 * size_t(d.e_nodes[at(e, 8, c)]);
#pragma unroll
  for (int i = 0; i < 3; ++i) vp[i] -= d.e_dv[at(e, 24, 3 * c + i)];
}

// applyElement's warm start of every exact body, node by node: each node
// subtracts its elements' corner changes in slot order, which is colour order,
// so the node sees the same subtractions in the same order as the colour by
// colour application.
__global__ void __launch_bounds__(NODE_BLOCK) k_lat_warm_gather(LatDev d, uint32_t substep) {
  const uint32_t i = blockIdx.x * NODE_BLOCK + threadIdx.x;
  if (i >= d.N) return;
  const uint32_t b = d.node_body[i];
  if (substep >= d.body_substeps[b] || !d.body_exact[b]) return;
  const uint32_t k0 = d.warm_start[i], k1 = d.warm_start[i + 1];
  if (k0 == k1) return;
  double *vp = d.v + 3 * size_t(i);
  double v0 = vp[0], v1 = vp[1], v2 = vp[2];
  for (uint32_t k = k0; k < k1; ++k) {
    const double *dv = d.e_dv + 3 * size_t(d.warm_entry[k]);  // entry = slot * 8 + corner
    v0 -= dv[0];
    v1 -= dv[1];
    v2 -= dv[2];
  }
  vp[0] = v0;
  vp[1] = v1;
  vp[2] = v2;
}

// returnMap (plasticflow.cc:110-155): radial return of the Hencky strain of
// the elastic deformation E to `yield`
__device__ bool return_map(const double *E, double shear, double yield, double *mapped, double *trial,
                           double *strain) {
  double left[9], w[3], V[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) left[3 * i + j] = E[3 * i] * E[3 * j] + E[3 * i + 1] * E[3 * j + 1] + E[3 * i + 2] * E[3 * j + 2];
  eigen_sym3(left, w, V);
  if (!(w[0] > 0.0)) return false;  // ascending: w[0] is the smallest
  double e[3], dev[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) e[k] = 0.5 * log(w[k]);
  const double mean = (e[0] + e[1] + e[2]) / 3.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) dev[k] = e[k] - mean;
  const double size = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  const double stress = sqrt(1.5) * 2.0 * shear * size;
  if (!(stress > yield) || size <= 1.0e-15) return false;
  double scale[3], corr[9];
#pragma unroll
  for (int k = 0; k < 3; ++k) scale[k] = exp(dev[k] * (yield / stress - 1.0));
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      corr[3 * i + j] = V[3 * i] * scale[0] * V[3 * j] + V[3 * i + 1] * scale[1] * V[3 * j + 1] +
                        V[3 * i + 2] * scale[2] * V[3 * j + 2];
  mul3(corr, E, mapped);
  *trial = stress;
  *strain = (stress - yield) / (3.0 * shear);
  return true;
}


// one Jacobi rotation of a symmetric 3x3 in the (p, q) plane; r is the third
// index (the rotation the host's sym_eigen3 applies, values only)
__device__ __forceinline__ void jacobi_rotate(double &app, double &aqq, double &apq, double &arp, double &arq) {
  if (apq == 0.0) return;
  const double theta = (aqq - app) / (2.0 * apq);
  const double t = fabs(theta) > 1.0e150 ? 0.5 / theta
                                         : (theta < 0.0 ? -1.0 : 1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
  const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
  app -= t * apq;
  aqq += t * apq;
  apq = 0.0;
  const double rp = arp, rq = arq;
  arp = c * rp - s * rq;
  arq = s * rp + c * rq;
}

__device__ double largest_eigenvalue(double a00, double a11, double a22, double a01, double a02, double a12) {
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = a01 * a01 + a02 * a02 + a12 * a12;
    const double diag = a00 * a00 + a11 * a11 + a22 * a22;
    if (!(off > 1.0e-40 * diag)) break;
    jacobi_rotate(a00, a11, a01, a02, a12);
    jacobi_rotate(a00, a22, a02, a01, a12);
    jacobi_rotate(a11, a22, a12, a01, a02);
  }
  return fmax(a00, fmax(a11, a22));
}

// overstretched (tearing.cc:488-503) after a step: one lane per cell; over
// its eight corner frames (the corner's three edges, now against rest) the
// largest principal stretch sqrt(lambda_max(F^T F)), flagged past the limit
__global__ void __launch_bounds__(NODE_BLOCK) k_lat_stretch(LatDev d) {
  const uint32_t c = blockIdx.x * NODE_BLOCK + threadIdx.x;
  if (c >= d.T) return;
  const double limit = d.t_limit[c];
  uint32_t torn = 0u;
  if (isfinite(limit)) {
    double px[8][3], pr[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double *xp = d.x + 3 * size_t(d.t_nodes[8 * size_t(c) + k]);
      const double *rp = d.t_rest + 24 * size_t(c) + 3 * k;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        px[k][i] = xp[i];
        pr[k][i] = rp[i];
      }
    }
    double best = 0.0;
#pragma unroll
    for (int corner = 0; corner < 8; ++corner) {
      double R[9], N[9];
#pragma unroll
      for (int e = 0; e < 3; ++e)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          R[3 * i + e] = pr[corner ^ (1 << e)][i] - pr[corner][i];
          N[3 * i + e] = px[corner ^ (1 << e)][i] - px[corner][i];
        }
      const double c0 = R[4] * R[8] - R[5] * R[7], c1 = R[5] * R[6] - R[3] * R[8], c2 = R[3] * R[7] - R[4] * R[6];
      const double det = R[0] * c0 + R[1] * c1 + R[2] * c2;
      if (!(fabs(det) > 0.0)) continue;
      const double id = 1.0 / det;
      const double Ri[9] = {c0 * id, (R[2] * R[7] - R[1] * R[8]) * id, (R[1] * R[5] - R[2] * R[4]) * id,
                            c1 * id, (R[0] * R[8] - R[2] * R[6]) * id, (R[2] * R[3] - R[0] * R[5]) * id,
                            c2 * id, (R[1] * R[6] - R[0] * R[7]) * id, (R[0] * R[4] - R[1] * R[3]) * id};
      double F[9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) F[3 * i + j] = N[3 * i] * Ri[j] + N[3 * i + 1] * Ri[3 + j] + N[3 * i + 2] * Ri[6 + j];
      double C[9];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) C[3 * i + j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
      const double lam = largest_eigenvalue(C[0], C[4], C[8], C[1], C[2], C[5]);
      best = fmax(best, sqrt(fmax(lam, 0.0)));
    }
    torn = best > limit ? 1u : 0u;
  }
  d.t_torn[c] = torn;
}

// flowLattice (plasticflow.cc:161-260) after a step: one lane per Gauss point
// of every yielding cell (point p at (signs of corner p) / sqrt(3)); a point
// whose P changes by more than 1e-12 (one-norm) flags its cell as moved
__global__ void __launch_bounds__(NODE_BLOCK) k_lat_flow(LatDev d) {
  const uint32_t t = blockIdx.x * NODE_BLOCK + threadIdx.x;
  if (t >= 8u * d.C) return;
  const uint32_t c = t >> 3, p = t & 7u;
  const double *pm = d.c_pmat + size_t(c) * 8;
  const double yield = pm[3];
  if (!isfinite(yield)) return;
  const double shear = pm[0], ratio = pm[2], hmod = pm[4], cap = pm[5], crush = pm[6], limit = pm[7];
  const bool crushable = isfinite(crush) && limit < 1.0;
  const double at[3] = {0.5773502691896258 * corner_sign(p, 0), 0.5773502691896258 * corner_sign(p, 1),
                        0.5773502691896258 * corner_sign(p, 2)};
  const double h0 = d.c_half[3 * size_t(c)], h1 = d.c_half[3 * size_t(c) + 1], h2 = d.c_half[3 * size_t(c) + 2];
  double F[9] = {};
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) {
    const double s0 = corner_sign(k, 0), s1 = corner_sign(k, 1), s2 = corner_sign(k, 2);
    const double g[3] = {0.125 * s0 * (1.0 + s1 * at[1]) * (1.0 + s2 * at[2]) / h0,
                         0.125 * s1 * (1.0 + s0 * at[0]) * (1.0 + s2 * at[2]) / h1,
                         0.125 * s2 * (1.0 + s0 * at[0]) * (1.0 + s1 * at[1]) / h2};
    const double *xp = d.x + 3 * size_t(d.c_nodes[8 * size_t(c) + k]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) F[3 * i + j] += xp[i] * g[j];
  }
  if (!(det3(F) > 0.0)) return;
  double *Pp = d.c_plastic + 72 * size_t(c) + 9 * p;
  double P[9], E[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) P[k] = Pp[k];
  mul3(F, P, E);
  const double hard = d.c_hard[8 * size_t(c) + p];
  const double start = yield + hmod * hard;
  const double room = crushable ? -log(det3(P) * limit) : 0.0;
  const bool squeezed = crushable && room > 0.0 && det3(E) < exp(-crush);
  double tE[9], tt = 0.0, ts = 0.0;
  const bool trial = return_map(E, shear, start, tE, &tt, &ts);
  const double now = trial ? (start + ratio * tt) / (1.0 + ratio) : start;
  double mE[9], mt = 0.0, ms = 0.0;
  const bool mapped = trial && return_map(E, shear, now, mE, &mt, &ms);
  double after[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) after[k] = mapped ? mE[k] : E[k];
  const double volume = det3(after);
  double compaction = 0.0;
  if (squeezed && volume > 0.0) compaction = fmin(-crush - log(volume), room);
  if (!mapped && !(compaction > 0.0)) return;
  const double grow = compaction > 0.0 ? exp(compaction / 3.0) : 1.0;
  double Fi[9], ga[9], next[9];
  inv3(F, Fi);
#pragma unroll
  for (int k = 0; k < 9; ++k) ga[k] = grow * after[k];
  mul3(Fi, ga, next);
  const double kept = det3(next);
  if (isfinite(cap) && kept > 0.0) {
    double st[9];
    inv3(next, st);
    st[0] -= 1.0;
    st[4] -= 1.0;
    st[8] -= 1.0;
    double size = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) size += st[k] * st[k];
    size = sqrt(size);
    if (size > cap) {
      double capped[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) capped[k] = (cap / size) * st[k];
      capped[0] += 1.0;
      capped[4] += 1.0;
      capped[8] += 1.0;
      const double cv = det3(capped);
      if (!(cv > 1.0e-6)) return;
      double ci[9];
      inv3(capped, ci);
      const double sc = cbrt(kept * cv);
#pragma unroll
      for (int k = 0; k < 9; ++k) next[k] = sc * ci[k];
    }
  }
  double change = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k) change += fabs(next[k] - P[k]);
  if (!(change > 1.0e-12)) return;
#pragma unroll
  for (int k = 0; k < 9; ++k) Pp[k] = next[k];
  d.c_hard[8 * size_t(c) + p] = hard + (mapped ? ms : 0.0);
  d.c_moved[c] = 1u;  // every writer stores the same value
}

// solveElement for one colour's elements: one block solve of the 18 rows with
// the volume row folded in through its Schur complement, applied to the
// corner velocities.  One lane per row, 18 lanes per element, three
// elements per wave; lanes 0..7 of an element also own its corners.  The
// first pass also finishes updateElement: image = Inv coupling and 1 / schur.
__global__ void __launch_bounds__(EBLOCK) k_lat_solve(LatDev d, uint32_t first, uint32_t count,
                                                      uint32_t substep, uint32_t first_pass) {
  __shared__ double s_b[EPW + 1][24];    // per corner T^T v_c
  __shared__ double s_vs[EPW + 1][8];    // per corner vg_c . v_c
  __shared__ double s_res[EPW + 1][R];   // residual rows, then the final solution
  __shared__ double s_prod[EPW + 1][R];  // coupling_r * (Inv residual)_r
  __shared__ double s_inv[EPW + 1][PACKED];  // the element's packed inverse
  __shared__ double s_cp[EPW + 1][R];    // coupling rows (first pass)
  __shared__ double s_pi[EPW + 1][R];    // coupling_r * image_r (first pass)
  const uint32_t slot = threadIdx.x / R, r = threadIdx.x - slot * R;
  const uint32_t t = blockIdx.x * EPW + slot;
  const uint32_t e = first + t;
  // loads are issued for every element in range; the work and the stores
  // only for elements whose body takes this substep
  const bool inrange = slot < EPW && t < count;
  const bool active = inrange && substep < (d.e_sub[e] & 0x7fffffffu);
  const bool corner = r < 8;
  const uint32_t c = r;
  double inv[9], T[9];
  double bend = 0.0, bias = 0.0, soft = 0.0, imp = 0.0, coupling = 0.0, image = 0.0;
  double gain = 0.0, rhs = 0.0, schur = 0.0, vimp = 0.0, w = 0.0;
  double vg[3] = {0.0, 0.0, 0.0}, vc[3] = {0.0, 0.0, 0.0};
  uint32_t node = 0;
  if (inrange) {
    // every global load first (the corner gather after its node index), then
    // the LDS writes, so the element waits for memory once
    if (corner) node = d.e_nodes[at(e, 8, c)];
    double tri[STAGE];
    load_inverse(d.inv_table + size_t(d.e_inv_id[e]) * PACKED, r, tri);
#pragma unroll
    for (int k = 0; k < 9; ++k) inv[k] = d.e_shape_inv[at(e, 9, k)];
    bend = d.e_bend[e];
    gain = d.e_scal[at(e, 3, 0)];
    rhs = d.e_scal[at(e, 3, 1)];
    schur = d.e_scal[at(e, 3, 2)];
    vimp = d.e_vimp[e];
    bias = d.e_bias[at(e, R, r)];
    soft = d.e_soft[at(e, R, r)];
    imp = d.e_impulse[at(e, R, r)];
    coupling = d.e_coupling[at(e, R, r)];
    if (!first_pass) image = d.e_image[at(e, R, r)];
    if (corner) {
#pragma unroll
      for (int k = 0; k < 9; ++k) T[k] = d.e_turn[at(e, 9, k)];
      w = d.e_inv_mass[at(e, 8, c)];
#pragma unroll
      for (int k = 0; k < 3; ++k) vg[k] = d.e_vgrad[at(e, 24, 3 * c + k)];
      const double *vp = d.v + 3 * size_t(node);
      vc[0] = vp[0];
      vc[1] = vp[1];
      vc[2] = vp[2];
    }
    store_inverse(tri, r, s_inv[slot]);
    s_cp[slot][r] = coupling;
    if (corner) {
      s_vs[slot][c] = vg[0] * vc[0] + vg[1] * vc[1] + vg[2] * vc[2];
#pragma unroll
      for (int k = 0; k < 3; ++k) s_b[slot][3 * c + k] = T[k] * vc[0] + T[3 + k] * vc[1] + T[6 + k] * vc[2];
    }
  }
  __syncthreads();
  if (active) {
    double acc = 0.0;
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
      double gk[3];
      grad_of(inv, k, gk);
      acc += row_term(r, k, gk, bend, &s_b[slot][3 * k]);
    }
    s_res[slot][r] = acc + (bias - soft * imp);
    if (first_pass) {
      // the rest of updateElement: image = Inv coupling and the Schur
      // complement of the volume row
      image = inverse_row(s_inv[slot], r, s_cp[slot]);
      s_pi[slot][r] = coupling * image;
    }
  }
  __syncthreads();
  double sol = 0.0;
  if (active) {
    sol = inverse_row(s_inv[slot], r, s_res[slot]);
    s_prod[slot][r] = coupling * sol;
    if (first_pass) {
      double coupled = 0.0;
#pragma unroll
      for (uint32_t k = 0; k < R; ++k) coupled += s_pi[slot][k];
      const double left = schur - coupled;  // schur holds vweight + gain here
      schur = d.e_volumetric[e] && left > 0.0 ? 1.0 / left : 0.0;
      d.e_image[at(e, R, r)] = image;
      if (r == 0) d.e_scal[at(e, 3, 2)] = schur;
    }
  }
  __syncthreads();
  double change = 0.0;
  if (active) {
    double coupled = 0.0, vspeed = 0.0;
#pragma unroll
    for (uint32_t k = 0; k < R; ++k) coupled += s_prod[slot][k];
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) vspeed += s_vs[slot][k];
    change = schur * ((rhs - gain * vimp + vspeed) - coupled);
    sol -= image * change;
    d.e_impulse[at(e, R, r)] = imp + sol;
    s_res[slot][r] = sol;
    if (r == 0) d.e_vimp[e] = vimp + change;
  }
  __syncthreads();
  if (active && corner && w != 0.0) {
    double gc[3], lf[3];
    grad_of(inv, c, gc);
    row_force(s_res[slot], gc, bend, c, lf);
    double *vp = d.v + 3 * size_t(node);  // only this element of its colour touches the node
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double f = T[3 * i] * lf[0] + T[3 * i + 1] * lf[1] + T[3 * i + 2] * lf[2];
      vp[i] = vc[i] - (f + vg[i] * change) * w;
    }
  }
}

}  // namespace

struct LatticeSet::Device {
  std::vector<void *> owned;
  LatDev d{};
  uint32_t colour_start[9] = {};
  std::vector<uint32_t> node0;               // first node of each body
  std::vector<std::vector<uint32_t>> slot_of;  // per body, per element: slot
  std::vector<uint32_t> body_substeps;         // per body, as laid out
  std::vector<uint8_t> body_exact;
  std::vector<uint32_t> active;                // [substep][colour]: elements taking that substep
  std::vector<uint32_t> active_total;          // [substep]: elements taking it, all colours
  uint32_t *frame_order = nullptr;             // slots by descending substeps (device)
  std::vector<uint32_t> cell0;                 // first plasticity cell of each body
  bool yields = false;                         // some body has a yielding lane
  bool tears = false;                          // some body has a tearing lane
  std::vector<uint32_t> slot_body;             // per slot: its body
  uint32_t inverses = 0;
  hipEvent_t start = nullptr, stop = nullptr;
  std::vector<hipEvent_t> marks;  // per-launch timing events (KERNEL_TIMING)
  // the captured launch sequence of one step, valid for these arrays and
  // this (substeps, exact flags) combination
  hipGraphExec_t graph = nullptr;
  uint32_t graph_substeps = 0;
  bool graph_exact = false, graph_inexact = false;

  template <class T>
  T *alloc(size_t n) {
    T *p = nullptr;
    HIPCHECK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    owned.push_back(p);
    return p;
  }
  template <class T>
  T *copy(const std::vector<T> &h) {
    T *p = alloc<T>(h.size());
    if (!h.empty()) HIPCHECK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
  }
  ~Device() {
    if (graph) (void)hipGraphExecDestroy(graph);
    for (void *p : owned) (void)hipFree(p);
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
    for (hipEvent_t e : marks) (void)hipEventDestroy(e);
  }
};

LatticeSet::LatticeSet(hipStream_t stream) : stream_(stream) {}

LatticeSet::~LatticeSet() {
  if (stream_) (void)hipStreamSynchronize(stream_);
}

uint32_t LatticeSet::add(LatticeBody &&body) {
  download();
  bodies_.push_back(std::move(body));
  device_valid_ = false;
  return uint32_t(bodies_.size() - 1);
}

LatticeBody &LatticeSet::body(uint32_t i) {
  download();
  return bodies_[i];
}

void LatticeSet::set_substeps(uint32_t i, uint32_t substeps) { bodies_[i].substeps = substeps; }

void LatticeSet::get_plastic(uint32_t i, double *plastic, double *hardening, uint32_t *present) {
  download();
  const LatticeBody &b = bodies_[i];
  if (present) *present = b.plastic.empty() ? 0u : 1u;
  if (b.plastic.empty()) return;
  if (plastic) std::memcpy(plastic, b.plastic.data(), b.plastic.size() * sizeof(double));
  if (hardening) std::memcpy(hardening, b.hardening.data(), b.hardening.size() * sizeof(double));
}

void LatticeSet::set_plastic(uint32_t i, const double *plastic, const double *hardening) {
  download();
  LatticeBody &b = bodies_[i];
  b.ensure_plastic();
  if (plastic) std::memcpy(b.plastic.data(), plastic, b.plastic.size() * sizeof(double));
  if (hardening) std::memcpy(b.hardening.data(), hardening, b.hardening.size() * sizeof(double));
  device_valid_ = false;
  flowed_ = false;
  refresh_pending_ = true;  // the next step rebuilds the elements of the cells whose P moved
}

void LatticeSet::get_state(uint32_t i, double *x, double *v) {
  LatticeBody &b = bodies_[i];
  const size_t n = 3 * size_t(b.nodes);
  if (device_valid_) {
    const size_t off = 3 * size_t(dev_->node0[i]);
    if (x) HIPCHECK(hipMemcpyAsync(b.x.data(), dev_->d.x + off, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (v) HIPCHECK(hipMemcpyAsync(b.v.data(), dev_->d.v + off, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
  }
  if (x) std::memcpy(x, b.x.data(), n * sizeof(double));
  if (v) std::memcpy(v, b.v.data(), n * sizeof(double));
}

void LatticeSet::set_state(uint32_t i, const double *x, const double *v) {
  LatticeBody &b = bodies_[i];
  const size_t n = 3 * size_t(b.nodes);
  if (x) std::memcpy(b.x.data(), x, n * sizeof(double));
  if (v) std::memcpy(b.v.data(), v, n * sizeof(double));
  if (device_valid_) {
    const size_t off = 3 * size_t(dev_->node0[i]);
    if (x) HIPCHECK(hipMemcpyAsync(dev_->d.x + off, b.x.data(), n * sizeof(double), hipMemcpyHostToDevice, stream_));
    if (v) HIPCHECK(hipMemcpyAsync(dev_->d.v + off, b.v.data(), n * sizeof(double), hipMemcpyHostToDevice, stream_));
    HIPCHECK(hipStreamSynchronize(stream_));
  }
}

void LatticeSet::download() {
  if (!device_valid_ || !dev_) return;
  const LatDev &d = dev_->d;
  const size_t E = d.E;
  std::vector<double> x(3 * size_t(d.N)), v(3 * size_t(d.N)), imp(size_t(R) * E), vimp(E), quat(4 * E);
  HIPCHECK(hipStreamSynchronize(stream_));
  HIPCHECK(hipMemcpy(x.data(), d.x, x.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(v.data(), d.v, v.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (E) {
    HIPCHECK(hipMemcpy(imp.data(), d.e_impulse, imp.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(vimp.data(), d.e_vimp, vimp.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(quat.data(), d.e_quat, quat.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  for (size_t b = 0; b < bodies_.size(); ++b) {
    LatticeBody &body = bodies_[b];
    const size_t off = 3 * size_t(dev_->node0[b]);
    std::copy(x.begin() + off, x.begin() + off + 3 * size_t(body.nodes), body.x.begin());
    std::copy(v.begin() + off, v.begin() + off + 3 * size_t(body.nodes), body.v.begin());
    for (size_t i = 0; i < body.elements.size(); ++i) {
      const size_t s = dev_->slot_of[b][i];
      HexElement &el = body.elements[i];
      for (uint32_t r = 0; r < R; ++r) el.impulse[r] = imp[s * R + r];
      el.vol_impulse = vimp[s];
      for (int k = 0; k < 4; ++k) el.quat[k] = quat[s * 4 + k];
    }
  }
  if (dev_->yields) {
    std::vector<double> P(72 * size_t(d.C)), a(8 * size_t(d.C));
    HIPCHECK(hipMemcpy(P.data(), d.c_plastic, P.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(a.data(), d.c_hard, a.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t b = 0; b < bodies_.size(); ++b) {
      LatticeBody &body = bodies_[b];
      if (body.plastic.empty()) continue;
      const size_t c0 = dev_->cell0[b];
      std::copy(P.begin() + 72 * c0, P.begin() + 72 * (c0 + body.hexes), body.plastic.begin());
      std::copy(a.begin() + 8 * c0, a.begin() + 8 * (c0 + body.hexes), body.hardening.begin());
    }
  }
}

void LatticeSet::upload(const std::vector<uint32_t> &substeps, const std::vector<uint8_t> &exact) {
  dev_.reset();
  auto dv = std::make_unique<Device>();
  const uint32_t B = uint32_t(bodies_.size());
  uint32_t N = 0, E = 0;
  dv->node0.resize(B);
  dv->slot_of.resize(B);
  for (uint32_t b = 0; b < B; ++b) {
    dv->node0[b] = N;
    N += bodies_[b].nodes;
    E += uint32_t(bodies_[b].elements.size());
    dv->slot_of[b].assign(bodies_[b].elements.size(), 0);
  }
  // slots: colour, then the bodies by descending substep count (ties in
  // body order), then cell, so the elements that take substep s form a
  // prefix of every colour's range
  std::vector<uint32_t> by_substeps(B);
  for (uint32_t b = 0; b < B; ++b) by_substeps[b] = b;
  std::stable_sort(by_substeps.begin(), by_substeps.end(),
                   [&](uint32_t a, uint32_t b) { return substeps[a] > substeps[b]; });
  const uint32_t smax = B ? substeps[by_substeps[0]] : 0u;
  std::vector<std::pair<uint32_t, uint32_t>> order;  // (body, element)
  order.reserve(E);
  dv->active.assign(8 * size_t(smax), 0u);
  for (uint32_t col = 0; col < 8; ++col) {
    dv->colour_start[col] = uint32_t(order.size());
    for (uint32_t b : by_substeps) {
      const LatticeBody &body = bodies_[b];
      for (uint32_t i = 0; i < body.elements.size(); ++i)
        if (body.hex_colour[body.elements[i].cell] == col) {
          dv->slot_of[b][i] = uint32_t(order.size());
          order.emplace_back(b, i);
          for (uint32_t s = 0; s < substeps[b]; ++s) ++dv->active[8 * size_t(s) + col];
        }
    }
  }
  dv->colour_start[8] = uint32_t(order.size());
  dv->active_total.assign(smax, 0u);
  for (uint32_t s = 0; s < smax; ++s)
    for (uint32_t col = 0; col < 8; ++col) dv->active_total[s] += dv->active[8 * size_t(s) + col];
  std::vector<uint32_t> frame_order(order.size());
  for (size_t s = 0; s < frame_order.size(); ++s) frame_order[s] = uint32_t(s);
  std::stable_sort(frame_order.begin(), frame_order.end(), [&](uint32_t a, uint32_t b) {
    return substeps[order[a].first] > substeps[order[b].first];
  });
  // nodes
  std::vector<double> x(3 * size_t(N)), v(3 * size_t(N));
  std::vector<uint32_t> node_body(N);
  std::vector<uint8_t> node_pinned(N);
  std::vector<double> body_h(B);
  std::vector<uint32_t> body_substeps(B);
  std::vector<uint8_t> body_exact(B);
  for (uint32_t b = 0; b < B; ++b) {
    const LatticeBody &body = bodies_[b];
    const size_t n0 = dv->node0[b];
    std::copy(body.x.begin(), body.x.end(), x.begin() + 3 * n0);
    std::copy(body.v.begin(), body.v.end(), v.begin() + 3 * n0);
    for (uint32_t i = 0; i < body.nodes; ++i) {
      node_body[n0 + i] = b;
      node_pinned[n0 + i] = body.pinned(i) ? 1 : 0;
    }
    body_h[b] = body.h;
  }
  // elements
  const size_t Es = E;
  std::vector<uint32_t> e_body(Es), e_nodes(8 * Es), e_inv_id(Es);
  std::vector<double> e_shape_inv(9 * Es), e_inv_mass(8 * Es), e_bend(Es), e_rest_hg(12 * Es),
      e_rate(R * Es), e_soft(R * Es), e_cap(R * Es), e_vol(3 * Es), e_impulse(R * Es), e_vimp(Es),
      e_quat(4 * Es);
  std::vector<uint8_t> e_volumetric(Es);
  std::unordered_map<std::string, uint32_t> interned;  // inverse bytes -> table entry
  std::vector<double> table;
  for (size_t s = 0; s < Es; ++s) {
    const uint32_t b = order[s].first;
    const HexElement &el = bodies_[b].elements[order[s].second];
    const uint32_t n0 = dv->node0[b];
    e_body[s] = b;
    for (int c = 0; c < 8; ++c) {
      e_nodes[s * 8 + c] = n0 + el.nodes[c];
      e_inv_mass[s * 8 + c] = el.inv_mass[c];
    }
    for (int k = 0; k < 9; ++k) e_shape_inv[s * 9 + k] = el.shape_inverse[k];
    e_bend[s] = el.bend;
    for (int k = 0; k < 12; ++k) e_rest_hg[s * 12 + k] = el.rest_hg[k];
    for (uint32_t r = 0; r < R; ++r) {
      e_rate[s * R + r] = el.rate[r];
      e_soft[s * R + r] = el.softness[r];
      e_cap[s * R + r] = el.cap[r];
      e_impulse[s * R + r] = el.impulse[r];
    }
    e_vol[s * 3] = el.vol_rate;
    e_vol[s * 3 + 1] = el.vol_softness;
    e_vol[s * 3 + 2] = el.vol_cap;
    e_volumetric[s] = el.volumetric ? 1 : 0;
    e_vimp[s] = el.vol_impulse;
    for (int k = 0; k < 4; ++k) e_quat[s * 4 + k] = el.quat[k];
    std::string key(reinterpret_cast<const char *>(el.inverse.data()), HEX_INVERSE * sizeof(double));
    auto it = interned.find(key);
    if (it == interned.end()) {
      it = interned.emplace(std::move(key), uint32_t(interned.size())).first;
      for (uint32_t i = 0; i < R; ++i)  // upper triangle, row by row
        for (uint32_t j = i; j < R; ++j) table.push_back(el.inverse[R * i + j]);
    }
    e_inv_id[s] = it->second;
  }
  dv->inverses = uint32_t(interned.size());
  // warm-start entries per node, in slot order
  std::vector<uint32_t> warm_start(size_t(N) + 1, 0), warm_entry;
  for (size_t s = 0; s < Es; ++s)
    for (uint32_t c = 0; c < 8; ++c)
      if (e_inv_mass[s * 8 + c] != 0.0) ++warm_start[size_t(e_nodes[s * 8 + c]) + 1];
  for (size_t n = 0; n < N; ++n) warm_start[n + 1] += warm_start[n];
  warm_entry.resize(warm_start[N]);
  {
    std::vector<uint32_t> fill(warm_start.begin(), warm_start.end() - 1);
    for (size_t s = 0; s < Es; ++s)
      for (uint32_t c = 0; c < 8; ++c)
        if (e_inv_mass[s * 8 + c] != 0.0) warm_entry[fill[e_nodes[s * 8 + c]]++] = uint32_t(s * 8 + c);
  }
  LatDev &d = dv->d;
  d.N = N;
  d.E = E;
  d.x = dv->copy(x);
  d.v = dv->copy(v);
  d.node_body = dv->copy(node_body);
  d.node_pinned = dv->copy(node_pinned);
  d.body_h = dv->copy(body_h);
  d.body_substeps = dv->alloc<uint32_t>(B);
  d.body_exact = dv->alloc<uint8_t>(B);
  d.e_body = dv->copy(e_body);
  d.e_sub = dv->alloc<uint32_t>(Es);
  dv->slot_body = e_body;
  d.e_nodes = dv->copy(e_nodes);
  d.e_inv_id = dv->copy(e_inv_id);
  d.e_shape_inv = dv->copy(e_shape_inv);
  d.e_inv_mass = dv->copy(e_inv_mass);
  d.e_bend = dv->copy(e_bend);
  d.e_rest_hg = dv->copy(e_rest_hg);
  d.e_rate = dv->copy(e_rate);
  d.e_soft = dv->copy(e_soft);
  d.e_cap = dv->copy(e_cap);
  d.e_vol = dv->copy(e_vol);
  d.e_volumetric = dv->copy(e_volumetric);
  d.inv_table = dv->copy(table);
  d.e_impulse = dv->copy(e_impulse);
  d.e_vimp = dv->copy(e_vimp);
  d.e_quat = dv->copy(e_quat);
  d.e_turn = dv->alloc<double>(9 * Es);
  d.e_bias = dv->alloc<double>(R * Es);
  d.e_vgrad = dv->alloc<double>(24 * Es);
  d.e_coupling = dv->alloc<double>(R * Es);
  d.e_image = dv->alloc<double>(R * Es);
  d.e_scal = dv->alloc<double>(3 * Es);
  d.e_dv = dv->alloc<double>(24 * Es);
  dv->frame_order = dv->copy(frame_order);
  d.warm_start = dv->copy(warm_start);
  d.warm_entry = dv->copy(warm_entry);
  // plasticity cells: every cell of every body, bodies in order
  {
    uint32_t C = 0;
    dv->cell0.resize(B);
    for (uint32_t b = 0; b < B; ++b) {
      dv->cell0[b] = C;
      C += bodies_[b].hexes;
      if (bodies_[b].yields) {
        bodies_[b].ensure_plastic();
        dv->yields = true;
      }
    }
    d.C = dv->yields ? C : 0u;
    if (dv->yields) {
      std::vector<uint32_t> c_nodes(8 * size_t(C));
      std::vector<double> c_half(3 * size_t(C)), c_pmat(8 * size_t(C)), c_plastic(72 * size_t(C)),
          c_hard(8 * size_t(C), 0.0);
      const double inf = std::numeric_limits<double>::infinity();
      for (uint32_t b = 0; b < B; ++b) {
        const LatticeBody &body = bodies_[b];
        const size_t c0 = dv->cell0[b];
        for (uint32_t c = 0; c < body.hexes; ++c) {
          const size_t g = c0 + c;
          for (int k = 0; k < 8; ++k) c_nodes[8 * g + k] = dv->node0[b] + body.hex[8 * size_t(c) + k];
          for (int k = 0; k < 3; ++k) c_half[3 * g + k] = double(0.5f * body.voxel[k]);
          const talus_mi_material &m = body.material_of(c);
          double shear, first;
          hex_material(m.youngs_modulus, m.poissons_ratio, &shear, &first);
          const double bulk = first + 2.0 * shear / 3.0;
          double *pm = &c_pmat[8 * g];
          pm[0] = shear;
          pm[1] = bulk;
          pm[2] = double(m.hardening_modulus) / (3.0 * shear);
          pm[3] = body.yields && std::isfinite(m.yield_stress) ? double(m.yield_stress) : inf;
          pm[4] = double(m.hardening_modulus);
          pm[5] = double(m.plastic_max);
          const bool crushable = std::isfinite(m.crush_stress) && m.crush_limit < 1.0f;
          pm[6] = crushable ? double(m.crush_stress) / bulk : inf;
          pm[7] = double(m.crush_limit);
          if (!body.plastic.empty()) {
            std::copy(body.plastic.begin() + 72 * size_t(c), body.plastic.begin() + 72 * size_t(c + 1),
                      c_plastic.begin() + 72 * g);
            std::copy(body.hardening.begin() + 8 * size_t(c), body.hardening.begin() + 8 * size_t(c + 1),
                      c_hard.begin() + 8 * g);
          } else {
            for (int q = 0; q < 8; ++q)
              c_plastic[72 * g + 9 * q] = c_plastic[72 * g + 9 * q + 4] = c_plastic[72 * g + 9 * q + 8] = 1.0;
          }
        }
      }
      d.c_nodes = dv->copy(c_nodes);
      d.c_half = dv->copy(c_half);
      d.c_pmat = dv->copy(c_pmat);
      d.c_plastic = dv->copy(c_plastic);
      d.c_hard = dv->copy(c_hard);
      d.c_moved = dv->copy(std::vector<uint32_t>(C, 0u));
    }
    // tearing cells: the same cells, when some lane tears
    for (uint32_t b = 0; b < B; ++b) dv->tears = dv->tears || bodies_[b].tears;
    d.T = dv->tears ? C : 0u;
    if (dv->tears) {
      std::vector<uint32_t> t_nodes(8 * size_t(C));
      std::vector<double> t_rest(24 * size_t(C)), t_limit(C);
      for (uint32_t b = 0; b < B; ++b) {
        const LatticeBody &body = bodies_[b];
        for (uint32_t c = 0; c < body.hexes; ++c) {
          const size_t g = dv->cell0[b] + c;
          for (int k = 0; k < 8; ++k) {
            const uint32_t n = body.hex[8 * size_t(c) + k];
            t_nodes[8 * g + k] = dv->node0[b] + n;
            for (int i = 0; i < 3; ++i) t_rest[24 * g + 3 * k + i] = body.rest[3 * size_t(n) + i];
          }
          t_limit[g] = body.tear_limit(c);
        }
      }
      d.t_nodes = dv->copy(t_nodes);
      d.t_rest = dv->copy(t_rest);
      d.t_limit = dv->copy(t_limit);
      d.t_torn = dv->alloc<uint32_t>(C);
    }
  }
  HIPCHECK(hipEventCreate(&dv->start));
  HIPCHECK(hipEventCreate(&dv->stop));
  // the bodies' substep counts and exact flags, per body and per slot
  HIPCHECK(hipMemcpy(const_cast<uint32_t *>(d.body_substeps), substeps.data(), B * sizeof(uint32_t),
                     hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(const_cast<uint8_t *>(d.body_exact), exact.data(), B, hipMemcpyHostToDevice));
  {
    std::vector<uint32_t> e_sub(Es);
    for (size_t s = 0; s < Es; ++s) {
      const uint32_t b = e_body[s];
      e_sub[s] = substeps[b] | (exact[b] ? 0x80000000u : 0u);
    }
    if (Es) HIPCHECK(hipMemcpy(const_cast<uint32_t *>(d.e_sub), e_sub.data(), Es * sizeof(uint32_t),
                               hipMemcpyHostToDevice));
  }
  dv->body_substeps = substeps;
  dv->body_exact = exact;
  dev_ = std::move(dv);
  device_valid_ = true;
}

void LatticeSet::step(const talus_mi_lattice_params &p, talus_mi_lattice_report &rep) {
  std::memset(&rep, 0, sizeof rep);
  const uint32_t B = uint32_t(bodies_.size());
  rep.lattices = B;
  if (B == 0) return;
  if (!(p.dt > 0.0) || !std::isfinite(p.dt)) throw std::invalid_argument("dt must be positive");
  std::vector<uint32_t> substeps(B);
  std::vector<double> h(B);
  std::vector<uint8_t> exact(B);
  bool rebuild = false;
  for (uint32_t b = 0; b < B; ++b) {
    substeps[b] = bodies_[b].island_substeps(p.dt, p.softstep_substeps);
    h[b] = p.dt / double(substeps[b]);
    exact[b] = !(h[b] * bodies_[b].stiffest() > 1.5) ? 1 : 0;
    if (bodies_[b].h != h[b] || bodies_[b].retopo) rebuild = true;
  }
  // refreshFlowed: the cells the last step's flow moved get their elements
  // rebuilt on the host (at the positions the step starts from)
  bool refresh = refresh_pending_;
  if (flowed_ && device_valid_ && dev_->d.C) {
    std::vector<uint32_t> moved(dev_->d.C);
    HIPCHECK(hipStreamSynchronize(stream_));
    HIPCHECK(hipMemcpy(moved.data(), dev_->d.c_moved, moved.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint32_t m : moved) refresh = refresh || m != 0u;
    if (refresh) download();
  }
  if (refresh) {
    for (uint32_t b = 0; b < B; ++b)  // a body whose substep changes or that tore is rebuilt whole below
      if (bodies_[b].h == h[b] && !bodies_[b].retopo) rep.refreshed += bodies_[b].refresh_flowed();
    device_valid_ = false;  // laid out again with the new elements and P
  }
  flowed_ = false;
  refresh_pending_ = false;

  if (rebuild) {
    download();
    for (uint32_t b = 0; b < B; ++b)
      if (bodies_[b].h != h[b]) {
        bodies_[b].build_elements(h[b]);  // fresh elements: impulses start at zero
        bodies_[b].retopo = false;
        ++rep.rebuilds;
      } else if (bodies_[b].retopo) {
        bodies_[b].rebuild_torn();  // the same substep: impulses carry over cell by cell
        ++rep.rebuilds;
      }
    device_valid_ = false;
  }
  // the slot order and the per-substep launch sizes follow the substep
  // counts: a change of them lays the device arrays out again
  if (device_valid_ && (dev_->body_substeps != substeps || dev_->body_exact != exact)) {
    download();
    device_valid_ = false;
  }
  if (!device_valid_) upload(substeps, exact);
  Device &dv = *dev_;
  LatDev &d = dv.d;
  uint32_t S = 0;
  bool any_exact = false, any_inexact = false;
  for (uint32_t b = 0; b < B; ++b) {
    any_exact = any_exact || exact[b];
    any_inexact = any_inexact || !exact[b];
    S = std::max(S, substeps[b]);
    rep.element_solves += uint64_t(substeps[b]) * PASSES * bodies_[b].elements.size();
  }
  rep.substeps = S;
  rep.elements = d.E;
  const uint32_t node_grid = (d.N + NODE_BLOCK - 1) / NODE_BLOCK;
  const bool timing = (p.flags & TALUS_MI_KERNEL_TIMING) != 0;
  const bool direct = timing || (p.flags & TALUS_MI_DIRECT_LAUNCH) != 0;
  // kernel timing: an event pair around each back-to-back run of one kind
  // (the frame and warm-start gather of a substep; its solve passes), so the
  // per-launch average is the kernels' own time and not the events'
  std::vector<std::pair<uint32_t, int>> timed;  // (first event, kind 0 update / 1 solve)
  auto open_span = [&]() -> uint32_t {
    const uint32_t k = uint32_t(2 * timed.size());
    while (dv.marks.size() < k + 2) {
      hipEvent_t ev;
      HIPCHECK(hipEventCreate(&ev));
      dv.marks.push_back(ev);
    }
    HIPCHECK(hipEventRecord(dv.marks[k], stream_));
    return k;
  };
  auto close_span = [&](uint32_t k, int kind) {
    HIPCHECK(hipEventRecord(dv.marks[k + 1], stream_));
    timed.emplace_back(k, kind);
  };
  auto enqueue = [&]() {
    // integrate of substep 0; each advance integrates the next substep
    hipLaunchKernelGGL(k_lat_integrate, dim3(node_grid), dim3(NODE_BLOCK), 0, stream_, d, 0u, p.gravity[0],
                       p.gravity[1], p.gravity[2]);
    for (uint32_t s = 0; s < S; ++s) {
      uint32_t span = 0;
      if (timing && any_exact) span = open_span();
      if (any_exact) {
        // every exact body's elements are updated before the passes: the
        // update reads positions only, so one launch covers every colour;
        // the warm starts are then gathered node by node
        const uint32_t live = dv.active_total[s];  // a prefix of the substep order
        if (live)
          hipLaunchKernelGGL(k_lat_frame, dim3((live + EBLOCK - 1) / EBLOCK), dim3(EBLOCK), 0, stream_, d,
                             dv.frame_order, 0u, live, s, 1u);
        hipLaunchKernelGGL(k_lat_warm_gather, dim3(node_grid), dim3(NODE_BLOCK), 0, stream_, d, s);
      }
      if (timing && any_exact) close_span(span, 0);
      // (a too-stiff body's per-colour update runs inside the solve span)
      if (timing) span = open_span();
      for (uint32_t pass = 0; pass < PASSES; ++pass)
        for (uint32_t col = 0; col < 8; ++col) {
          const uint32_t first = dv.colour_start[col], count = dv.colour_start[col + 1] - first;
          if (!count) continue;
          if (pass == 0 && any_inexact) {
            hipLaunchKernelGGL(k_lat_frame, dim3((count + EBLOCK - 1) / EBLOCK), dim3(EBLOCK), 0, stream_, d,
                               static_cast<const uint32_t *>(nullptr), first, count, s, 0u);
            hipLaunchKernelGGL(k_lat_warm, dim3((count + 7) / 8), dim3(EBLOCK), 0, stream_, d, first, count, s,
                               0u);
          }
          const uint32_t live = dv.active[8 * size_t(s) + col];  // a prefix of the colour's range
          if (live)
            hipLaunchKernelGGL(k_lat_solve, dim3((live + EPW - 1) / EPW), dim3(EBLOCK), 0, stream_, d, first,
                               live, s, pass == 0 ? 1u : 0u);
        }
      if (timing) close_span(span, 1);
      hipLaunchKernelGGL(k_lat_advance, dim3(node_grid), dim3(NODE_BLOCK), 0, stream_, d, s, p.gravity[0],
                         p.gravity[1], p.gravity[2]);
    }
  };
  for (uint32_t col = 0; col < 8; ++col)
    if (dv.colour_start[col + 1] > dv.colour_start[col]) {
      for (uint32_t s = 0; s < S; ++s) rep.solve_launches += dv.active[8 * size_t(s) + col] ? PASSES : 0;
      rep.update_launches += S * (any_inexact ? 2 : 0);  // a too-stiff body's frame + warm start per colour
    }
  if (any_exact && d.E) rep.update_launches += 2 * S;  // frame over every element, warm-start gather
  rep.inverses = dv.inverses;
  HIPCHECK(hipEventRecord(dv.start, stream_));
  if (direct) {
    enqueue();
    HIPCHECK(hipGetLastError());
  } else {
    if (!dv.graph || dv.graph_substeps != S || dv.graph_exact != any_exact || dv.graph_inexact != any_inexact) {
      if (dv.graph) {
        HIPCHECK(hipGraphExecDestroy(dv.graph));
        dv.graph = nullptr;
      }
      hipGraph_t g = nullptr;
      HIPCHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      enqueue();
      const hipError_t launched = hipGetLastError();
      const hipError_t ended = hipStreamEndCapture(stream_, &g);
      HIPCHECK(launched);
      HIPCHECK(ended);
      const hipError_t made = hipGraphInstantiate(&dv.graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIPCHECK(made);
      dv.graph_substeps = S;
      dv.graph_exact = any_exact;
      dv.graph_inexact = any_inexact;
    }
    HIPCHECK(hipGraphLaunch(dv.graph, stream_));
  }
  // flowLattice once per step, after the substeps
  if (d.C) {
    hipLaunchKernelGGL(k_lat_flow, dim3((8 * d.C + NODE_BLOCK - 1) / NODE_BLOCK), dim3(NODE_BLOCK), 0, stream_,
                       d);
    HIPCHECK(hipGetLastError());
    flowed_ = true;
  }
  // overstretched cells, flagged on the device at the step's end
  if (d.T) {
    hipLaunchKernelGGL(k_lat_stretch, dim3((d.T + NODE_BLOCK - 1) / NODE_BLOCK), dim3(NODE_BLOCK), 0, stream_, d);
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(hipEventRecord(dv.stop, stream_));
  HIPCHECK(hipEventSynchronize(dv.stop));
  float ms = 0.0f;
  HIPCHECK(hipEventElapsedTime(&ms, dv.start, dv.stop));
  rep.device_ms = ms;
  for (const auto &t : timed) {
    float k = 0.0f;
    HIPCHECK(hipEventElapsedTime(&k, dv.marks[t.first], dv.marks[t.first + 1]));
    (t.second ? rep.solve_ms : rep.update_ms) += k;
  }
  // tearCells on the host for the flagged cells (strainstep.cc:826-830);
  // the torn bodies are laid out again, and their elements rebuilt, at the
  // next step
  if (d.T) {
    std::vector<uint32_t> flags(d.T);
    HIPCHECK(hipMemcpy(flags.data(), d.t_torn, flags.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (std::find(flags.begin(), flags.end(), 1u) != flags.end()) {
      download();
      std::vector<uint32_t> cells;
      for (uint32_t b = 0; b < B; ++b) {
        LatticeBody &body = bodies_[b];
        cells.clear();
        for (uint32_t c = 0; c < body.hexes; ++c)
          if (flags[dv.cell0[b] + c]) cells.push_back(c);
        if (cells.empty()) continue;
        rep.torn += uint32_t(cells.size());
        const uint32_t made = body.tear_cells(cells.data(), uint32_t(cells.size()));
        rep.node_copies += made;
        if (made) device_valid_ = false;
      }
      // the flow flags are not read once the arrays are laid out again: the
      // bodies that did not tear compare P with their built P instead
      if (!device_valid_ && flowed_) refresh_pending_ = true;
    }
  }
}

uint32_t LatticeSet::tear(uint32_t i, const uint32_t *cells, uint32_t count) {
  download();
  const uint32_t made = bodies_.at(i).tear_cells(cells, count);
  if (made) {
    device_valid_ = false;
    if (flowed_) refresh_pending_ = true;  // as after a tear inside a step
  }
  return made;
}

void LatticeSet::cell_stretch(uint32_t i, uint32_t cell, double *stretch, double *direction) {
  download();
  bodies_.at(i).cell_stretch(cell, stretch, direction);
}

}  // namespace tmi
