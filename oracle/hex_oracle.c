/*
 * hex_oracle.c -- FP64 CPU restatement of the strainstep hex-lattice solve.
 *
 * TEST INFRASTRUCTURE ONLY (see hex_oracle.h).
 *
 * Reference (paths under src/talus/):
 *   dynamics/strainstep/element.cc:22-27    springOf
 *   dynamics/strainstep/element.cc:42-99    symmetricInverse (Gauss-Jordan,
 *                                           no pivoting, inert rows zeroed)
 *   dynamics/strainstep/element.cc:113-158  addRows / rowForce
 *   dynamics/strainstep/element.cc:162-242  corner signs, hourglass signs,
 *                                           elementMaterial, cellFrequency,
 *                                           cornerRest, rowJacobian
 *   dynamics/strainstep/element.cc:244-342  makeElement
 *   dynamics/strainstep/element.cc:344-372  rotationOf
 *   dynamics/strainstep/element.cc:374-459  updateElement
 *   dynamics/strainstep/element.cc:461-523  applyElement / solveElement
 *   dynamics/strainstep/strainstep.cc:56-126, 138-149, 551-567  pinned nodes,
 *                                           integrate/advance, substep rule,
 *                                           inverse masses
 *   dynamics/strainstep/strainstep.cc:569-842  the substep loop (elements
 *                                           only)
 *   dynamics/bodyset/bodyset.cc:13-86       makeLattice (positions, velocity,
 *                                           node colours, cell size)
 *   dynamics/strainstep/plasticflow.cc:62-72  authoredRest
 *   deformable/solver/solver.cc:2185-2316   buildMesh, hex elements
 *
 * The lattice turn (plasticflow.cc:342-389) seeds each element's rotation;
 * the reference takes U V^T of an Eigen JacobiSVD of the mass-weighted
 * moment.  Here the same proper rotation is found with the rotationOf
 * iteration run from the identity (it maximises tr(R^T M), as U V^T does when
 * det M > 0); it only seeds rotationOf, which then iterates to 1e-9.
 */
#include "hex_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ROWS ORC_HEX_ROWS

static const double MAX_CORRECTIVE_SPEED = 3.0; /* element.cc:11 */
static const uint32_t ROTATION_ITERATIONS = 16; /* element.cc:12 */
static const double ROTATION_TOLERANCE = 1.0e-9;
static const double RESOLVED = 1.5;             /* element.h:14 */
static const uint32_t MOST_SUBSTEPS = 64;       /* element.h:15 */
static const double MAX_SPEED = 400.0;          /* strainstep.cc:56 */
static const uint32_t PASSES = 4;               /* strainstep.cc:58 */

static void signs_of(uint32_t c, double *s) {
  s[0] = (c & 1u) ? 1.0 : -1.0;
  s[1] = (c & 2u) ? 1.0 : -1.0;
  s[2] = (c & 4u) ? 1.0 : -1.0;
}

double orc_hourglass_sign(uint32_t mode, uint32_t corner) {
  double s[3];
  signs_of(corner, s);
  switch (mode) {
    case 0: return s[0] * s[1];
    case 1: return s[1] * s[2];
    case 2: return s[2] * s[0];
    default: return s[0] * s[1] * s[2];
  }
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - b[1] * a[2];
  c[1] = a[2] * b[0] - b[2] * a[0];
  c[2] = a[0] * b[1] - b[0] * a[1];
}

/* y = M x, M row-major */
static void mv3(const double *M, const double *x, double *y) {
  for (int r = 0; r < 3; ++r) y[r] = M[3 * r] * x[0] + M[3 * r + 1] * x[1] + M[3 * r + 2] * x[2];
}

/* y = M^T x */
static void mtv3(const double *M, const double *x, double *y) {
  for (int r = 0; r < 3; ++r) y[r] = M[r] * x[0] + M[3 + r] * x[1] + M[6 + r] * x[2];
}

/* adjugate inverse of a row-major 3x3 (entries of adj(A) times 1/det) */
static void inverse3(const double *a, double *inv) {
  const double det = orc_det3(a);
  const double s = 1.0 / det;
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

void orc_hex_corner_rest(const double *corners, double *shape, double *hourglass) {
  memset(shape, 0, 9 * sizeof(double));
  memset(hourglass, 0, 12 * sizeof(double));
  for (uint32_t c = 0; c < 8; ++c) {
    double s[3];
    signs_of(c, s);
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) shape[3 * r + k] += corners[3 * c + r] * s[k] / 8.0;
    for (uint32_t m = 0; m < 4; ++m) {
      const double g = orc_hourglass_sign(m, c);
      for (int r = 0; r < 3; ++r) hourglass[3 * m + r] += g * corners[3 * c + r];
    }
  }
}

void orc_hex_material(float youngs, float poisson, double *shear, double *first) {
  double nu = (double)poisson;
  nu = nu < 0.0 ? 0.0 : (nu > 0.499 ? 0.499 : nu);
  const double e = (double)youngs > 0.0 ? (double)youngs : 0.0;
  *shear = e / (2.0 * (1.0 + nu));
  *first = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
}

double orc_hex_cell_frequency(float youngs, float poisson, float density, double cell) {
  double nu = (double)poisson;
  nu = nu < 0.0 ? 0.0 : (nu > 0.499 ? 0.499 : nu);
  const double e = (double)youngs > 0.0 ? (double)youngs : 0.0;
  return sqrt(e / ((1.0 + nu) * (double)density)) / cell;
}

/* Jacobian of row `row` with respect to a corner's local velocity */
static void row_jacobian(const orc_hex *e, uint32_t corner, uint32_t row, double *j) {
  const double *c = e->gradient + 3 * corner;
  j[0] = j[1] = j[2] = 0.0;
  switch (row) {
    case 0: j[0] = c[0]; return;
    case 1: j[1] = c[1]; return;
    case 2: j[2] = c[2]; return;
    case 3: j[0] = 0.5 * c[1]; j[1] = 0.5 * c[0]; return;
    case 4: j[1] = 0.5 * c[2]; j[2] = 0.5 * c[1]; return;
    case 5: j[0] = 0.5 * c[2]; j[2] = 0.5 * c[0]; return;
    default: j[(row - 6) % 3] = e->bend * orc_hourglass_sign((row - 6) / 3, corner); return;
  }
}

/* rows += J_corner * local for all 18 rows */
static void add_rows(double *rows, const orc_hex *e, uint32_t corner, const double *l) {
  const double *c = e->gradient + 3 * corner;
  rows[0] += l[0] * c[0];
  rows[1] += l[1] * c[1];
  rows[2] += l[2] * c[2];
  rows[3] += 0.5 * (l[0] * c[1] + l[1] * c[0]);
  rows[4] += 0.5 * (l[1] * c[2] + l[2] * c[1]);
  rows[5] += 0.5 * (l[2] * c[0] + l[0] * c[2]);
  for (uint32_t m = 0; m < 4; ++m) {
    const double sign = orc_hourglass_sign(m, corner) * e->bend;
    rows[6 + 3 * m] += sign * l[0];
    rows[7 + 3 * m] += sign * l[1];
    rows[8 + 3 * m] += sign * l[2];
  }
}

/* J_corner^T * impulse, the corner's force in the element frame */
static void row_force(const orc_hex *e, uint32_t corner, const double *imp, double *out) {
  const double *c = e->gradient + 3 * corner;
  out[0] = imp[0] * c[0] + 0.5 * (imp[3] * c[1] + imp[5] * c[2]);
  out[1] = imp[1] * c[1] + 0.5 * (imp[3] * c[0] + imp[4] * c[2]);
  out[2] = imp[2] * c[2] + 0.5 * (imp[4] * c[1] + imp[5] * c[0]);
  for (uint32_t m = 0; m < 4; ++m) {
    const double sign = orc_hourglass_sign(m, corner) * e->bend;
    out[0] += sign * imp[6 + 3 * m];
    out[1] += sign * imp[7 + 3 * m];
    out[2] += sign * imp[8 + 3 * m];
  }
}

/* critically damped soft spring of the row (theta = 1, zeta = 1) */
static void spring_of(double omega, double zeta, double h, double theta, double *rate,
                      double *softness) {
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

static void symmetric_inverse(const double *block, const int *inert, double *out) {
  double m[ROWS * ROWS];
  memcpy(m, block, sizeof m);
  memset(out, 0, ROWS * ROWS * sizeof(double));
  for (uint32_t r = 0; r < ROWS; ++r) {
    if (inert[r]) {
      for (uint32_t k = 0; k < ROWS; ++k) {
        m[ROWS * r + k] = 0.0;
        m[ROWS * k + r] = 0.0;
      }
      m[(ROWS + 1) * r] = 1.0;
    }
    out[(ROWS + 1) * r] = 1.0;
  }
  for (uint32_t col = 0; col < ROWS; ++col) {
    const double pivot = m[(ROWS + 1) * col];
    if (!(pivot > 0.0)) {
      memset(out, 0, ROWS * ROWS * sizeof(double));
      return;
    }
    for (uint32_t k = 0; k < ROWS; ++k) {
      m[ROWS * col + k] /= pivot;
      out[ROWS * col + k] /= pivot;
    }
    for (uint32_t r = 0; r < ROWS; ++r) {
      const double f = m[ROWS * r + col];
      if (r == col || f == 0.0) continue;
      for (uint32_t k = 0; k < ROWS; ++k) {
        m[ROWS * r + k] -= f * m[ROWS * col + k];
        out[ROWS * r + k] -= f * out[ROWS * col + k];
      }
    }
  }
  for (uint32_t r = 0; r < ROWS; ++r)
    if (inert[r])
      for (uint32_t k = 0; k < ROWS; ++k) {
        out[ROWS * r + k] = 0.0;
        out[ROWS * k + r] = 0.0;
      }
}

/* ElementDamping (element.h:33-38) and rowDamping (element.cc:29-40): the
 * row's damping ratio from the material's target at the lattice's lowest
 * angular frequency, 1 when either is missing */
static double clampd(double v, double lo, double hi);

typedef struct {
  double target, lowest, theta, own;
} damping_t;
static const damping_t UNDAMPED = {0.0, 0.0, 1.0, 0.0};

static double row_damping(const damping_t *d, double omega) {
  if (!isfinite(d->target) || !isfinite(d->lowest) || !(d->lowest > 0.0)) return 1.0;
  const double z = d->target * omega / d->lowest - d->own * omega / 2.0;
  return clampd(z, 0.0, 1.0);
}

static int hex_make_damped(orc_hex *e, uint32_t cell, const uint32_t *nodes, const double *inv_mass,
                           const double *shape, const double *hourglass, double volume, double shear,
                           double first, double h, const damping_t *damping);

int orc_hex_make(orc_hex *e, uint32_t cell, const uint32_t *nodes, const double *inv_mass,
                 const double *shape, const double *hourglass, double volume, double shear,
                 double first, double h) {
  return hex_make_damped(e, cell, nodes, inv_mass, shape, hourglass, volume, shear, first, h, &UNDAMPED);
}

static int hex_make_damped(orc_hex *e, uint32_t cell, const uint32_t *nodes, const double *inv_mass,
                           const double *shape, const double *hourglass, double volume, double shear,
                           double first, double h, const damping_t *damping) {
  const double det = orc_det3(shape);
  if (!(fabs(det) > 0.0) || !(shear > 0.0) || !(volume > 0.0)) return 0;
  memset(e, 0, sizeof *e);
  double inv[9];
  inverse3(shape, inv);
  e->cell = cell;
  e->size = cbrt(8.0 * fabs(det));
  e->bend = 1.0 / (4.0 * e->size);
  memcpy(e->rest_hg, hourglass, sizeof e->rest_hg);
  for (uint32_t c = 0; c < 8; ++c) {
    double s[3], g[3];
    signs_of(c, s);
    e->nodes[c] = nodes[c];
    e->inv_mass[c] = inv_mass[c];
    mtv3(inv, s, g); /* shape^-T s */
    for (int r = 0; r < 3; ++r) e->gradient[3 * c + r] = g[r] / 8.0;
  }
  double block[ROWS * ROWS];
  memset(block, 0, sizeof block);
  double vweight = 0.0;
  for (uint32_t c = 0; c < 8; ++c) {
    const double w = e->inv_mass[c];
    if (w == 0.0) continue;
    vweight += w * dot3(e->gradient + 3 * c, e->gradient + 3 * c);
    for (uint32_t r = 0; r < ROWS; ++r) {
      double jl[3];
      row_jacobian(e, c, r, jl);
      for (uint32_t k = 0; k < ROWS; ++k) {
        double jr[3];
        row_jacobian(e, c, k, jr);
        block[ROWS * r + k] += w * dot3(jl, jr);
      }
    }
  }
  if (!(vweight > 0.0)) return 0;
  const double hg = shear * (3.0 * first + 2.0 * shear) / (3.0 * (first + shear));
  int inert[ROWS];
  for (uint32_t r = 0; r < ROWS; ++r) {
    const double modulus = r < 3 ? 2.0 * shear : (r < 6 ? 4.0 * shear : hg);
    const double weight = block[(ROWS + 1) * r];
    const double omega = sqrt(modulus * volume * weight);
    e->stiffness[r] = modulus * volume;
    inert[r] = 0;
    if (omega > 0.0) {
      double rate, soft;
      spring_of(omega, row_damping(damping, omega), h, damping->theta, &rate, &soft);
      e->rate[r] = rate;
      e->softness[r] = weight * soft;
      block[(ROWS + 1) * r] += e->softness[r];
      e->cap[r] = MAX_CORRECTIVE_SPEED / (rate * e->size);
    } else {
      inert[r] = 1;
      e->cap[r] = INFINITY;
    }
  }
  symmetric_inverse(block, inert, e->inverse);
  e->vol_stiffness = first * volume;
  e->vol_cap = INFINITY;
  const double vomega = sqrt(e->vol_stiffness * vweight);
  if (vomega > 0.0) {
    double rate, soft;
    spring_of(vomega, row_damping(damping, vomega), h, damping->theta, &rate, &soft);
    e->vol_rate = rate;
    e->vol_softness = soft;
    e->vol_cap = rate > 0.0 ? MAX_CORRECTIVE_SPEED / (rate * e->size) : INFINITY;
    e->volumetric = 1;
  }
  e->quat[0] = 1.0;
  e->turn[0] = e->turn[4] = e->turn[8] = 1.0;
  return 1;
}

/* ---- quaternions (w, x, y, z) ------------------------------------------ */

static void quat_normalize(double *q) {
  const double len = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (len <= 0.0) {
    q[0] = 1.0;
    q[1] = q[2] = q[3] = 0.0;
    return;
  }
  const double s = 1.0 / len;
  for (int k = 0; k < 4; ++k) q[k] *= s;
}

/* Hamilton product p q */
static void quat_mul(const double *p, const double *q, double *out) {
  out[0] = p[0] * q[0] - p[1] * q[1] - p[2] * q[2] - p[3] * q[3];
  out[1] = p[0] * q[1] + p[1] * q[0] + p[2] * q[3] - p[3] * q[2];
  out[2] = p[0] * q[2] + p[2] * q[0] + p[3] * q[1] - p[1] * q[3];
  out[3] = p[0] * q[3] + p[3] * q[0] + p[1] * q[2] - p[2] * q[1];
}

void orc_quat_matrix(const double *q, double *R) {
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

void orc_rotation_of(const double *F, const double *guess, uint32_t iterations, double *out) {
  double q[4] = {guess[0], guess[1], guess[2], guess[3]};
  quat_normalize(q);
  for (uint32_t it = 0; it < iterations; ++it) {
    double T[9];
    orc_quat_matrix(q, T);
    /* torque = sum_k T_col_k x F_col_k, scale = |sum_k T_col_k . F_col_k| */
    double torque[3] = {0.0, 0.0, 0.0}, tr = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double tc[3] = {T[k], T[3 + k], T[6 + k]};
      const double fc[3] = {F[k], F[3 + k], F[6 + k]};
      double cr[3];
      cross3(tc, fc, cr);
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
    const double scale = fabs(tr) + 1.0e-9;
    const double om[3] = {torque[0] / scale, torque[1] / scale, torque[2] / scale};
    const double angle = sqrt(dot3(om, om));
    if (!(angle > 1.0e-12)) break;
    const double half = angle * 0.5, s = sin(half);
    const double turn[4] = {cos(half), om[0] / angle * s, om[1] / angle * s, om[2] / angle * s};
    double next[4];
    quat_mul(turn, q, next);
    memcpy(q, next, sizeof q);
    quat_normalize(q);
    if (angle <= ROTATION_TOLERANCE) break;
  }
  memcpy(out, q, sizeof q);
}

/* ---- element state ------------------------------------------------------ */

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (hi < v ? hi : v); }

void orc_hex_update(orc_hex *e, const double *pos) {
  double F[9] = {0}, hgv[12] = {0};
  for (uint32_t c = 0; c < 8; ++c) {
    const double *x = pos + 3 * (size_t)e->nodes[c];
    const double *g = e->gradient + 3 * c;
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) F[3 * r + k] += x[r] * g[k];
    for (uint32_t m = 0; m < 4; ++m) {
      const double sg = orc_hourglass_sign(m, c);
      for (int r = 0; r < 3; ++r) hgv[3 * m + r] += sg * x[r];
    }
  }
  orc_rotation_of(F, e->quat, ROTATION_ITERATIONS, e->quat);
  orc_quat_matrix(e->quat, e->turn);
  const double *T = e->turn;
  double L[9]; /* T^T F */
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) L[3 * r + k] = T[r] * F[k] + T[3 + r] * F[3 + k] + T[6 + r] * F[6 + k];
  const double det = orc_det3(F);
  const int inverted = det < 0.0;
  double raw[ROWS];
  raw[0] = L[0] - 1.0;
  raw[1] = L[4] - 1.0;
  raw[2] = L[8] - 1.0;
  raw[3] = 0.5 * (L[1] + L[3]);
  raw[4] = 0.5 * (L[5] + L[7]);
  raw[5] = 0.5 * (L[6] + L[2]);
  for (uint32_t m = 0; m < 4; ++m) {
    double b[3];
    mtv3(T, hgv + 3 * m, b);
    for (int k = 0; k < 3; ++k) raw[6 + 3 * m + k] = (b[k] - e->rest_hg[3 * m + k]) * e->bend;
  }
  for (uint32_t r = 0; r < ROWS; ++r) {
    const double cap = inverted ? e->cap[r] : (e->cap[r] < 1.0 ? 1.0 : e->cap[r]);
    e->strain[r] = clampd(raw[r], -cap, cap);
  }
  /* cofactor columns: F1 x F2, F2 x F0, F0 x F1 */
  double cof[9];
  {
    const double f0[3] = {F[0], F[3], F[6]}, f1[3] = {F[1], F[4], F[7]}, f2[3] = {F[2], F[5], F[8]};
    double c0[3], c1[3], c2[3];
    cross3(f1, f2, c0);
    cross3(f2, f0, c1);
    cross3(f0, f1, c2);
    for (int r = 0; r < 3; ++r) {
      cof[3 * r] = c0[r];
      cof[3 * r + 1] = c1[r];
      cof[3 * r + 2] = c2[r];
    }
  }
  double vweight = 0.0;
  memset(e->coupling, 0, sizeof e->coupling);
  for (uint32_t c = 0; c < 8; ++c) {
    double vg[3], b[3], l[3];
    mv3(cof, e->gradient + 3 * c, vg);
    const double w = e->inv_mass[c];
    memcpy(e->vol_gradient + 3 * c, vg, sizeof vg);
    vweight += w * dot3(vg, vg);
    mtv3(T, vg, b);
    for (int k = 0; k < 3; ++k) l[k] = b[k] * w;
    add_rows(e->coupling, e, c, l);
  }
  const double vcap = inverted ? e->vol_cap : (e->vol_cap < 1.0 ? 1.0 : e->vol_cap);
  const double volume = clampd(det - 1.0, -vcap, vcap);
  double coupled = 0.0;
  for (uint32_t r = 0; r < ROWS; ++r) {
    double t = 0.0;
    for (uint32_t k = 0; k < ROWS; ++k) t += e->inverse[ROWS * r + k] * e->coupling[k];
    e->image[r] = t;
    coupled += e->coupling[r] * t;
  }
  e->gain = vweight * e->vol_softness;
  const double schur = vweight + e->gain - coupled;
  e->schur = e->volumetric && schur > 0.0 ? 1.0 / schur : 0.0;
  e->rhs = volume * e->vol_rate;
}

void orc_hex_apply(const orc_hex *e, double *vel, const double *impulse, double vol_impulse) {
  for (uint32_t c = 0; c < 8; ++c) {
    const double w = e->inv_mass[c];
    if (w == 0.0) continue;
    double lf[3], f[3];
    row_force(e, c, impulse, lf);
    mv3(e->turn, lf, f);
    double *v = vel + 3 * (size_t)e->nodes[c];
    for (int k = 0; k < 3; ++k) v[k] -= (f[k] + e->vol_gradient[3 * c + k] * vol_impulse) * w;
  }
}

void orc_hex_solve(orc_hex *e, double *vel) {
  double res[ROWS];
  memset(res, 0, sizeof res);
  double vspeed = 0.0;
  for (uint32_t c = 0; c < 8; ++c) {
    const double *v = vel + 3 * (size_t)e->nodes[c];
    double b[3];
    vspeed += dot3(e->vol_gradient + 3 * c, v);
    mtv3(e->turn, v, b);
    add_rows(res, e, c, b);
  }
  for (uint32_t r = 0; r < ROWS; ++r)
    res[r] += e->rate[r] * e->strain[r] - e->softness[r] * e->impulse[r];
  double sol[ROWS];
  double coupled = 0.0;
  for (uint32_t r = 0; r < ROWS; ++r) {
    double t = 0.0;
    for (uint32_t k = 0; k < ROWS; ++k) t += e->inverse[ROWS * r + k] * res[k];
    sol[r] = t;
    coupled += e->coupling[r] * t;
  }
  const double volume = e->rhs - e->gain * e->vol_impulse + vspeed;
  const double change = e->schur * (volume - coupled);
  for (uint32_t r = 0; r < ROWS; ++r) {
    sol[r] -= e->image[r] * change;
    e->impulse[r] += sol[r];
  }
  e->vol_impulse += change;
  orc_hex_apply(e, vel, sol, change);
}

/* ---- hex lattice mesh --------------------------------------------------- */

/* laneOf (solver.cc:30-37): a lane past the table reads the default material */
static const orc_material *mat_of(const orc_material *m, uint32_t lanes, uint32_t lane,
                                  orc_material *fallback) {
  if (lane < lanes) return &m[lane];
  memset(fallback, 0, sizeof *fallback);
  fallback->friction = 0.5f;
  fallback->density = 1000.0f;
  fallback->strength = INFINITY;
  fallback->poissons_ratio = 0.3f;
  fallback->yield_stress = INFINITY;
  fallback->damping_ratio = INFINITY;
  fallback->tear_strain = INFINITY;
  fallback->plastic_max = INFINITY;
  fallback->crush_stress = INFINITY;
  fallback->crush_limit = 1.0f;
  return fallback;
}

int orc_build_hex_lattice(const orc_lattice *L, const orc_material *materials, uint32_t lanes,
                          const orc_support *supports, uint32_t nsupports, orc_hex_mesh *out) {
  memset(out, 0, sizeof *out);
  if (L->res[0] < 2 || L->res[1] < 2 || L->res[2] < 2) return -1;
  const uint32_t cx = L->res[0] - 1, cy = L->res[1] - 1, cz = L->res[2] - 1;
  const size_t samples = (size_t)L->res[0] * L->res[1] * L->res[2];
  uint32_t *node_of = (uint32_t *)malloc(samples * sizeof(uint32_t));
  for (size_t s = 0; s < samples; ++s) node_of[s] = UINT32_MAX;
  uint32_t nodes = 0, hexes = 0;
  for (uint32_t z = 0; z < cz; ++z)
    for (uint32_t y = 0; y < cy; ++y)
      for (uint32_t x = 0; x < cx; ++x) {
        const size_t cell = x + (size_t)cx * (y + (size_t)cy * z);
        if (L->solid && !L->solid[cell]) continue;
        ++hexes;
        for (uint32_t c = 0; c < 8; ++c) {
          const size_t s = (x + (c & 1u)) + (size_t)L->res[0] * ((y + ((c >> 1) & 1u)) +
                                                                 (size_t)L->res[1] * (z + ((c >> 2) & 1u)));
          if (node_of[s] == UINT32_MAX) node_of[s] = nodes++;
        }
      }
  if (nodes == 0) {
    free(node_of);
    return -1;
  }
  out->nodes = nodes;
  out->hexes = hexes;
  for (int k = 0; k < 3; ++k) out->voxel[k] = L->voxel[k];
  out->rest = (double *)calloc(3 * (size_t)nodes, sizeof(double));
  out->mass = (float *)calloc(nodes, sizeof(float));
  out->anchored = (uint8_t *)calloc(nodes, 1);
  out->hex = (uint32_t *)malloc(8 * (size_t)hexes * sizeof(uint32_t));
  out->volume = (float *)malloc((size_t)hexes * sizeof(float));
  out->lane = (uint32_t *)malloc((size_t)hexes * sizeof(uint32_t));
  float *restf = (float *)malloc(3 * (size_t)nodes * sizeof(float));
  for (uint32_t az = 0; az < L->res[2]; ++az)
    for (uint32_t ay = 0; ay < L->res[1]; ++ay)
      for (uint32_t ax = 0; ax < L->res[0]; ++ax) {
        const size_t s = ax + (size_t)L->res[0] * (ay + (size_t)L->res[1] * az);
        const uint32_t n = node_of[s];
        if (n == UINT32_MAX) continue;
        float p[3];
        if (L->sample_pos) {
          p[0] = L->sample_pos[3 * s + 0];
          p[1] = L->sample_pos[3 * s + 1];
          p[2] = L->sample_pos[3 * s + 2];
        } else {
          p[0] = L->origin[0] + L->voxel[0] * (float)ax;
          p[1] = L->origin[1] + L->voxel[1] * (float)ay;
          p[2] = L->origin[2] + L->voxel[2] * (float)az;
        }
        for (int k = 0; k < 3; ++k) {
          restf[3 * n + k] = p[k];
          out->rest[3 * n + k] = (double)p[k];
        }
      }
  for (uint32_t n = 0; n < nodes; ++n)
    for (uint32_t v = 0; v < nsupports; ++v) {
      const float *M = supports[v].world_to_support;
      int inside = 1;
      for (int r = 0; r < 3 && inside; ++r) {
        const float l = M[r * 4 + 0] * restf[3 * n + 0] + M[r * 4 + 1] * restf[3 * n + 1] +
                        M[r * 4 + 2] * restf[3 * n + 2] + M[r * 4 + 3];
        if (!(fabsf(l) <= supports[v].half[r])) inside = 0;
      }
      if (inside) {
        out->anchored[n] = 1;
        break;
      }
    }
  uint32_t h = 0;
  orc_material fb;
  for (uint32_t z = 0; z < cz; ++z)
    for (uint32_t y = 0; y < cy; ++y)
      for (uint32_t x = 0; x < cx; ++x) {
        const size_t cell = x + (size_t)cx * (y + (size_t)cy * z);
        if (L->solid && !L->solid[cell]) continue;
        const uint32_t lane = L->cell_lane ? L->cell_lane[cell] : 0u;
        const float density = mat_of(materials, lanes, lane, &fb)->density;
        const float vol = 1.0f * L->voxel[0] * L->voxel[1] * L->voxel[2];
        const float share = density * vol * 0.125f;
        for (uint32_t c = 0; c < 8; ++c) {
          const size_t s = (x + (c & 1u)) + (size_t)L->res[0] * ((y + ((c >> 1) & 1u)) +
                                                                 (size_t)L->res[1] * (z + ((c >> 2) & 1u)));
          out->hex[8 * (size_t)h + c] = node_of[s];
          out->mass[node_of[s]] += share;
        }
        out->volume[h] = vol;
        out->lane[h] = lane;
        ++h;
      }
  free(restf);
  free(node_of);
  return 0;
}

void orc_hex_mesh_free(orc_hex_mesh *m) {
  if (!m) return;
  free(m->rest);
  free(m->mass);
  free(m->anchored);
  free(m->hex);
  free(m->volume);
  free(m->lane);
  memset(m, 0, sizeof *m);
}

/* ---- lattice body ------------------------------------------------------- */

struct orc_lattice_body {
  uint32_t nodes, hexes, lanes;
  double *rest;
  float *mass;
  uint8_t *anchored;
  uint32_t *hex;
  float *volume;
  uint32_t *lane;
  orc_material *materials;
  float voxel[3];
  double cell; /* min voxel */
  double *x, *v;
  uint8_t *node_colour, *hex_colour;
  uint32_t substeps; /* LatticeBody::substeps */
  /* cached elements (StrainLattice) */
  double h;          /* 0 = not built */
  orc_hex *elements;
  uint32_t count;
  /* plasticity: per Gauss point P (9) and plastic strain, the P the elements
   * were last built from, and per cell the rest the element was built on */
  double *plastic, *hardening, *cached;
  double *rest_shape, *rest_hg, *rest_vol; /* 9, 12, 1 per cell; volume 0 = none */
  /* tearing: per-cell crack-face bits (NULL before the first tear), and the
   * mesh changed since the elements were built */
  uint8_t *crack;
  int retopo;
  /* PreparedMesh::surfaceNodes as a per-node flag (NULL until first asked
   * for or torn) */
  uint8_t *surface;
  uint32_t surface_nodes; /* nodes the flags cover */
  /* the lowest angular frequency the elements were damped for
   * (StrainLattice::lowest; 0 = not found yet) */
  double lowest;
};

orc_lattice_body *orc_lattice_create(const orc_hex_mesh *mesh, const orc_material *materials,
                                     uint32_t lanes, const double *offset,
                                     const double *velocity) {
  if (lanes == 0) return NULL; /* strainstep.cc:87-89 throws without lanes */
  orc_lattice_body *b = (orc_lattice_body *)calloc(1, sizeof *b);
  b->nodes = mesh->nodes;
  b->hexes = mesh->hexes;
  b->lanes = lanes;
  const size_t n = mesh->nodes;
  b->rest = (double *)malloc(3 * n * sizeof(double));
  memcpy(b->rest, mesh->rest, 3 * n * sizeof(double));
  b->mass = (float *)malloc(n * sizeof(float));
  memcpy(b->mass, mesh->mass, n * sizeof(float));
  b->anchored = (uint8_t *)malloc(n);
  memcpy(b->anchored, mesh->anchored, n);
  b->hex = (uint32_t *)malloc(8 * (size_t)mesh->hexes * sizeof(uint32_t));
  memcpy(b->hex, mesh->hex, 8 * (size_t)mesh->hexes * sizeof(uint32_t));
  b->volume = (float *)malloc((size_t)mesh->hexes * sizeof(float));
  memcpy(b->volume, mesh->volume, (size_t)mesh->hexes * sizeof(float));
  b->lane = (uint32_t *)malloc((size_t)mesh->hexes * sizeof(uint32_t));
  memcpy(b->lane, mesh->lane, (size_t)mesh->hexes * sizeof(uint32_t));
  b->materials = (orc_material *)malloc(lanes * sizeof(orc_material));
  memcpy(b->materials, materials, lanes * sizeof(orc_material));
  double vox[3] = {mesh->voxel[0], mesh->voxel[1], mesh->voxel[2]};
  for (int k = 0; k < 3; ++k) b->voxel[k] = mesh->voxel[k];
  b->cell = vox[0] < vox[1] ? vox[0] : vox[1];
  if (vox[2] < b->cell) b->cell = vox[2];
  b->x = (double *)malloc(3 * n * sizeof(double));
  b->v = (double *)malloc(3 * n * sizeof(double));
  b->node_colour = (uint8_t *)malloc(n);
  b->hex_colour = (uint8_t *)malloc(mesh->hexes ? mesh->hexes : 1);
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k)
      if ((double)(float)b->rest[3 * i + k] < lo[k]) lo[k] = (double)(float)b->rest[3 * i + k];
  for (size_t i = 0; i < n; ++i) {
    uint32_t colour = 0;
    for (int k = 0; k < 3; ++k) {
      b->x[3 * i + k] = b->rest[3 * i + k] + offset[k];
      b->v[3 * i + k] = b->anchored[i] ? 0.0 : velocity[k];
      const double index = round((b->rest[3 * i + k] - lo[k]) / vox[k]);
      colour |= ((uint32_t)index & 1u) << k;
    }
    b->node_colour[i] = (uint8_t)colour;
  }
  for (uint32_t c = 0; c < mesh->hexes; ++c) b->hex_colour[c] = b->node_colour[b->hex[8 * (size_t)c]];
  const size_t H = mesh->hexes ? mesh->hexes : 1;
  b->rest_shape = (double *)calloc(9 * H, sizeof(double));
  b->rest_hg = (double *)calloc(12 * H, sizeof(double));
  b->rest_vol = (double *)calloc(H, sizeof(double));
  return b;
}

void orc_lattice_free(orc_lattice_body *b) {
  if (!b) return;
  free(b->rest);
  free(b->mass);
  free(b->anchored);
  free(b->hex);
  free(b->volume);
  free(b->lane);
  free(b->materials);
  free(b->x);
  free(b->v);
  free(b->node_colour);
  free(b->hex_colour);
  free(b->elements);
  free(b->plastic);
  free(b->hardening);
  free(b->cached);
  free(b->rest_shape);
  free(b->rest_hg);
  free(b->rest_vol);
  free(b->crack);
  free(b->surface);
  free(b);
}

void orc_lattice_set_substeps(orc_lattice_body *b, uint32_t substeps) { b->substeps = substeps; }

void orc_lattice_get_state(const orc_lattice_body *b, double *x, double *v) {
  if (x) memcpy(x, b->x, 3 * (size_t)b->nodes * sizeof(double));
  if (v) memcpy(v, b->v, 3 * (size_t)b->nodes * sizeof(double));
}

void orc_lattice_set_state(orc_lattice_body *b, const double *x, const double *v) {
  if (x) memcpy(b->x, x, 3 * (size_t)b->nodes * sizeof(double));
  if (v) memcpy(b->v, v, 3 * (size_t)b->nodes * sizeof(double));
}

void orc_lattice_colours(const orc_lattice_body *b, uint8_t *node_colour, uint8_t *hex_colour) {
  if (node_colour) memcpy(node_colour, b->node_colour, b->nodes);
  if (hex_colour) memcpy(hex_colour, b->hex_colour, b->hexes);
}

uint32_t orc_lattice_elements(const orc_lattice_body *b) { return b->count; }

void orc_lattice_impulses(const orc_lattice_body *b, uint32_t *cells, double *impulses) {
  for (uint32_t i = 0; i < b->count; ++i) {
    if (cells) cells[i] = b->elements[i].cell;
    if (impulses) {
      memcpy(impulses + 19 * (size_t)i, b->elements[i].impulse, ROWS * sizeof(double));
      impulses[19 * (size_t)i + ROWS] = b->elements[i].vol_impulse;
    }
  }
}

static const orc_material *lane_mat(const orc_lattice_body *b, uint32_t cell) {
  const uint32_t lane = b->lane[cell];
  return &b->materials[lane < b->lanes ? lane : b->lanes - 1];
}

static int pinned(const orc_lattice_body *b, size_t node) {
  return !(b->mass[node] > 0.0f) || b->anchored[node] != 0;
}

static double stiffest(const orc_lattice_body *b) {
  double f = 0.0;
  for (uint32_t c = 0; c < b->hexes; ++c) {
    const orc_material *m = lane_mat(b, c);
    const double g = orc_hex_cell_frequency(m->youngs_modulus, m->poissons_ratio, m->density, b->cell);
    f = f < g ? g : f; /* std::max(f, g) */
  }
  return f;
}

uint32_t orc_lattice_substeps(const orc_lattice_body *b, double dt, uint32_t softstep) {
  const double f = stiffest(b);
  uint32_t needed = 1;
  if (f > 0.0 && isfinite(f)) {
    needed = (uint32_t)clampd(ceil(dt * f / RESOLVED), 1.0, (double)MOST_SUBSTEPS);
    while (needed < MOST_SUBSTEPS && dt / (double)needed * f > RESOLVED) ++needed;
  }
  uint32_t s = softstep > 1 ? softstep : 1;
  if (b->substeps > s) s = b->substeps;
  if (needed > s) s = needed;
  return s;
}

/* the proper rotation of the mass-weighted moment of positions about the
 * centre of mass against rest positions (the lattice turn) */
static void lattice_turn(const orc_lattice_body *b, double *q) {
  double c[3] = {0, 0, 0}, rc[3] = {0, 0, 0}, total = 0.0;
  for (size_t i = 0; i < b->nodes; ++i) {
    const double m = b->mass[i];
    for (int k = 0; k < 3; ++k) {
      c[k] += m * b->x[3 * i + k];
      rc[k] += m * b->rest[3 * i + k];
    }
    total += m;
  }
  q[0] = 1.0;
  q[1] = q[2] = q[3] = 0.0;
  if (!(total > 0.0)) return;
  for (int k = 0; k < 3; ++k) {
    c[k] /= total;
    rc[k] /= total;
  }
  double M[9] = {0};
  for (size_t i = 0; i < b->nodes; ++i) {
    const double m = b->mass[i];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k)
        M[3 * r + k] += m * (b->x[3 * i + r] - c[r]) * (b->rest[3 * i + k] - rc[k]);
  }
  const double id[4] = {1.0, 0.0, 0.0, 0.0};
  orc_rotation_of(M, id, 64, q);
}

/* ---- plasticity ---------------------------------------------------------- */

static const double GAUSS_OFFSET = 0.5773502691896258; /* plasticflow.cc:14 */

static void mm3(const double *A, const double *B, double *C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

void orc_hex_shape_gradients(const double *at, const double *half, double *grad) {
  for (uint32_t c = 0; c < 8; ++c) {
    double s[3];
    signs_of(c, s);
    grad[3 * c + 0] = 0.125 * s[0] * (1.0 + s[1] * at[1]) * (1.0 + s[2] * at[2]) / half[0];
    grad[3 * c + 1] = 0.125 * s[1] * (1.0 + s[0] * at[0]) * (1.0 + s[2] * at[2]) / half[1];
    grad[3 * c + 2] = 0.125 * s[2] * (1.0 + s[0] * at[0]) * (1.0 + s[1] * at[1]) / half[2];
  }
}

int orc_hex_return_map(const double *E, double shear, double yield, double *mapped, double *trial,
                       double *strain) {
  double left[9], w[3], V[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) left[3 * i + j] = E[3 * i] * E[3 * j] + E[3 * i + 1] * E[3 * j + 1] + E[3 * i + 2] * E[3 * j + 2];
  orc_eigen_sym3(left, w, V);
  const double lo = w[0] < w[1] ? (w[0] < w[2] ? w[0] : w[2]) : (w[1] < w[2] ? w[1] : w[2]);
  if (!(lo > 0.0)) return 0;
  double e[3], dev[3];
  for (int k = 0; k < 3; ++k) e[k] = 0.5 * log(w[k]);
  const double mean = (e[0] + e[1] + e[2]) / 3.0;
  for (int k = 0; k < 3; ++k) dev[k] = e[k] - mean;
  const double size = sqrt(dev[0] * dev[0] + dev[1] * dev[1] + dev[2] * dev[2]);
  const double stress = sqrt(1.5) * 2.0 * shear * size;
  if (!(stress > yield) || size <= 1.0e-15) return 0;
  double scale[3];
  for (int k = 0; k < 3; ++k) scale[k] = exp(dev[k] * (yield / stress - 1.0));
  double corr[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t = 0.0;
      for (int k = 0; k < 3; ++k) t += V[3 * i + k] * scale[k] * V[3 * j + k];
      corr[3 * i + j] = t;
    }
  mm3(corr, E, mapped);
  *trial = stress;
  *strain = (stress - yield) / (3.0 * shear);
  return 1;
}

void orc_hex_fitted_jacobian(const double *shape, const double *hourglass, uint32_t point, double *J) {
  double s[3];
  signs_of(point, s);
  const double at[3] = {GAUSS_OFFSET * s[0], GAUSS_OFFSET * s[1], GAUSS_OFFSET * s[2]};
  const double slopes[12] = {at[1], at[0], 0.0, 0.0, at[2], at[1], at[2], 0.0, at[0],
                             at[1] * at[2], at[0] * at[2], at[0] * at[1]};
  memcpy(J, shape, 9 * sizeof(double));
  for (uint32_t m = 0; m < 4; ++m)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) J[3 * i + j] += hourglass[3 * m + i] * slopes[3 * m + j] / 8.0;
}

int orc_hex_fit_holds(const double *shape, const double *hourglass) {
  for (uint32_t p = 0; p < 8; ++p) {
    double J[9];
    orc_hex_fitted_jacobian(shape, hourglass, p, J);
    if (!(orc_det3(J) > 0.0)) return 0;
  }
  return 1;
}

static void authored_rest(const orc_lattice_body *b, uint32_t cell, double *shape, double *hg, double *vol) {
  double corners[24];
  for (uint32_t c = 0; c < 8; ++c)
    for (int k = 0; k < 3; ++k) corners[3 * c + k] = b->rest[3 * (size_t)b->hex[8 * (size_t)cell + c] + k];
  orc_hex_corner_rest(corners, shape, hg);
  *vol = (double)b->volume[cell];
}

void orc_lattice_plastic_rest(const orc_lattice_body *b, uint32_t cell, double *shape, double *hg,
                              double *volume) {
  double ashape[9], ahg[12], avol;
  authored_rest(b, cell, ashape, ahg, &avol);
  if (!b->plastic) {
    memcpy(shape, ashape, sizeof ashape);
    memcpy(hg, ahg, sizeof ahg);
    *volume = avol;
    return;
  }
  double stretch[8][9], mean[9] = {0}, vol = 0.0;
  for (uint32_t p = 0; p < 8; ++p) {
    inverse3(b->plastic + 9 * (8 * (size_t)cell + p), stretch[p]);
    for (int k = 0; k < 9; ++k) mean[k] += stretch[p][k] / 8.0;
    vol += orc_det3(stretch[p]) / 8.0;
  }
  const double factor = sqrt(3.0) / 2.0;
  mm3(mean, ashape, shape);
  *volume = avol * vol;
  memset(hg, 0, 12 * sizeof(double));
  for (uint32_t p = 0; p < 8; ++p) {
    double target[9], s[3];
    mm3(stretch[p], ashape, target);
    signs_of(p, s);
    const double d0[3] = {s[1], s[0], 0.0}, d1[3] = {0.0, s[2], s[1]}, d2[3] = {s[2], 0.0, s[0]};
    const double d3[3] = {s[1] * s[2], s[0] * s[2], s[0] * s[1]};
    double t[3];
    mv3(target, d0, t);
    for (int k = 0; k < 3; ++k) hg[k] += factor * t[k];
    mv3(target, d1, t);
    for (int k = 0; k < 3; ++k) hg[3 + k] += factor * t[k];
    mv3(target, d2, t);
    for (int k = 0; k < 3; ++k) hg[6 + k] += factor * t[k];
    mv3(target, d3, t);
    for (int k = 0; k < 3; ++k) hg[9 + k] += t[k];
  }
  double fitted = 0.0;
  for (uint32_t p = 0; p < 8; ++p) {
    double J[9];
    orc_hex_fitted_jacobian(shape, hg, p, J);
    fitted += orc_det3(J) / 8.0;
  }
  const double ratio = vol * orc_det3(ashape) / fitted;
  const double size = isfinite(ratio) && ratio > 0.0 ? cbrt(ratio) : 1.0;
  for (int k = 0; k < 9; ++k) shape[k] *= size;
  for (int k = 0; k < 12; ++k) hg[k] *= size;
}

void orc_lattice_write_fitted(orc_lattice_body *b, uint32_t cell, const double *shape, const double *hg) {
  if (!b->plastic) return;
  double ashape[9], ahg[12], avol;
  authored_rest(b, cell, ashape, ahg, &avol);
  for (uint32_t p = 0; p < 8; ++p) {
    double J[9], Ji[9];
    orc_hex_fitted_jacobian(shape, hg, p, J);
    inverse3(J, Ji);
    mm3(ashape, Ji, b->plastic + 9 * (8 * (size_t)cell + p));
  }
}

uint32_t orc_lattice_has_plastic(const orc_lattice_body *b) { return b->plastic != NULL; }

void orc_lattice_get_plastic(const orc_lattice_body *b, double *plastic, double *hardening) {
  if (!b->plastic) return;
  if (plastic) memcpy(plastic, b->plastic, 72 * (size_t)b->hexes * sizeof(double));
  if (hardening) memcpy(hardening, b->hardening, 8 * (size_t)b->hexes * sizeof(double));
}

void orc_lattice_set_plastic(orc_lattice_body *b, const double *plastic, const double *hardening) {
  if (!b->plastic) {
    b->plastic = (double *)malloc(72 * (size_t)b->hexes * sizeof(double));
    b->hardening = (double *)calloc(8 * (size_t)b->hexes, sizeof(double));
    b->cached = (double *)malloc(72 * (size_t)b->hexes * sizeof(double));
    for (size_t i = 0; i < 8 * (size_t)b->hexes; ++i) {
      memset(b->plastic + 9 * i, 0, 9 * sizeof(double));
      b->plastic[9 * i] = b->plastic[9 * i + 4] = b->plastic[9 * i + 8] = 1.0;
    }
    memcpy(b->cached, b->plastic, 72 * (size_t)b->hexes * sizeof(double));
  }
  if (plastic) memcpy(b->plastic, plastic, 72 * (size_t)b->hexes * sizeof(double));
  if (hardening) memcpy(b->hardening, hardening, 8 * (size_t)b->hexes * sizeof(double));
}

static double frobenius3(const double *m) {
  double t = 0.0;
  for (int k = 0; k < 9; ++k) t += m[k] * m[k];
  return sqrt(t);
}

double orc_lattice_flow(orc_lattice_body *b) {
  int yields = 0;
  for (uint32_t c = 0; c < b->hexes; ++c) yields = yields || isfinite(lane_mat(b, c)->yield_stress);
  if (!yields) return 0.0;
  if (!b->plastic) orc_lattice_set_plastic(b, NULL, NULL);
  double largest = 0.0;
  const double half[3] = {(double)(0.5f * (float)b->voxel[0]), (double)(0.5f * (float)b->voxel[1]),
                          (double)(0.5f * (float)b->voxel[2])};
  for (uint32_t cell = 0; cell < b->hexes; ++cell) {
    const orc_material *m = lane_mat(b, cell);
    if (!isfinite(m->yield_stress)) continue;
    double shear, first;
    orc_hex_material(m->youngs_modulus, m->poissons_ratio, &shear, &first);
    const double bulk = first + 2.0 * shear / 3.0;
    const double ratio = (double)m->hardening_modulus / (3.0 * shear);
    const int crushable = isfinite(m->crush_stress) && m->crush_limit < 1.0f;
    const double crush = (double)m->crush_stress / bulk;
    for (uint32_t p = 0; p < 8; ++p) {
      const size_t at = 8 * (size_t)cell + p;
      double s[3], grad[24];
      signs_of(p, s);
      const double xi[3] = {GAUSS_OFFSET * s[0], GAUSS_OFFSET * s[1], GAUSS_OFFSET * s[2]};
      orc_hex_shape_gradients(xi, half, grad);
      double F[9] = {0};
      for (uint32_t c = 0; c < 8; ++c) {
        const double *x = b->x + 3 * (size_t)b->hex[8 * (size_t)cell + c];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) F[3 * i + j] += x[i] * grad[3 * c + j];
      }
      if (!(orc_det3(F) > 0.0)) continue;
      double P[9], E[9];
      memcpy(P, b->plastic + 9 * at, sizeof P);
      mm3(F, P, E);
      const double start = (double)m->yield_stress + (double)m->hardening_modulus * b->hardening[at];
      const double room = crushable ? -log(orc_det3(P) * (double)m->crush_limit) : 0.0;
      const int squeezed = crushable && room > 0.0 && orc_det3(E) < exp(-crush);
      double tE[9], tt, ts;
      const int trial = orc_hex_return_map(E, shear, start, tE, &tt, &ts);
      const double now = trial ? (start + ratio * tt) / (1.0 + ratio) : start;
      double mE[9], mt = 0.0, ms = 0.0;
      const int mapped = trial ? orc_hex_return_map(E, shear, now, mE, &mt, &ms) : 0;
      const double *after = mapped ? mE : E;
      const double volume = orc_det3(after);
      double compaction = 0.0;
      if (squeezed && volume > 0.0) {
        compaction = -crush - log(volume);
        if (room < compaction) compaction = room;
      }
      if (!mapped && !(compaction > 0.0)) continue;
      const double grow = compaction > 0.0 ? exp(compaction / 3.0) : 1.0;
      double Fi[9], ga[9], next[9];
      inverse3(F, Fi);
      for (int k = 0; k < 9; ++k) ga[k] = grow * after[k];
      mm3(Fi, ga, next);
      const double kept = orc_det3(next);
      const double cap = (double)m->plastic_max;
      if (isfinite(cap) && kept > 0.0) {
        double st[9];
        inverse3(next, st);
        st[0] -= 1.0;
        st[4] -= 1.0;
        st[8] -= 1.0;
        const double size = frobenius3(st);
        if (size > cap) {
          double capped[9];
          for (int k = 0; k < 9; ++k) capped[k] = (cap / size) * st[k];
          capped[0] += 1.0;
          capped[4] += 1.0;
          capped[8] += 1.0;
          const double cv = orc_det3(capped);
          if (!(cv > 1.0e-6)) continue;
          double ci[9];
          inverse3(capped, ci);
          const double sc = cbrt(kept * cv);
          for (int k = 0; k < 9; ++k) next[k] = sc * ci[k];
        }
      }
      double change = 0.0;
      for (int k = 0; k < 9; ++k) change += fabs(next[k] - P[k]);
      if (!(change > 1.0e-12)) continue;
      if (largest < change) largest = change;
      memcpy(b->plastic + 9 * at, next, sizeof next);
      b->hardening[at] += mapped ? ms : 0.0;
    }
  }
  return largest;
}

/* ---- lowest mode (strainstep/lowestmode.cc) ----------------------------- */

#define LOWEST_ITERATIONS 300u /* lowestmode.cc:16 */
#define LOWEST_TOLERANCE 1.0e-12

static double mass_dot(const double *a, const double *b, const double *m, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += a[i] * m[i] * b[i];
  return s;
}

static void deflate(double *x, const double *rigid, uint32_t modes, const double *m, size_t n) {
  for (uint32_t k = 0; k < modes; ++k) {
    const double *mode = rigid + k * n;
    const double c = mass_dot(mode, x, m, n);
    for (size_t i = 0; i < n; ++i) x[i] -= c * mode[i];
  }
}

/* the order the free coordinates are numbered in: nodes by rest position
 * along the lattice's longest extent, then the other axes, then id, so the
 * band spans one cross-section of the body */
typedef struct {
  double key[3];
  uint32_t node;
} node_key;

static int node_key_cmp(const void *x, const void *y) {
  const node_key *a = (const node_key *)x, *b = (const node_key *)y;
  for (int k = 0; k < 3; ++k) {
    if (a->key[k] < b->key[k]) return -1;
    if (b->key[k] < a->key[k]) return 1;
  }
  return a->node < b->node ? -1 : (a->node > b->node ? 1 : 0);
}

static void band_order(const double *rest, size_t nodes, uint32_t *order) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < nodes; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = rest[3 * i + k] < lo[k] ? rest[3 * i + k] : lo[k];
      hi[k] = rest[3 * i + k] > hi[k] ? rest[3 * i + k] : hi[k];
    }
  int axes[3] = {0, 1, 2};
  for (int k = 1; k < 3; ++k) /* longest extent first (first of equals) */
    if (hi[k] - lo[k] > hi[axes[0]] - lo[axes[0]]) axes[0] = k;
  for (int k = 0, at = 1; k < 3; ++k)
    if (k != axes[0]) axes[at++] = k;
  node_key *keys = (node_key *)malloc((nodes ? nodes : 1) * sizeof(node_key));
  for (size_t i = 0; i < nodes; ++i) {
    for (int k = 0; k < 3; ++k) keys[i].key[k] = rest[3 * i + axes[k]];
    keys[i].node = (uint32_t)i;
  }
  qsort(keys, nodes, sizeof(node_key), node_key_cmp);
  for (size_t i = 0; i < nodes; ++i) order[i] = keys[i].node;
  free(keys);
}

/* lowestAngular (lowestmode.cc:31-203): inverse iteration on K v = w^2 M v
 * over the free coordinates (pinned and massless nodes left out) with the
 * rows' stiffnesses along their local Jacobians, rigid modes deflated and a
 * small mass shift when nothing is held.  The shifted stiffness is factored
 * as a band L D L^T with the coordinates numbered along the longest extent
 * (the reference factors it sparse);
 * INFINITY without free coordinates, 0 when a pivot vanishes. */
static double lowest_angular(const orc_hex *el, uint32_t count, const double *mass, const uint8_t *pinned_node,
                             const double *rest, size_t nodes) {
  int64_t *index = (int64_t *)malloc((nodes ? 3 * nodes : 1) * sizeof(int64_t));
  double *w = (double *)malloc((nodes ? 3 * nodes : 1) * sizeof(double));
  size_t *owner = (size_t *)malloc((nodes ? 3 * nodes : 1) * sizeof(size_t));
  size_t n = 0;
  int held = 0;
  uint32_t *order = (uint32_t *)malloc((nodes ? nodes : 1) * sizeof(uint32_t));
  band_order(rest, nodes, order);
  for (size_t i = 0; i < 3 * nodes; ++i) index[i] = -1;
  for (size_t at = 0; at < nodes; ++at) {
    const size_t node = order[at];
    if (pinned_node[node]) {
      held = 1;
      continue;
    }
    if (!(mass[node] > 0.0)) continue;
    for (int a = 0; a < 3; ++a) {
      index[3 * node + a] = (int64_t)n;
      w[n] = mass[node];
      owner[n++] = 3 * node + (size_t)a;
    }
  }
  free(order);
  if (n == 0) {
    free(index);
    free(w);
    free(owner);
    return INFINITY;
  }
  size_t bw = 0;
  for (uint32_t e = 0; e < count; ++e)
    for (int i = 0; i < 24; ++i) {
      const int64_t p = index[3 * (size_t)el[e].nodes[i / 3] + i % 3];
      if (p < 0) continue;
      for (int j = 0; j < 24; ++j) {
        const int64_t q = index[3 * (size_t)el[e].nodes[j / 3] + j % 3];
        if (q >= 0 && (size_t)(p > q ? p - q : q - p) > bw) bw = (size_t)(p > q ? p - q : q - p);
      }
    }
  const size_t W = bw + 1; /* band row p holds (p, p - k), k = 0..bw */
  double *K = (double *)calloc(n * W, sizeof(double));
  for (uint32_t e = 0; e < count; ++e)
    for (uint32_t row = 0; row < ROWS + 1; ++row) {
      const double st = row < ROWS ? el[e].stiffness[row] : el[e].vol_stiffness;
      if (!(st > 0.0)) continue;
      double line[24];
      for (uint32_t c = 0; c < 8; ++c) {
        if (row < ROWS)
          row_jacobian(&el[e], c, row, line + 3 * c);
        else
          memcpy(line + 3 * c, el[e].gradient + 3 * c, 3 * sizeof(double));
      }
      for (int i = 0; i < 24; ++i) {
        const int64_t p = index[3 * (size_t)el[e].nodes[i / 3] + i % 3];
        if (p < 0 || line[i] == 0.0) continue;
        for (int j = 0; j < 24; ++j) {
          const int64_t q = index[3 * (size_t)el[e].nodes[j / 3] + j % 3];
          if (q >= 0 && q <= p && line[j] != 0.0) K[(size_t)p * W + (size_t)(p - q)] += st * line[i] * line[j];
        }
      }
    }
  double *rigid = (double *)malloc(6 * n * sizeof(double));
  uint32_t modes = 0;
  if (!held) {
    double centre[3] = {0.0, 0.0, 0.0}, total = 0.0;
    for (size_t i = 0; i < n; i += 3) {
      const size_t node = owner[i] / 3;
      for (int a = 0; a < 3; ++a) centre[a] += w[i] * rest[3 * node + a];
      total += w[i];
    }
    for (int a = 0; a < 3; ++a) centre[a] /= total;
    for (int mode = 0; mode < 6; ++mode) {
      double *shape = rigid + modes * n;
      for (size_t i = 0; i < n; ++i) {
        const size_t node = owner[i] / 3;
        const int axis = (int)(owner[i] % 3);
        double dir[3] = {0.0, 0.0, 0.0};
        if (mode < 3) {
          dir[mode] = 1.0;
        } else {
          double spin[3] = {0.0, 0.0, 0.0}, r[3];
          spin[mode - 3] = 1.0;
          for (int a = 0; a < 3; ++a) r[a] = rest[3 * node + a] - centre[a];
          cross3(spin, r, dir);
        }
        shape[i] = dir[axis];
      }
      deflate(shape, rigid, modes, w, n);
      const double norm = sqrt(mass_dot(shape, shape, w, n));
      if (norm > 0.0) {
        for (size_t i = 0; i < n; ++i) shape[i] /= norm;
        ++modes;
      }
    }
  }
  double traceK = 0.0, sumw = 0.0;
  for (size_t i = 0; i < n; ++i) {
    traceK += K[i * W];
    sumw += w[i];
  }
  const double shift = held ? 0.0 : 1.0e-6 * traceK / sumw;
  double *L = (double *)malloc(n * W * sizeof(double));
  memcpy(L, K, n * W * sizeof(double));
  for (size_t i = 0; i < n; ++i) L[i * W] += shift * w[i];
  double *D = (double *)malloc(n * sizeof(double));
  int ok = 1;
  for (size_t j = 0; j < n && ok; ++j) { /* band L D L^T, in place */
    const size_t lo = j > bw ? j - bw : 0;
    double d = L[j * W];
    for (size_t k = lo; k < j; ++k) d -= L[j * W + (j - k)] * L[j * W + (j - k)] * D[k];
    if (d == 0.0) {
      ok = 0;
      break;
    }
    D[j] = d;
    const size_t hi = j + bw < n - 1 ? j + bw : n - 1;
    for (size_t i = j + 1; i <= hi; ++i) {
      double v = L[i * W + (i - j)];
      for (size_t k = (i > bw ? i - bw : 0) > lo ? (i > bw ? i - bw : 0) : lo; k < j; ++k)
        v -= L[i * W + (i - k)] * L[j * W + (j - k)] * D[k];
      L[i * W + (i - j)] = v / d;
    }
  }
  double value = INFINITY;
  if (ok) {
    double *x = (double *)malloc(n * sizeof(double));
    double *z = (double *)malloc(n * sizeof(double));
    double *kz = (double *)malloc(n * sizeof(double));
    for (size_t i = 0; i < n; ++i) x[i] = sin(1.7 * (double)i + 0.3) + 0.5;
    double previous = INFINITY;
    for (uint32_t it = 0; it < LOWEST_ITERATIONS; ++it) {
      deflate(x, rigid, modes, w, n);
      const double norm = sqrt(mass_dot(x, x, w, n));
      for (size_t i = 0; i < n; ++i) x[i] /= norm;
      for (size_t i = 0; i < n; ++i) { /* z = (K + shift M)^-1 M x */
        double v = x[i] * w[i];
        for (size_t k = i > bw ? i - bw : 0; k < i; ++k) v -= L[i * W + (i - k)] * z[k];
        z[i] = v;
      }
      for (size_t i = 0; i < n; ++i) z[i] /= D[i];
      for (size_t i = n; i-- > 0;) {
        double v = z[i];
        const size_t hi = i + bw < n - 1 ? i + bw : n - 1;
        for (size_t k = i + 1; k <= hi; ++k) v -= L[k * W + (k - i)] * z[k];
        z[i] = v;
      }
      deflate(z, rigid, modes, w, n);
      for (size_t i = 0; i < n; ++i) kz[i] = 0.0;
      for (size_t p = 0; p < n; ++p) { /* K z from the lower band */
        kz[p] += K[p * W] * z[p];
        for (size_t k = 1; k <= bw && k <= p; ++k) {
          const double a = K[p * W + k];
          kz[p] += a * z[p - k];
          kz[p - k] += a * z[p];
        }
      }
      double zkz = 0.0;
      for (size_t i = 0; i < n; ++i) zkz += z[i] * kz[i];
      value = zkz / mass_dot(z, z, w, n);
      memcpy(x, z, n * sizeof(double));
      if (fabs(value - previous) <= LOWEST_TOLERANCE * fabs(value)) break;
      previous = value;
    }
    free(x);
    free(z);
    free(kz);
  }
  free(index);
  free(w);
  free(owner);
  free(K);
  free(L);
  free(D);
  free(rigid);
  if (!ok) return 0.0;
  return sqrt(value > 0.0 ? value : 0.0);
}

static double remaining(const double *before, const double *after, int n) {
  double was = 0.0, now = 0.0;
  for (int k = 0; k < n; ++k) {
    was += before[k] * before[k];
    now += after[k] * after[k];
  }
  if (!(was > 0.0)) return 1.0;
  const double r = sqrt(now / was);
  return r < 1.0 ? r : 1.0;
}

static void inverse_masses(const orc_lattice_body *b, double *inv) {
  for (size_t i = 0; i < b->nodes; ++i) inv[i] = (!pinned(b, i) && b->mass[i] > 0.0f) ? 1.0 / (double)b->mass[i] : 0.0;
}

/* cellDamping (strainstep.cc:151-160): the lane's target ratio at the
 * lattice's lowest frequency, with the implicit weight theta lowered toward
 * the midpoint rule for a body solved exactly */
static damping_t cell_damping(const orc_lattice_body *b, uint32_t cell, double h, double lowest) {
  const double target = lane_mat(b, cell)->damping_ratio;
  damping_t d = {target, isfinite(lowest) ? lowest : 0.0, 1.0, 0.0};
  const int exact = !(h * stiffest(b) > RESOLVED);
  if (exact && lowest > 0.0 && isfinite(lowest) && isfinite(target)) {
    d.theta = clampd(2.0 * target / (h * lowest), 0.5, 1.0); /* MIDPOINT, strainstep.cc:58 */
    d.own = d.theta * h;
  }
  return d;
}

static int make_cell(const orc_lattice_body *b, const double *inv, uint32_t cell, const double *shape,
                     const double *hg, double volume, double h, const damping_t *damping, orc_hex *out) {
  const uint32_t *nd = b->hex + 8 * (size_t)cell;
  double w[8], shear, first;
  for (uint32_t c = 0; c < 8; ++c) w[c] = inv[nd[c]];
  const orc_material *m = lane_mat(b, cell);
  orc_hex_material(m->youngs_modulus, m->poissons_ratio, &shear, &first);
  return hex_make_damped(out, cell, nd, w, shape, hg, volume, shear, first, h, damping);
}

/* the rest each cell's element is built on (restOf, strainstep.cc:128-136) */
static void rest_for(const orc_lattice_body *b, uint32_t cell, double *shape, double *hg, double *vol) {
  orc_lattice_plastic_rest(b, cell, shape, hg, vol);
  if (!orc_hex_fit_holds(shape, hg) && b->rest_vol[cell] > 0.0) {
    memcpy(shape, b->rest_shape + 9 * (size_t)cell, 9 * sizeof(double));
    memcpy(hg, b->rest_hg + 12 * (size_t)cell, 12 * sizeof(double));
    *vol = b->rest_vol[cell];
  }
}

/* lowestAngular of the body's undamped elements on their rests
 * (strainstep.cc:303-331) */
static double plain_lowest(const orc_lattice_body *b, const double *inv, double h) {
  orc_hex *plain = (orc_hex *)malloc((b->hexes ? b->hexes : 1) * sizeof(orc_hex));
  uint32_t count = 0;
  for (uint32_t cell = 0; cell < b->hexes; ++cell) {
    double shape[9], hg[12], vol;
    rest_for(b, cell, shape, hg, &vol);
    if (make_cell(b, inv, cell, shape, hg, vol, h, &UNDAMPED, &plain[count])) ++count;
  }
  double *mass = (double *)malloc((b->nodes ? b->nodes : 1) * sizeof(double));
  uint8_t *held = (uint8_t *)malloc(b->nodes ? b->nodes : 1);
  for (size_t i = 0; i < b->nodes; ++i) {
    mass[i] = b->mass[i];
    held[i] = (uint8_t)pinned(b, i);
  }
  const double omega = lowest_angular(plain, count, mass, held, b->rest, b->nodes);
  free(plain);
  free(mass);
  free(held);
  return omega;
}

double orc_lattice_lowest_angular(const orc_lattice_body *b) {
  double *inv = (double *)calloc(b->nodes ? b->nodes : 1, sizeof(double));
  inverse_masses(b, inv);
  const double omega = plain_lowest(b, inv, 1.0 / 240.0);
  free(inv);
  return omega;
}

double orc_lattice_lowest(const orc_lattice_body *b) { return b->lowest; }

/* refreshFlowed (strainstep.cc:371-451): rebuilds the elements of the cells
 * whose plastic state moved since their element was built */
static void refresh_flowed(orc_lattice_body *b) {
  if (!b->plastic) return;
  double *inv = (double *)malloc((b->nodes ? b->nodes : 1) * sizeof(double));
  inverse_masses(b, inv);
  for (uint32_t i = 0; i < b->count; ++i) {
    orc_hex *old = &b->elements[i];
    const uint32_t cell = old->cell;
    if (!memcmp(b->plastic + 72 * (size_t)cell, b->cached + 72 * (size_t)cell, 72 * sizeof(double))) continue;
    double shape[9], hg[12], vol;
    orc_lattice_plastic_rest(b, cell, shape, hg, &vol);
    if (!orc_hex_fit_holds(shape, hg)) continue;
    orc_hex fresh;
    const damping_t damping = cell_damping(b, cell, b->h, b->lowest);
    if (!make_cell(b, inv, cell, shape, hg, vol, b->h, &damping, &fresh)) continue;
    memcpy(fresh.quat, old->quat, sizeof fresh.quat);
    memcpy(fresh.impulse, old->impulse, sizeof fresh.impulse);
    fresh.vol_impulse = old->vol_impulse;
    orc_hex_update(old, b->x);
    orc_hex_update(&fresh, b->x);
    const double strain = remaining(old->strain, fresh.strain, 6);
    const double bend = remaining(old->strain + 6, fresh.strain + 6, ROWS - 6);
    for (uint32_t r = 0; r < ROWS; ++r) fresh.impulse[r] *= r < 6 ? strain : bend;
    orc_lattice_write_fitted(b, cell, shape, hg);
    memcpy(b->cached + 72 * (size_t)cell, b->plastic + 72 * (size_t)cell, 72 * sizeof(double));
    memcpy(b->rest_shape + 9 * (size_t)cell, shape, sizeof shape);
    memcpy(b->rest_hg + 12 * (size_t)cell, hg, sizeof hg);
    b->rest_vol[cell] = vol;
    *old = fresh;
  }
  free(inv);
}

static void build_elements(orc_lattice_body *b, double h) {
  double *inv = (double *)calloc(b->nodes ? b->nodes : 1, sizeof(double));
  inverse_masses(b, inv);
  double turn[4];
  lattice_turn(b, turn);
  /* a damped lattice keeps the lowest frequency it was first built with
   * (strainstep.cc:300-331) */
  int damped = 0;
  for (uint32_t cell = 0; cell < b->hexes; ++cell) damped = damped || isfinite(lane_mat(b, cell)->damping_ratio);
  if (damped && !(b->lowest > 0.0 && isfinite(b->lowest))) b->lowest = plain_lowest(b, inv, h);
  orc_hex *fresh = (orc_hex *)malloc((b->hexes ? b->hexes : 1) * sizeof(orc_hex));
  uint32_t count = 0;
  for (uint32_t cell = 0; cell < b->hexes; ++cell) {
    /* restOf (strainstep.cc:128-136): the plastic rest, or the rest kept from
     * the last build when the fit does not hold */
    double shape[9], hg[12], vol;
    rest_for(b, cell, shape, hg, &vol);
    const damping_t damping = cell_damping(b, cell, h, b->lowest);
    if (make_cell(b, inv, cell, shape, hg, vol, h, &damping, &fresh[count])) {
      memcpy(fresh[count].quat, turn, sizeof turn);
      orc_quat_matrix(turn, fresh[count].turn);
      ++count;
    }
    const int holds = orc_hex_fit_holds(shape, hg);
    memcpy(b->rest_shape + 9 * (size_t)cell, shape, sizeof shape);
    memcpy(b->rest_hg + 12 * (size_t)cell, hg, sizeof hg);
    b->rest_vol[cell] = holds ? vol : 0.0;
  }
  if (b->plastic) memcpy(b->cached, b->plastic, 72 * (size_t)b->hexes * sizeof(double));
  free(inv);
  free(b->elements);
  b->elements = fresh;
  b->count = count;
  b->h = h;
}

/* ---- cell tearing (strainstep/tearing.cc) -------------------------------- */

#define MIN_PIECE_CELLS 2u /* tearing.h:13 */

/* FACES (tearing.cc:27-34): the corners of face 2 axis + side */
static const uint32_t FACES[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5},
                                     {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};

void orc_sym_eigen3(const double *A, double *w, double *V) {
  double a[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      a[r][c] = A[3 * r + c];
      V[3 * r + c] = r == c ? 1.0 : 0.0;
    }
  static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (!(off > 1.0e-40 * diag)) break;
    for (int k = 0; k < 3; ++k) {
      const int p = P[k], q = Q[k];
      const double apq = a[p][q];
      if (apq == 0.0) continue;
      const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
      const double t = fabs(theta) > 1.0e150 ? 0.5 / theta
                                             : (theta < 0.0 ? -1.0 : 1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
      const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
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

static void rest_of(const orc_lattice_body *b, uint32_t node, double *r) {
  for (int k = 0; k < 3; ++k) r[k] = b->rest[3 * (size_t)node + k];
}

/* stretchOf (tearing.cc:258-300): per corner the frame of its three edges,
 * F = now rest^-1, and the largest eigenvalue of F^T F */
void orc_lattice_cell_stretch(const orc_lattice_body *b, uint32_t cell, double *stretch, double *direction) {
  const uint32_t *nd = b->hex + 8 * (size_t)cell;
  double best = 0.0, dir[3] = {1.0, 0.0, 0.0};
  for (uint32_t corner = 0; corner < 8; ++corner) {
    const uint32_t base = nd[corner];
    double R[9], N[9]; /* row-major, edge e in column e */
    for (uint32_t e = 0; e < 3; ++e) {
      const uint32_t other = nd[corner ^ (1u << e)];
      for (int k = 0; k < 3; ++k) {
        R[3 * k + e] = b->rest[3 * (size_t)other + k] - b->rest[3 * (size_t)base + k];
        N[3 * k + e] = b->x[3 * (size_t)other + k] - b->x[3 * (size_t)base + k];
      }
    }
    const double det = R[0] * (R[4] * R[8] - R[5] * R[7]) - R[1] * (R[3] * R[8] - R[5] * R[6]) +
                       R[2] * (R[3] * R[7] - R[4] * R[6]);
    if (!(fabs(det) > 0.0)) continue;
    double Ri[9], F[9], C[9], w[3], V[9];
    inverse3(R, Ri);
    mm3(N, Ri, F);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) C[3 * r + c] = F[r] * F[c] + F[3 + r] * F[3 + c] + F[6 + r] * F[6 + c];
    orc_sym_eigen3(C, w, V);
    int top = 0;
    for (int k = 1; k < 3; ++k)
      if (w[k] > w[top]) top = k;
    const double s = sqrt(w[top] > 0.0 ? w[top] : 0.0);
    if (s > best) {
      double v[3] = {V[top], V[3 + top], V[6 + top]};
      const double len = sqrt(dot3(v, v));
      int big = 0;
      for (int k = 1; k < 3; ++k)
        if (fabs(v[k]) > fabs(v[big])) big = k;
      const double sign = v[big] < 0.0 ? -1.0 : 1.0;
      for (int k = 0; k < 3; ++k) dir[k] = sign * v[k] / len;
      best = s;
    }
  }
  *stretch = best;
  memcpy(direction, dir, sizeof dir);
}

uint32_t orc_lattice_overstretched(const orc_lattice_body *b, uint32_t *cells) {
  uint32_t n = 0;
  for (uint32_t c = 0; c < b->hexes; ++c) {
    const double limit = 1.0 + (double)lane_mat(b, c)->tear_strain;
    double s, d[3];
    if (!isfinite(limit)) continue;
    orc_lattice_cell_stretch(b, c, &s, d);
    if (s > limit) cells[n++] = c;
  }
  return n;
}

static void face_of(const orc_lattice_body *b, uint32_t cell, uint32_t face, uint32_t *out) {
  for (int k = 0; k < 4; ++k) out[k] = b->hex[8 * (size_t)cell + FACES[face][k]];
  for (int i = 1; i < 4; ++i) /* sorted */
    for (int j = i; j > 0 && out[j] < out[j - 1]; --j) {
      const uint32_t t = out[j];
      out[j] = out[j - 1];
      out[j - 1] = t;
    }
}

static int face_holds(const uint32_t *f, uint32_t node) {
  return f[0] == node || f[1] == node || f[2] == node || f[3] == node;
}

static int hex_has(const orc_lattice_body *b, uint32_t cell, uint32_t node) {
  for (int k = 0; k < 8; ++k)
    if (b->hex[8 * (size_t)cell + k] == node) return 1;
  return 0;
}

/* cellsAt (tearing.cc:58-70): the cells holding node, ascending */
static uint32_t cells_at(const orc_lattice_body *b, uint32_t node, uint32_t *out) {
  uint32_t n = 0;
  for (uint32_t c = 0; c < b->hexes; ++c)
    if (hex_has(b, c, node)) out[n++] = c;
  return n;
}

static uint32_t shared_nodes(const orc_lattice_body *b, uint32_t x, uint32_t y) {
  uint32_t n = 0;
  for (int k = 0; k < 8; ++k) n += hex_has(b, y, b->hex[8 * (size_t)x + k]) ? 1u : 0u;
  return n;
}

static uint32_t root_of(uint32_t *parent, uint32_t at) {
  while (parent[at] != at) {
    parent[at] = parent[parent[at]];
    at = parent[at];
  }
  return at;
}

/* groupsOf (tearing.cc:92-137): cells of the fan on one side joined through
 * shared faces (four or more shared nodes); groups numbered side 0 first,
 * each side in fan order */
static uint32_t groups_of(const orc_lattice_body *b, const uint32_t *fan, uint32_t n, const uint8_t *sides,
                          uint32_t *group) {
  uint32_t *parent = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  uint32_t *number = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  for (uint32_t i = 0; i < n; ++i) {
    parent[i] = i;
    number[i] = UINT32_MAX;
  }
  for (uint32_t i = 0; i < n; ++i)
    for (uint32_t j = i + 1; j < n; ++j)
      if (sides[i] == sides[j] && shared_nodes(b, fan[i], fan[j]) >= 4) {
        const uint32_t x = root_of(parent, i), y = root_of(parent, j);
        if (x != y) parent[x > y ? x : y] = x < y ? x : y;
      }
  uint32_t count = 0;
  for (uint8_t side = 0; side < 2; ++side)
    for (uint32_t i = 0; i < n; ++i) {
      if (sides[i] != side) continue;
      const uint32_t r = root_of(parent, i);
      if (number[r] == UINT32_MAX) number[r] = count++;
      group[i] = number[r];
    }
  free(parent);
  free(number);
  return count;
}

/* reachable (tearing.cc:139-180): how many cells the seeds reach through
 * faces that do not hold the pivot, stopping at limit */
static size_t reachable(const orc_lattice_body *b, const uint32_t *seeds, uint32_t nseeds, uint32_t pivot,
                        size_t limit) {
  uint8_t *seen = (uint8_t *)calloc(b->hexes ? b->hexes : 1, 1);
  uint32_t *stack = (uint32_t *)malloc((b->hexes ? b->hexes : 1) * sizeof(uint32_t));
  size_t count = 0, top = 0;
  for (uint32_t i = 0; i < nseeds; ++i)
    if (!seen[seeds[i]]) {
      seen[seeds[i]] = 1;
      ++count;
      stack[top++] = seeds[i];
    }
  while (top > 0 && count < limit) {
    const uint32_t cell = stack[--top];
    for (uint32_t face = 0; face < 6; ++face) {
      uint32_t f[4];
      face_of(b, cell, face, f);
      if (face_holds(f, pivot)) continue;
      for (uint32_t other = 0; other < b->hexes; ++other) {
        if (seen[other]) continue;
        for (uint32_t g = 0; g < 6; ++g) {
          uint32_t o[4];
          face_of(b, other, g, o);
          if (!memcmp(o, f, sizeof o)) {
            seen[other] = 1;
            ++count;
            stack[top++] = other;
            break;
          }
        }
      }
    }
  }
  free(seen);
  free(stack);
  return count;
}

typedef struct {
  uint32_t *copy, *original;
  uint32_t n, cap;
} copies_t;

/* splitNode (tearing.cc:182-212): a copy of node with its rest, state and
 * colour, no mass and not anchored */
static uint32_t split_node(orc_lattice_body *b, uint32_t node) {
  const uint32_t copy = b->nodes;
  const size_t n = (size_t)b->nodes + 1;
  b->rest = (double *)realloc(b->rest, 3 * n * sizeof(double));
  b->x = (double *)realloc(b->x, 3 * n * sizeof(double));
  b->v = (double *)realloc(b->v, 3 * n * sizeof(double));
  b->mass = (float *)realloc(b->mass, n * sizeof(float));
  b->anchored = (uint8_t *)realloc(b->anchored, n);
  b->node_colour = (uint8_t *)realloc(b->node_colour, n);
  for (int k = 0; k < 3; ++k) {
    b->rest[3 * (size_t)copy + k] = b->rest[3 * (size_t)node + k];
    b->x[3 * (size_t)copy + k] = b->x[3 * (size_t)node + k];
    b->v[3 * (size_t)copy + k] = b->v[3 * (size_t)node + k];
  }
  b->mass[copy] = 0.0f;
  b->anchored[copy] = 0;
  b->node_colour[copy] = b->node_colour[node];
  b->nodes = (uint32_t)n;
  return copy;
}

/* split (tearing.cc:214-249): every group past the first gets a copy of the
 * node; anchored nodes never split */
static int split(orc_lattice_body *b, uint32_t node, const uint32_t *fan, uint32_t n, const uint32_t *group,
                 uint32_t count, copies_t *copies) {
  if (count < 2 || b->anchored[node]) return 0;
  for (uint32_t piece = 1; piece < count; ++piece) {
    const uint32_t copy = split_node(b, node);
    if (copies->n == copies->cap) {
      copies->cap = copies->cap ? 2 * copies->cap : 16;
      copies->copy = (uint32_t *)realloc(copies->copy, copies->cap * sizeof(uint32_t));
      copies->original = (uint32_t *)realloc(copies->original, copies->cap * sizeof(uint32_t));
    }
    copies->copy[copies->n] = copy;
    copies->original[copies->n++] = node;
    for (uint32_t at = 0; at < n; ++at) {
      if (group[at] != piece) continue;
      uint32_t *nd = b->hex + 8 * (size_t)fan[at];
      for (int k = 0; k < 8; ++k)
        if (nd[k] == node) nd[k] = copy;
    }
  }
  return 1;
}

typedef struct {
  double d;
  uint32_t node;
} by_distance;

static int by_distance_cmp(const void *x, const void *y) {
  const by_distance *a = (const by_distance *)x, *c = (const by_distance *)y;
  if (a->d < c->d) return -1;
  if (c->d < a->d) return 1;
  return a->node < c->node ? -1 : (a->node > c->node ? 1 : 0);
}

/* crackThroughCell (tearing.cc:302-376): the corner nearest the plane
 * through the cell's rest centre across the stretch whose cells lie on both
 * sides, each side-group of them reaching a whole piece, is split */
static int crack_through_cell(orc_lattice_body *b, uint32_t cell, copies_t *copies) {
  double s, across[3], centre[3] = {0.0, 0.0, 0.0};
  orc_lattice_cell_stretch(b, cell, &s, across);
  uint32_t corners[8];
  memcpy(corners, b->hex + 8 * (size_t)cell, sizeof corners);
  for (int k = 0; k < 8; ++k) {
    double r[3];
    rest_of(b, corners[k], r);
    for (int a = 0; a < 3; ++a) centre[a] += r[a] / 8.0;
  }
  by_distance order[8];
  for (int k = 0; k < 8; ++k) {
    double r[3];
    rest_of(b, corners[k], r);
    for (int a = 0; a < 3; ++a) r[a] -= centre[a];
    order[k].d = fabs(dot3(r, across));
    order[k].node = corners[k];
  }
  qsort(order, 8, sizeof order[0], by_distance_cmp);
  uint32_t *fan = (uint32_t *)malloc(b->hexes * sizeof(uint32_t));
  uint32_t *group = (uint32_t *)malloc(b->hexes * sizeof(uint32_t));
  uint32_t *seeds = (uint32_t *)malloc(b->hexes * sizeof(uint32_t));
  uint8_t *sides = (uint8_t *)malloc(b->hexes);
  int done = 0;
  for (int k = 0; k < 8 && !done; ++k) {
    const uint32_t node = order[k].node;
    if (b->anchored[node]) continue;
    const uint32_t n = cells_at(b, node, fan);
    int low = 0, high = 0;
    double rn[3];
    rest_of(b, node, rn);
    for (uint32_t at = 0; at < n; ++at) {
      double offset = 0.0;
      for (int c = 0; c < 8; ++c) {
        double r[3];
        rest_of(b, b->hex[8 * (size_t)fan[at] + c], r);
        for (int a = 0; a < 3; ++a) r[a] -= rn[a];
        offset += dot3(r, across);
      }
      sides[at] = offset > 0.0 ? 1 : 0;
      low = low || sides[at] == 0;
      high = high || sides[at] == 1;
    }
    if (!low || !high) continue;
    const uint32_t count = groups_of(b, fan, n, sides, group);
    int whole = 1;
    for (uint32_t piece = 0; piece < count && whole; ++piece) {
      uint32_t ns = 0;
      for (uint32_t at = 0; at < n; ++at)
        if (group[at] == piece) seeds[ns++] = fan[at];
      whole = reachable(b, seeds, ns, node, MIN_PIECE_CELLS) >= MIN_PIECE_CELLS;
    }
    if (whole && split(b, node, fan, n, group, count, copies)) done = 1;
  }
  free(fan);
  free(group);
  free(seeds);
  free(sides);
  return done;
}

/* splitPinched (tearing.cc:378-408): every node of a cell touching a split
 * node, in ascending order, is split into its face-connected groups */
static void split_pinched(orc_lattice_body *b, copies_t *copies) {
  uint8_t *near = (uint8_t *)calloc(b->nodes, 1);
  for (uint32_t i = 0; i < copies->n; ++i) near[copies->copy[i]] = near[copies->original[i]] = 1;
  const uint32_t count = b->nodes;
  uint8_t *candidate = (uint8_t *)calloc(count, 1);
  for (uint32_t c = 0; c < b->hexes; ++c) {
    const uint32_t *nd = b->hex + 8 * (size_t)c;
    int touches = 0;
    for (int k = 0; k < 8; ++k) touches = touches || near[nd[k]];
    if (touches)
      for (int k = 0; k < 8; ++k) candidate[nd[k]] = 1;
  }
  free(near);
  uint32_t *fan = (uint32_t *)malloc(b->hexes * sizeof(uint32_t));
  uint32_t *group = (uint32_t *)malloc(b->hexes * sizeof(uint32_t));
  uint8_t *sides = (uint8_t *)calloc(b->hexes, 1);
  for (uint32_t node = 0; node < count; ++node) {
    if (!candidate[node]) continue;
    const uint32_t n = cells_at(b, node, fan);
    if (n < 2) continue;
    const uint32_t groups = groups_of(b, fan, n, sides, group);
    (void)split(b, node, fan, n, group, groups, copies);
  }
  free(candidate);
  free(fan);
  free(group);
  free(sides);
}

/* recomputeMasses (tearing.cc:410-424), in single precision */
static void recompute_masses(orc_lattice_body *b) {
  for (uint32_t i = 0; i < b->nodes; ++i) b->mass[i] = 0.0f;
  for (uint32_t c = 0; c < b->hexes; ++c) {
    const float density = b->lane[c] < b->lanes ? b->materials[b->lane[c]].density : 1000.0f;
    const float share = density * b->volume[c] * 0.125f;
    for (int k = 0; k < 8; ++k) b->mass[b->hex[8 * (size_t)c + k]] += share;
  }
}

/* the surface nodes buildMesh marks (solver.cc:2369-2413): the corners of
 * every cell face no other cell shares (for a lattice the grid builds, the
 * faces whose neighbour cell is not solid or off the grid) */
static void ensure_surface(orc_lattice_body *b) {
  if (b->surface) return;
  b->surface = (uint8_t *)calloc(b->nodes ? b->nodes : 1, 1);
  for (uint32_t c = 0; c < b->hexes; ++c)
    for (uint32_t face = 0; face < 6; ++face) {
      uint32_t f[4];
      face_of(b, c, face, f);
      int shared = 0;
      for (uint32_t o = 0; o < b->hexes && !shared; ++o) {
        if (o == c) continue;
        for (uint32_t g = 0; g < 6 && !shared; ++g) {
          uint32_t h[4];
          face_of(b, o, g, h);
          shared = !memcmp(f, h, sizeof f);
        }
      }
      if (!shared)
        for (int k = 0; k < 4; ++k) b->surface[f[k]] = 1;
    }
}

/* markCracks (tearing.cc:426-484): the faces of each cell whose grid
 * neighbour's facing face no longer holds the same nodes, their nodes put on
 * the surface, then each copy of a surface node */
static void mark_cracks(orc_lattice_body *b, const copies_t *copies) {
  b->surface = (uint8_t *)realloc(b->surface, b->nodes);
  for (uint32_t i = b->surface_nodes; i < b->nodes; ++i) b->surface[i] = 0;
  b->surface_nodes = b->nodes;
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  for (uint32_t i = 0; i < b->nodes; ++i)
    for (int k = 0; k < 3; ++k) {
      const float r = (float)b->rest[3 * (size_t)i + k];
      lo[k] = r < lo[k] ? r : lo[k];
    }
  int32_t *index = (int32_t *)malloc(3 * (size_t)b->hexes * sizeof(int32_t));
  for (uint32_t c = 0; c < b->hexes; ++c)
    for (int k = 0; k < 3; ++k)
      index[3 * (size_t)c + k] =
          (int32_t)roundf(((float)b->rest[3 * (size_t)b->hex[8 * (size_t)c] + k] - lo[k]) / b->voxel[k]);
  if (!b->crack) b->crack = (uint8_t *)calloc(b->hexes ? b->hexes : 1, 1);
  for (uint32_t c = 0; c < b->hexes; ++c) {
    b->crack[c] = 0;
    for (uint32_t axis = 0; axis < 3; ++axis)
      for (uint32_t side = 0; side < 2; ++side) {
        int32_t next[3] = {index[3 * (size_t)c], index[3 * (size_t)c + 1], index[3 * (size_t)c + 2]};
        next[axis] += side == 0 ? -1 : 1;
        uint32_t found = UINT32_MAX; /* the first cell at that index (unordered_map::emplace keeps it) */
        for (uint32_t o = 0; o < b->hexes && found == UINT32_MAX; ++o)
          if (index[3 * (size_t)o] == next[0] && index[3 * (size_t)o + 1] == next[1] &&
              index[3 * (size_t)o + 2] == next[2])
            found = o;
        if (found == UINT32_MAX) continue;
        uint32_t mine[4], theirs[4];
        face_of(b, c, 2 * axis + side, mine);
        face_of(b, found, 2 * axis + (1 - side), theirs);
        if (memcmp(mine, theirs, sizeof mine)) {
          b->crack[c] |= (uint8_t)(1u << (2 * axis + side));
          for (int k = 0; k < 4; ++k) b->surface[mine[k]] = 1;
        }
      }
  }
  free(index);
  for (uint32_t i = 0; i < copies->n; ++i)
    if (b->surface[copies->original[i]]) b->surface[copies->copy[i]] = 1;
}

uint32_t orc_lattice_tear(orc_lattice_body *b, const uint32_t *cells, uint32_t count) {
  if (count == 0) return 0;
  ensure_surface(b);
  b->surface_nodes = b->nodes;
  copies_t copies = {NULL, NULL, 0, 0};
  for (uint32_t i = 0; i < count; ++i) (void)crack_through_cell(b, cells[i], &copies);
  const uint32_t made0 = copies.n;
  if (made0 == 0) return 0;
  split_pinched(b, &copies);
  recompute_masses(b);
  mark_cracks(b, &copies);
  b->retopo = 1;
  const uint32_t made = copies.n;
  free(copies.copy);
  free(copies.original);
  return made;
}

uint32_t orc_lattice_surface(orc_lattice_body *b, uint32_t *nodes) {
  if (!b->surface) {
    ensure_surface(b);
    b->surface_nodes = b->nodes;
  }
  uint32_t n = 0;
  for (uint32_t i = 0; i < b->nodes; ++i)
    if (b->surface[i]) {
      if (nodes) nodes[n] = i;
      ++n;
    }
  return n;
}

uint32_t orc_lattice_pieces(const orc_lattice_body *b) {
  uint32_t *parent = (uint32_t *)malloc((b->hexes ? b->hexes : 1) * sizeof(uint32_t));
  uint32_t *owner = (uint32_t *)malloc((b->nodes ? b->nodes : 1) * sizeof(uint32_t));
  for (uint32_t c = 0; c < b->hexes; ++c) parent[c] = c;
  for (uint32_t i = 0; i < b->nodes; ++i) owner[i] = UINT32_MAX;
  for (uint32_t c = 0; c < b->hexes; ++c)
    for (int k = 0; k < 8; ++k) {
      const uint32_t node = b->hex[8 * (size_t)c + k];
      if (owner[node] == UINT32_MAX) {
        owner[node] = c;
        continue;
      }
      const uint32_t x = root_of(parent, owner[node]), y = root_of(parent, c);
      if (x != y) parent[x > y ? x : y] = x < y ? x : y;
    }
  uint32_t pieces = 0;
  for (uint32_t c = 0; c < b->hexes; ++c) pieces += root_of(parent, c) == c ? 1u : 0u;
  free(parent);
  free(owner);
  return pieces;
}

uint32_t orc_lattice_topology(const orc_lattice_body *b, uint32_t *hex_nodes, float *node_mass,
                              uint8_t *crack_faces) {
  if (hex_nodes) memcpy(hex_nodes, b->hex, 8 * (size_t)b->hexes * sizeof(uint32_t));
  if (node_mass) memcpy(node_mass, b->mass, (size_t)b->nodes * sizeof(float));
  if (crack_faces) {
    if (b->crack)
      memcpy(crack_faces, b->crack, b->hexes);
    else
      memset(crack_faces, 0, b->hexes);
  }
  return b->nodes;
}

uint32_t orc_lattice_step(orc_lattice_body *b, double dt, const double *g, uint32_t softstep,
                          uint32_t colored) {
  const uint32_t substeps = orc_lattice_substeps(b, dt, softstep);
  const double h = dt / (double)substeps;
  if (b->h != h) {
    build_elements(b, h); /* fresh elements: impulses start at zero */
  } else if (b->retopo) {
    /* the mesh tore: every element is built again and keeps its cell's
     * impulses and rotation (strainstep.cc:610-623) */
    orc_hex *old = b->elements;
    const uint32_t was = b->count;
    b->elements = NULL;
    build_elements(b, h);
    if (was == b->count)
      for (uint32_t i = 0; i < was; ++i)
        if (b->elements[i].cell == old[i].cell) {
          memcpy(b->elements[i].impulse, old[i].impulse, sizeof old[i].impulse);
          b->elements[i].vol_impulse = old[i].vol_impulse;
          memcpy(b->elements[i].quat, old[i].quat, sizeof old[i].quat);
          orc_quat_matrix(old[i].quat, b->elements[i].turn);
        }
    free(old);
  } else {
    refresh_flowed(b);
  }
  b->retopo = 0;
  const int exact = !(h * stiffest(b) > RESOLVED);
  /* sweep order: cell order, or colour by colour */
  uint32_t *order = (uint32_t *)malloc((b->count ? b->count : 1) * sizeof(uint32_t));
  uint32_t n = 0;
  if (colored) {
    for (uint32_t col = 0; col < 8; ++col)
      for (uint32_t i = 0; i < b->count; ++i)
        if (b->hex_colour[b->elements[i].cell] == col) order[n++] = i;
  } else {
    for (uint32_t i = 0; i < b->count; ++i) order[n++] = i;
  }
  for (uint32_t s = 0; s < substeps; ++s) {
    for (size_t i = 0; i < b->nodes; ++i)
      for (int k = 0; k < 3; ++k) b->v[3 * i + k] = pinned(b, i) ? 0.0 : b->v[3 * i + k] + h * g[k];
    if (exact)
      for (uint32_t i = 0; i < n; ++i) {
        orc_hex *e = &b->elements[order[i]];
        orc_hex_update(e, b->x);
        orc_hex_apply(e, b->v, e->impulse, e->vol_impulse);
      }
    for (uint32_t pass = 0; pass < PASSES; ++pass)
      for (uint32_t i = 0; i < n; ++i) {
        orc_hex *e = &b->elements[order[i]];
        if (pass == 0 && !exact) {
          orc_hex_update(e, b->x);
          orc_hex_apply(e, b->v, e->impulse, e->vol_impulse);
        }
        orc_hex_solve(e, b->v);
      }
    for (size_t i = 0; i < b->nodes; ++i) {
      double *v = b->v + 3 * i;
      const double speed = sqrt(dot3(v, v));
      if (speed > MAX_SPEED)
        for (int k = 0; k < 3; ++k) v[k] *= MAX_SPEED / speed;
      for (int k = 0; k < 3; ++k) b->x[3 * i + k] += h * v[k];
    }
  }
  free(order);
  (void)orc_lattice_flow(b);
  /* overstretched cells tear at the step's end (strainstep.cc:826-830) */
  uint32_t *torn = (uint32_t *)malloc((b->hexes ? b->hexes : 1) * sizeof(uint32_t));
  const uint32_t ntorn = orc_lattice_overstretched(b, torn);
  if (ntorn) (void)orc_lattice_tear(b, torn, ntorn);
  free(torn);
  return substeps;
}
