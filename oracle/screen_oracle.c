/* screen_oracle.c -- FP64 CPU restatement (test infrastructure only) of the
 * FP64 neo-Hookean hex, the lattice energy ledger and the screen factor.
 * What each part restates (paths relative to the Talus tree):
 *   orc_snh_moduli          src/talus/dynamics/neohookean/neohookean.cc:82-90
 *   Gauss gradients         neohookean.cc:92-123 (points at +-1/sqrt(3))
 *   orc_snh_cell_energy     neohookean.cc:125-151
 *   orc_snh_node_terms      neohookean.cc:153-196
 *   orc_snh_rest_stiffness  neohookean.cc:198-229
 *   orc_snh_volume_ratio    neohookean.cc:231-240
 *   orc_lattice_energy      src/talus/dynamics/ledger/ledger.cc:9-52
 *   orc_rest_body_of        src/talus/dynamics/screen/screen.cc:26-83
 *   orc_screen_load         screen.cc:85-254 (factorScreen + screenLoad; a
 *                           dense Cholesky factor in place of Eigen's
 *                           SimplicialLDLT: the same solution of the same
 *                           SPD system)
 * Pinned by tests/test_screen_cpu.py: rigid motion, finite differences of
 * the energy and of the gradients, the affine closed form against the tet
 * oracle's stable neo-Hookean, det F for the volume ratio, point-mass
 * closed forms for the rest body, F L / (E A) for a squeezed block, no dent
 * for a rigid load, and the residual of the solved system. */
#include "screen_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GAUSS 0.5773502691896258
#define SCREEN_SHIFT 1.0e-9
#define ANCHOR_SCALE 1.0e6

static void corner_signs(uint32_t c, double *s) {
  s[0] = (c & 1u) ? 1.0 : -1.0;
  s[1] = (c & 2u) ? 1.0 : -1.0;
  s[2] = (c & 4u) ? 1.0 : -1.0;
}

/* shape gradients of the 8 corners at Gauss point p */
static void gauss_gradients(uint32_t p, const double *half, double *g) {
  double s[3], at[3];
  corner_signs(p, s);
  for (int k = 0; k < 3; ++k) at[k] = GAUSS * s[k];
  orc_hex_shape_gradients(at, half, g);
}

static double det3(const double *a) {
  return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
         a[2] * (a[3] * a[7] - a[4] * a[6]);
}

/* cofactor: column k of C is the cross product of F's other two columns in
 * cyclic order, so C = det(F) F^-T */
static void cofactor(const double *F, double *C) {
  C[0] = F[4] * F[8] - F[5] * F[7];
  C[1] = F[5] * F[6] - F[3] * F[8];
  C[2] = F[3] * F[7] - F[4] * F[6];
  C[3] = F[2] * F[7] - F[1] * F[8];
  C[4] = F[0] * F[8] - F[2] * F[6];
  C[5] = F[1] * F[6] - F[0] * F[7];
  C[6] = F[1] * F[5] - F[2] * F[4];
  C[7] = F[2] * F[3] - F[0] * F[5];
  C[8] = F[0] * F[4] - F[1] * F[3];
}

/* J - 1 from the displacement gradient: tr H + (tr^2 H - H:H^T)/2 + det H */
static double volume_change(const double *H) {
  const double tr = H[0] + H[4] + H[8];
  double hht = 0.0;
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) hht += H[3 * r + k] * H[3 * k + r];
  return tr + 0.5 * (tr * tr - hht) + det3(H);
}

/* H = sum_c u_c g_c^T */
static void offset_of(const double *u, const double *g, double *H) {
  memset(H, 0, 9 * sizeof(double));
  for (int c = 0; c < 8; ++c)
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) H[3 * r + k] += u[3 * c + r] * g[3 * c + k];
}

void orc_snh_moduli(float youngs, float poisson, double *shear, double *shifted) {
  double nu = (double)poisson;
  nu = nu < 0.0 ? 0.0 : (nu > 0.49 ? 0.49 : nu);
  const double E = (double)youngs;
  const double first = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  *shear = E / (2.0 * (1.0 + nu));
  *shifted = first + *shear;
}

double orc_snh_cell_energy(const double *deformed, const double *rest, const double *half, double volume,
                           double shear, double shifted) {
  double u[24], g[24], H[9];
  for (int i = 0; i < 24; ++i) u[i] = deformed[i] - rest[i];
  const double w = 0.125 * volume;
  double energy = 0.0;
  for (uint32_t p = 0; p < 8; ++p) {
    gauss_gradients(p, half, g);
    offset_of(u, g, H);
    const double change = volume_change(H);
    double hh = 0.0;
    for (int k = 0; k < 9; ++k) hh += H[k] * H[k];
    const double tr = H[0] + H[4] + H[8];
    energy += w * (0.5 * shear * hh - shear * (change - tr) + 0.5 * shifted * change * change);
  }
  return energy;
}

void orc_snh_node_terms(const double *deformed, const double *rest, const double *half, double volume,
                        double shear, double shifted, uint32_t corner, double *gradient, double *hessian) {
  double u[24], g[24], H[9], F[9], C[9], S[9];
  for (int i = 0; i < 24; ++i) u[i] = deformed[i] - rest[i];
  const double w = 0.125 * volume;
  memset(gradient, 0, 3 * sizeof(double));
  memset(hessian, 0, 9 * sizeof(double));
  for (uint32_t p = 0; p < 8; ++p) {
    gauss_gradients(p, half, g);
    offset_of(u, g, H);
    for (int k = 0; k < 9; ++k) F[k] = H[k] + ((k % 4 == 0) ? 1.0 : 0.0);
    cofactor(F, C);
    const double change = volume_change(H);
    for (int k = 0; k < 9; ++k) S[k] = shear * (F[k] - C[k]) + shifted * change * C[k];
    const double *gc = &g[3 * corner];
    double b[3];
    for (int r = 0; r < 3; ++r) {
      double sg = 0.0;
      b[r] = 0.0;
      for (int k = 0; k < 3; ++k) {
        sg += S[3 * r + k] * gc[k];
        b[r] += C[3 * r + k] * gc[k];
      }
      gradient[r] += w * sg;
    }
    const double gg = gc[0] * gc[0] + gc[1] * gc[1] + gc[2] * gc[2];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k)
        hessian[3 * r + k] += w * ((r == k ? shear * gg : 0.0) + shifted * b[r] * b[k]);
  }
}

void orc_snh_rest_stiffness(const double *half, double volume, double shear, double shifted, double *K) {
  double g[24];
  const double w = 0.125 * volume;
  memset(K, 0, 576 * sizeof(double));
  for (uint32_t p = 0; p < 8; ++p) {
    gauss_gradients(p, half, g);
    for (int c = 0; c < 8; ++c)
      for (int o = 0; o < 8; ++o) {
        const double *a = &g[3 * c], *b = &g[3 * o];
        const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
        for (int r = 0; r < 3; ++r)
          for (int k = 0; k < 3; ++k)
            K[(3 * c + r) * 24 + 3 * o + k] +=
                w * ((r == k ? shear * ab : 0.0) + shear * b[r] * a[k] + (shifted - shear) * a[r] * b[k]);
      }
  }
}

double orc_snh_volume_ratio(const double *deformed, const double *rest, const double *half) {
  double u[24], g[24], H[9];
  const double centre[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 24; ++i) u[i] = deformed[i] - rest[i];
  orc_hex_shape_gradients(centre, half, g);
  offset_of(u, g, H);
  return 1.0 + volume_change(H);
}

static void corners_of(const orc_hex_mesh *m, uint32_t cell, const double *x, double *out) {
  for (int c = 0; c < 8; ++c) {
    const uint32_t n = m->hex[8 * (size_t)cell + c];
    for (int r = 0; r < 3; ++r) out[3 * c + r] = x[3 * (size_t)n + r];
  }
}

static void half_of(const orc_hex_mesh *m, double *half) {
  for (int k = 0; k < 3; ++k) half[k] = 0.5 * (double)m->voxel[k];
}

void orc_lattice_energy(const orc_hex_mesh *m, const double *shear, const double *shifted,
                        const double *x, const double *v, const double *gravity, double *kinetic,
                        double *potential, double *elastic) {
  double ke = 0.0, pe = 0.0, el = 0.0, half[3];
  half_of(m, half);
  for (uint32_t n = 0; n < m->nodes; ++n) {
    const double mass = (double)m->mass[n];
    const double *vn = &v[3 * (size_t)n], *xn = &x[3 * (size_t)n];
    if (!(m->anchored && m->anchored[n])) ke += 0.5 * mass * (vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
    pe -= mass * (gravity[0] * xn[0] + gravity[1] * xn[1] + gravity[2] * xn[2]);
  }
  for (uint32_t cell = 0; cell < m->hexes; ++cell) {
    double d[24], r[24], shift[3];
    corners_of(m, cell, x, d);
    corners_of(m, cell, m->rest, r);
    for (int k = 0; k < 3; ++k) shift[k] = d[k] - r[k];
    for (int c = 0; c < 8; ++c)
      for (int k = 0; k < 3; ++k) d[3 * c + k] -= shift[k];
    el += orc_snh_cell_energy(d, r, half, (double)m->volume[cell], shear[cell], shifted[cell]);
  }
  *kinetic = ke;
  *potential = pe;
  *elastic = el;
}

static int inv3(const double *a, double *o) {
  double c[9];
  cofactor(a, c);
  const double d = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
  if (d == 0.0) return 0;
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) o[3 * r + k] = c[3 * k + r] / d;
  return 1;
}

void orc_rest_body_of(const orc_hex_mesh *m, orc_rest_body *out) {
  double low[3] = {INFINITY, INFINITY, INFINITY}, high[3] = {-INFINITY, -INFINITY, -INFINITY};
  memset(out, 0, sizeof *out);
  for (uint32_t n = 0; n < m->nodes; ++n) {
    const double mass = (double)m->mass[n];
    out->total += mass;
    for (int k = 0; k < 3; ++k) {
      const double p = m->rest[3 * (size_t)n + k];
      out->centre[k] += mass * p;
      low[k] = p < low[k] ? p : low[k];
      high[k] = p > high[k] ? p : high[k];
    }
  }
  for (int k = 0; k < 3; ++k) out->centre[k] /= out->total;
  for (uint32_t n = 0; n < m->nodes; ++n) {
    const double mass = (double)m->mass[n];
    double a[3];
    for (int k = 0; k < 3; ++k) a[k] = m->rest[3 * (size_t)n + k] - out->centre[k];
    const double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 3; ++k) out->inertia[3 * r + k] += mass * ((r == k ? aa : 0.0) - a[r] * a[k]);
  }
  /* eigenvalues ascending, eigenvectors as columns; a reflection turns the
   * last axis */
  double w[3], V[9];
  orc_sym_eigen3(out->inertia, w, V);
  int order[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (w[order[j]] < w[order[i]]) {
        const int t = order[i];
        order[i] = order[j];
        order[j] = t;
      }
  for (int c = 0; c < 3; ++c) {
    out->moments[c] = w[order[c]];
    for (int r = 0; r < 3; ++r) out->axes[3 * r + c] = V[3 * r + order[c]];
  }
  if (det3(out->axes) < 0.0)
    for (int r = 0; r < 3; ++r) out->axes[3 * r + 2] = -out->axes[3 * r + 2];
  double thin = INFINITY;
  for (int k = 0; k < 3; ++k) thin = (high[k] - low[k]) < thin ? (high[k] - low[k]) : thin;
  out->thickness = thin;
}

int orc_screen_load(const orc_hex_mesh *m, const double *shear, const double *shifted, const uint32_t *held,
                    uint32_t nheld, const double *forces, double *displacement, double *dent) {
  const size_t N = m->nodes, D = 3 * N;
  if (D == 0 || D > 6000) return 0;
  double *A = (double *)calloc(D * D, sizeof(double));
  double *b = (double *)calloc(D, sizeof(double));
  uint8_t *hold = (uint8_t *)calloc(N, 1);
  if (!A || !b || !hold) {
    free(A);
    free(b);
    free(hold);
    return 0;
  }
  double half[3], K[576];
  half_of(m, half);
  for (uint32_t cell = 0; cell < m->hexes; ++cell) {
    orc_snh_rest_stiffness(half, (double)m->volume[cell], shear[cell], shifted[cell], K);
    for (int c = 0; c < 8; ++c)
      for (int o = 0; o < 8; ++o)
        for (int r = 0; r < 3; ++r)
          for (int k = 0; k < 3; ++k) {
            const size_t i = 3 * (size_t)m->hex[8 * (size_t)cell + c] + r;
            const size_t j = 3 * (size_t)m->hex[8 * (size_t)cell + o] + k;
            A[i * D + j] += K[(3 * c + r) * 24 + 3 * o + k];
          }
  }
  double largest = -INFINITY, heaviest = 0.0;
  for (size_t i = 0; i < D; ++i) largest = A[i * D + i] > largest ? A[i * D + i] : largest;
  for (size_t n = 0; n < N; ++n) heaviest = (double)m->mass[n] > heaviest ? (double)m->mass[n] : heaviest;
  int anchored = 0;
  for (size_t n = 0; n < N; ++n) hold[n] = m->anchored && m->anchored[n] ? 1 : 0;
  for (uint32_t k = 0; k < nheld; ++k) hold[held[k]] = 1;
  for (size_t n = 0; n < N; ++n) {
    anchored = anchored || hold[n];
    for (int r = 0; r < 3; ++r) {
      const size_t i = 3 * n + r;
      A[i * D + i] += hold[n] ? ANCHOR_SCALE * largest : SCREEN_SHIFT * largest / heaviest * (double)m->mass[n];
    }
  }
  orc_rest_body rb;
  orc_rest_body_of(m, &rb);
  double force[3] = {0, 0, 0}, torque[3] = {0, 0, 0}, linear[3] = {0, 0, 0}, angular[3] = {0, 0, 0}, Iinv[9];
  inv3(rb.inertia, Iinv);
  for (size_t n = 0; n < N; ++n) {
    const double *f = &forces[3 * n];
    double a[3];
    for (int k = 0; k < 3; ++k) a[k] = m->rest[3 * n + k] - rb.centre[k];
    for (int k = 0; k < 3; ++k) force[k] += f[k];
    torque[0] += a[1] * f[2] - a[2] * f[1];
    torque[1] += a[2] * f[0] - a[0] * f[2];
    torque[2] += a[0] * f[1] - a[1] * f[0];
  }
  if (!anchored) {
    for (int k = 0; k < 3; ++k) linear[k] = force[k] / rb.total;
    for (int r = 0; r < 3; ++r) angular[r] = Iinv[3 * r] * torque[0] + Iinv[3 * r + 1] * torque[1] + Iinv[3 * r + 2] * torque[2];
  }
  for (size_t n = 0; n < N; ++n) {
    if (hold[n]) continue;
    double a[3];
    for (int k = 0; k < 3; ++k) a[k] = m->rest[3 * n + k] - rb.centre[k];
    const double wa[3] = {angular[1] * a[2] - angular[2] * a[1], angular[2] * a[0] - angular[0] * a[2],
                          angular[0] * a[1] - angular[1] * a[0]};
    const double mass = (double)m->mass[n];
    for (int k = 0; k < 3; ++k) b[3 * n + k] = forces[3 * n + k] - mass * (linear[k] + wa[k]);
  }
  /* dense Cholesky A = L L^T in place (lower triangle) */
  int ok = 1;
  for (size_t j = 0; j < D && ok; ++j) {
    double s = A[j * D + j];
    for (size_t k = 0; k < j; ++k) s -= A[j * D + k] * A[j * D + k];
    if (!(s > 0.0)) {
      ok = 0;
      break;
    }
    const double l = sqrt(s);
    A[j * D + j] = l;
    for (size_t i = j + 1; i < D; ++i) {
      double t = A[i * D + j];
      for (size_t k = 0; k < j; ++k) t -= A[i * D + k] * A[j * D + k];
      A[i * D + j] = t / l;
    }
  }
  if (ok) {
    for (size_t i = 0; i < D; ++i) {
      double t = b[i];
      for (size_t k = 0; k < i; ++k) t -= A[i * D + k] * b[k];
      b[i] = t / A[i * D + i];
    }
    for (size_t i = D; i-- > 0;) {
      double t = b[i];
      for (size_t k = i + 1; k < D; ++k) t -= A[k * D + i] * b[k];
      b[i] = t / A[i * D + i];
    }
    memcpy(displacement, b, D * sizeof(double));
    if (!anchored) {
      double shift[3] = {0, 0, 0}, moment[3] = {0, 0, 0}, turn[3];
      for (size_t n = 0; n < N; ++n)
        for (int k = 0; k < 3; ++k) shift[k] += (double)m->mass[n] * displacement[3 * n + k];
      for (int k = 0; k < 3; ++k) shift[k] /= rb.total;
      for (size_t n = 0; n < N; ++n) {
        double a[3], d[3];
        for (int k = 0; k < 3; ++k) {
          a[k] = m->rest[3 * n + k] - rb.centre[k];
          d[k] = displacement[3 * n + k] - shift[k];
        }
        const double mass = (double)m->mass[n];
        moment[0] += mass * (a[1] * d[2] - a[2] * d[1]);
        moment[1] += mass * (a[2] * d[0] - a[0] * d[2]);
        moment[2] += mass * (a[0] * d[1] - a[1] * d[0]);
      }
      for (int r = 0; r < 3; ++r) turn[r] = Iinv[3 * r] * moment[0] + Iinv[3 * r + 1] * moment[1] + Iinv[3 * r + 2] * moment[2];
      for (size_t n = 0; n < N; ++n) {
        double a[3];
        for (int k = 0; k < 3; ++k) a[k] = m->rest[3 * n + k] - rb.centre[k];
        const double ta[3] = {turn[1] * a[2] - turn[2] * a[1], turn[2] * a[0] - turn[0] * a[2],
                              turn[0] * a[1] - turn[1] * a[0]};
        for (int k = 0; k < 3; ++k) displacement[3 * n + k] -= shift[k] + ta[k];
      }
    }
    double most = 0.0;
    for (size_t n = 0; n < N; ++n) {
      const double *d = &displacement[3 * n];
      const double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      most = l > most ? l : most;
    }
    *dent = most / rb.thickness;
  }
  free(A);
  free(b);
  free(hold);
  return ok;
}
